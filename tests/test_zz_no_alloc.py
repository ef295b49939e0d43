"""The allocation-free sort entry point (mpm_sort_pairs_noalloc) on device
memory, against the allocating mpm_sort_pairs on the same keys: bit-exact
keys and permutation, on consecutive calls that share the static scratch
array, the later ones smaller."""
import ctypes as C

import numpy as np
import pytest

from mpm_amd import engine
from mpm_amd import lib as L

pytestmark = pytest.mark.gpu


def _noalloc(keys, key_bits, alias=False):
    import torch
    lib = L.load()
    n = len(keys)
    dk = torch.from_numpy(keys.view(np.int32).copy()).cuda()
    ko = dk if alias else torch.empty(n, dtype=torch.int32, device="cuda")
    perm = torch.empty(n, dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    rc = lib.mpm_sort_pairs_noalloc(C.c_void_p(dk.data_ptr()), n, key_bits, C.c_void_p(ko.data_ptr()),
                                    C.c_void_p(perm.data_ptr()), C.c_void_p(stream))
    torch.cuda.synchronize()
    assert rc == L.MPM_OK, rc
    return ko.cpu().numpy().view(np.uint32), perm.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("alias", [False, True])
def test_noalloc_sort_matches_allocating_sort(gpu, alias):
    rng = np.random.default_rng(11)
    # (n, key bits): three passes (result copied back from the scratch
    # buffers), then a smaller two-pass call over the stale scratch, then a
    # one-pass and a zero-bit call.
    for n, bits in [(400_000, 24), (50_001, 16), (7_777, 8), (1_000, 0)]:
        hi = 1 << bits if bits else 1
        # few distinct keys, so the stability of ties is exercised
        keys = rng.integers(0, min(hi, 4096), n, dtype=np.uint64).astype(np.uint32)
        keys = (keys * np.uint32(max(1, hi // 4096))) & np.uint32(hi - 1 if bits else 0)
        ref_k, ref_p = engine.sort_pairs(keys, bits)
        got_k, got_p = _noalloc(keys, bits, alias=alias)
        assert np.array_equal(got_k, ref_k), (n, bits)
        assert np.array_equal(got_p, ref_p), (n, bits)
        assert np.array_equal(got_p, np.argsort(keys, kind="stable")), (n, bits)


def test_noalloc_sort_rejects_sizes_beyond_the_scratch(gpu):
    lib = L.load()
    big = 1 << 20
    rc = lib.mpm_sort_pairs_noalloc(C.c_void_p(1), big + 1, 32, C.c_void_p(1), C.c_void_p(1), None)
    assert rc == L.MPM_ERR_CAPACITY
    assert lib.mpm_sort_pairs_noalloc(None, 10, 32, None, None, None) == L.MPM_ERR_INVALID_ARGUMENT
    assert lib.mpm_sort_pairs_noalloc(None, 10, 33, None, None, None) == L.MPM_ERR_INVALID_ARGUMENT
    assert lib.mpm_sort_pairs_noalloc(None, 0, 32, None, None, None) == L.MPM_OK
