"""The service protocol's packet layout through the C ABI (host only): the
reference's packet() (gpu/server.py:17-20) and the worker's size checks
(src/three/gpu-worker.js:36-42)."""
import ctypes as C
import json
import struct

import numpy as np
import pytest

from mpm_amd import lib as L
from mpm_amd import wire


def reference_layout(meta, data):
    """u32 LE header length + compact JSON + spaces to 4 bytes + float32 LE rows."""
    head = json.dumps(meta, separators=(",", ":"), ensure_ascii=False, allow_nan=False).encode("utf-8")
    return struct.pack("<I", len(head)) + head + b" " * ((-len(head)) % 4) + np.asarray(data, "<f4").tobytes()


@pytest.mark.parametrize("extra", ["", "a", "ab", "abc", "温度"])
def test_packet_matches_the_service_layout_for_every_padding(extra):
    rng = np.random.default_rng(len(extra))
    data = rng.normal(size=(7, 8)).astype(np.float32)
    meta = dict(type="frame", count=10, capacity=64, renderCount=7, sampleStride=2, note=extra, mass=1.5)
    buf = wire.packet(meta, data)
    assert buf == reference_layout(meta, data)
    assert len(buf) % 4 == 0
    m, d = wire.decode(buf)
    assert m == meta and np.array_equal(d, data)


def test_empty_frame_and_header_only():
    meta = dict(count=0, capacity=8, renderCount=0)
    buf = wire.packet(meta, np.zeros((0, 8), np.float32))
    assert buf == reference_layout(meta, np.zeros((0, 8), np.float32))
    assert wire.decode(buf)[1].shape == (0, 8)


def test_decode_applies_the_worker_checks():
    data = np.zeros((3, 8), np.float32)
    good = wire.packet(dict(count=3, capacity=3, renderCount=3), data)
    wire.decode(good)
    for bad_meta in (dict(count=3, capacity=2, renderCount=3), dict(count=2, capacity=3, renderCount=3),
                     dict(count=3.0, capacity=3, renderCount=3), dict(count=3, capacity=3, renderCount=-1)):
        with pytest.raises(ValueError):
            wire.decode(wire.packet(bad_meta, data))
    with pytest.raises(ValueError):
        wire.decode(good[:-4])  # a truncated row
    with pytest.raises(ValueError):
        wire.decode(struct.pack("<I", 1000) + good[4:])  # header length past the end
    # renderCount defaults to count (gpu-worker.js:39)
    wire.decode(wire.packet(dict(count=3, capacity=3), data))


def test_packet_rejects_nan_in_the_header_and_small_buffers():
    with pytest.raises(ValueError):
        wire.packet(dict(count=0, capacity=0, mass=float("nan")), np.zeros((0, 8)))
    lib = L.load()
    length = C.c_uint64()
    out = (C.c_uint8 * 8)()
    head = b'{"count":0}'
    assert lib.mpm_frame_packet(head, len(head), None, 0, out, 8, C.byref(length)) == L.MPM_ERR_CAPACITY
    assert length.value == 4 + 11 + 1


def test_ready_message_names_the_features():
    msg = json.loads(wire.ready_message("MI355X"))
    assert msg["type"] == "ready" and msg["backend"] == dict(arch="hip", device="MI355X")
    assert "solidSampling" in msg["features"] and "discretization" in msg["features"]
    assert "customPhaseChange" not in msg["features"]  # the worker then refuses melting scenes itself


def test_preview_layout():
    lib = L.load()
    s, r = C.c_uint32(), C.c_uint32()
    for n, limit, want in ((1_000_000, 100_000, (10, 100_000)), (54, 100_000, (1, 54)), (0, 10, (1, 0)),
                           (100_001, 100_000, (2, 50_001))):
        assert lib.mpm_preview_layout(n, limit, C.byref(s), C.byref(r)) == L.MPM_OK
        assert (s.value, r.value) == want
    assert lib.mpm_preview_layout(5, 0, C.byref(s), C.byref(r)) == L.MPM_ERR_INVALID_ARGUMENT
