"""GPU parity of the grid update on grids large enough for its wide paths,
through the C ABI.

The piece compaction reads one flag per lane up to 2^17 pieces of 16 nodes
and 16 flags per lane (a 16-byte load and a wave scan) above; the grid update
then runs several workgroups per CU, each over a contiguous range of the
step's piece list, a quarter-wave per piece. In a batch the four pieces of a
wave can belong to different scenes, whose sums go to their own slots. A piece listed twice or missed would break the exact
equality of fixed-point grid mass and particle mass, and move particles.

Bars: positions against the FP64 oracle atol .003 rtol .001 (the reference's
CPU-GPU bar, gpu/test_solver.py:42); a batch scene against the same scene on
the single-scene engine: relative L2 of displacement and velocity <= 1e-5;
grid mass equal to particle mass exactly."""
import numpy as np
import pytest

from mpm_amd import Engine, Material, scenes
from oracle import binding as o

pytestmark = pytest.mark.gpu

E_REF = 1.1 / 64 * 16000


def pieces(cells):
    # a domain of `cells` cells per axis, its lower corner on a cell boundary,
    # has cells + 5 nodes per axis (two below, three above)
    return ((cells + 5) ** 3 + 15) // 16


def test_large_grid_matches_oracle(gpu):
    lower, upper, h = (0, 0, 0), (2048, 2048, 2048), 8.0  # 256^3 cells
    assert pieces(256) > 1 << 17
    rng = np.random.default_rng(17)
    a = np.arange(10) * 4.0
    Z, Y, X = np.meshgrid(a, a, a, indexing="ij")
    block = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    # three clusters far apart, so the list holds pieces of distant rows
    x = np.concatenate([block + c for c in ((300, 400, 300), (1600, 900, 700), (1000, 1700, 1750))])
    x = x + rng.uniform(-.5, .5, x.shape)
    v = rng.normal(0, 3, x.shape)
    eng = Engine(lower, upper, h, len(x), sort_every=4)  # past the 60-cell strict domain
    eng.set_materials([Material(0, E_REF, .4 * E_REF, .4 * E_REF)])
    eng.upload(x, v, 1.1, 64.0)
    orc = o.Scene(x, v, 1.1, 1.0, lower=lower, upper=upper, h=h)
    for _ in range(12):
        eng.step(2e-3)
        orc.step(2e-3)
    np.testing.assert_allclose(eng.download()["positions"], orc.positions(), atol=.003, rtol=.001)
    c = eng.conservation()
    assert c["grid_mass_q"] == c["particle_mass_q"]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_batch_over_large_total_grid(gpu):
    # Three 96^3-cell scenes: 193k pieces in the batch, 64k in each scene
    # alone, so the batch takes the 16-flag compaction and the single-scene
    # engines the one-flag path.
    bs = scenes.batched_scenes(nscenes=3, counts=(2_000, 4_000), grids=(96,), seed=7)
    assert 3 * pieces(96) > 1 << 17 > pieces(96)
    steps = 20
    b = bs.batch(sort_every=8)
    x0 = [b.download(i)["positions"] for i in range(len(bs.scenes))]
    b.step(steps)
    b.synchronize()
    for i, sc in enumerate(bs.scenes):
        eng = sc.engine(sort_every=8)
        eng.step(sc.dt, steps)
        got, want = b.download(i), eng.download()
        assert rel_l2(got["positions"] - x0[i], want["positions"] - x0[i]) <= 1e-5
        assert rel_l2(got["velocities"], want["velocities"]) <= 1e-5
        c = b.conservation(i)
        assert c["grid_mass_q"] == c["particle_mass_q"] and c["nonfinite"] == 0, i
