"""GPU parity of the step when the lane order of a wave is far from its cell
order, against the FP64 oracle through the C ABI.

The fused P2G groups the lanes of each wave by base cell (up to 8 cells by
minimum passes, the rest in lane order) before summing runs of equal cell.
These scenes reach the paths a dense, freshly sorted scene does not:
- about one particle per cell, so a wave spans far more than 8 cells;
- a re-bin period longer than the run while every particle crosses cell
  boundaries, so the slot order drifts from the cell order.

Bars: positions atol .003 rtol .001 (the reference's CPU-GPU bar,
gpu/test_solver.py:42); grid mass equal to particle mass exactly."""
import numpy as np
import pytest

from mpm_amd import Engine, Material
from oracle import binding as o

pytestmark = pytest.mark.gpu

LOWER, UPPER, H = (-64, -48, -64), (64, 72, 64), 8.0
E_REF = 1.1 / 64 * 16000


def run(x, v, steps, dt, sort_every):
    n = len(x)
    eng = Engine(LOWER, UPPER, H, n, strict_domain=True, sort_every=sort_every)
    eng.set_materials([Material(0, E_REF, .4 * E_REF, .4 * E_REF)])
    eng.upload(x, v, 1.1, 64.0)
    orc = o.Scene(x, v, 1.1, 1.0, lower=LOWER, upper=UPPER, h=H)
    for _ in range(steps):
        eng.step(dt)
        orc.step(dt)
    return eng, orc


def test_sparse_waves_span_many_cells(gpu):
    rng = np.random.default_rng(21)
    x = rng.uniform(-36, 36, (900, 3)) + [0, 12, 0]  # 0.9 particles per cell of 8
    v = rng.normal(0, 3, (900, 3))
    eng, orc = run(x, v, 12, 2e-3, 1)
    np.testing.assert_allclose(eng.download()["positions"], orc.positions(), atol=.003, rtol=.001)
    c = eng.conservation()
    assert c["grid_mass_q"] == c["particle_mass_q"]


@pytest.mark.parametrize("sort_every", [1, 1000])
def test_drifted_slot_order(gpu, sort_every):
    # A lattice block at 2 particles per cell axis moving 0.8 cells over the
    # run: with sort_every 1000 only the upload sorts, so by the end most
    # waves hold particles of cells their neighbours left. Both periods must
    # give the oracle's trajectory.
    a = np.arange(10) * 4.0 - 18
    Z, Y, X = np.meshgrid(a, a + 12, a, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + np.random.default_rng(5).uniform(-.5, .5, (1000, 3))
    v = np.tile([110.0, -60.0, 80.0], (1000, 1)) + np.random.default_rng(6).normal(0, 2, (1000, 3))
    eng, orc = run(x, v, 30, 2e-3, sort_every)
    np.testing.assert_allclose(eng.download()["positions"], orc.positions(), atol=.003, rtol=.001)
    c = eng.conservation()
    assert c["grid_mass_q"] == c["particle_mass_q"]
