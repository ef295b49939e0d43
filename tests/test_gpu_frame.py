"""GPU test of the frame wrapper (mpm_advance_frame: Simulation3.step's
all-solid substeps around MPM3.step with regularizeSolids) against the FP64
oracle's restatement, in the reference's units (cell 8, particle volume 1 in
s.volume, mass 1.1, mu = lambda = 0.4 E with E = m/(64 V) 16000).

Scene: two parts of 3x3x3 lattice blocks at spacing 4, with clumped pairs
(0.6 apart), one exactly coincident pair, a pinned particle, and a floor
bound at y = -2 that the falling blocks reach. Bars: the same substep counts;
positions atol .003 rtol .001 (the reference's CPU-GPU bar,
gpu/test_solver.py:42) after two frames; velocities within 2e-3 of the
largest speed; the pinned particle does not move."""
import numpy as np
import pytest

from mpm_amd import Engine, Material
from oracle import binding as o

pytestmark = pytest.mark.gpu

LOWER, UPPER, H = (-64, -48, -64), (64, 72, 64), 8.0


def clumped_scene():
    pts, part = [], []
    for k, (x0, z0) in enumerate(((-12, 0), (12, 0))):
        for z in range(3):
            for y in range(3):
                for x in range(3):
                    pts.append([x0 - 4 + 4 * x, 4 * y, z0 - 4 + 4 * z])
                    part.append(k)
    pts = np.array(pts, float)
    part = np.array(part)
    pts[1] = pts[0] + [0.6, 0, 0]        # clumped pair
    pts[30] = pts[29] + [0, 0.6, 0.3]    # clumped pair in the other part
    pts[5] = pts[4]                      # coincident pair
    pinned = np.zeros(len(pts), np.uint8)
    pinned[13] = 1
    return pts, part, pinned


def test_frame_matches_oracle(gpu):
    pts, part, pinned = clumped_scene()
    n = len(pts)
    E = 1.1 / 64 * 16000
    eng = Engine(LOWER, UPPER, H, n, strict_domain=True)
    eng.set_materials([Material(0, E, .4 * E, .4 * E), Material(0, E, .4 * E, .4 * E)])
    v0 = (0, -150, 0)  # the lowest layer crosses the floor bound within the first frame
    eng.upload(pts, v0, 1.1, 64.0, material=part.astype(np.uint8), pinned=pinned)
    orc = o.Scene(pts, v0, 1.1, 1.0, lower=LOWER, upper=UPPER, h=H, part=part, pinned=pinned)
    bounds = ((-60, -2, -60), (60, 68, 60))
    for frame in range(2):
        ng, adv_g = eng.advance_frame(1 / 60, bounds_lo=bounds[0], bounds_hi=bounds[1])
        no, adv_o = orc.advance_frame(1 / 60, bounds=bounds)
        assert ng == no and adv_g == pytest.approx(adv_o, rel=1e-9), (frame, ng, no)
    st = eng.download()
    x = st["positions"]
    np.testing.assert_allclose(x, orc.positions(), atol=.003, rtol=.001)
    v = st["velocities"]
    assert np.abs(v - orc.velocities()).max() <= 2e-3 * np.abs(orc.velocities()).max()
    np.testing.assert_array_equal(x[13], pts[13].astype(np.float32))
    assert x[:, 1].min() >= -2.0  # the floor bound held
    # the regularization separated the clumped pairs
    assert np.linalg.norm(x[1] - x[0]) > 0.6 and np.linalg.norm(x[5] - x[4]) > 0
    c = eng.conservation()
    assert c["grid_mass_q"] == c["particle_mass_q"]


def test_frame_rejects_bad_settings(gpu):
    pts, part, pinned = clumped_scene()
    eng = Engine(LOWER, UPPER, H, len(pts), strict_domain=True)
    eng.upload(pts, 0, 1.1, 64.0)
    from mpm_amd import lib as L
    with pytest.raises(L.MpmError):
        eng.advance_frame(0.0)
    with pytest.raises(L.MpmError):
        eng.advance_frame(1 / 60, gravity_mode=7)
    with pytest.raises(L.MpmError):
        eng.advance_frame(1 / 60, bounds_lo=(0, 0, 0), bounds_hi=(-1, 1, 1))
    small = Engine((0, 0, 0), (1, 1, 1), 1 / 32, 8)  # SI cell: the spacing rule is in the reference's units
    small.upload(np.full((8, 3), .5), 0, 1.0, 1e-6)
    with pytest.raises(L.MpmError):
        small.advance_frame(1e-3)
    small.advance_frame(1e-3, anti_clump=0)
