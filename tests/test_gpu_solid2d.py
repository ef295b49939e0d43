"""GPU parity of the 2D elastic MLS-MPM (mpm2_*, the reference's SolidMPM,
src/solid-mpm.js) and of its all-solid frame loop (src/simulation.js:240-300)
against the FP64 oracle, through the C ABI.

The reference's own 2D tests (tests/solid-mpm.test.mjs) run on the device
with their bars; trajectories against the oracle use the reference's
CPU-GPU bar, positions atol .003 rtol .001 (gpu/test_solver.py:42), with
the same substep counts; grid mass equals particle mass exactly."""
import math

import numpy as np
import pytest

from mpm_amd import solid2d
from mpm_amd.solid2d import Solid2D, panel, settings
from oracle import binding as o

pytestmark = pytest.mark.gpu


def width(x):
    return float(x[:, 0].max() - x[:, 0].min())


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


# tests/solid-mpm.test.mjs:20-26
def test_uniform_translation_is_kept(gpu):
    pts, _ = panel(-15, -15, 15, 15)
    e = Solid2D(len(pts))
    e.upload(pts, (7, -3))
    e.step(.001)
    st = e.download()
    assert np.abs(st["vx"] - 7).max() < 1e-5 and np.abs(st["vy"] + 3).max() < 1e-5
    assert e.stats()["strain"] < 1e-6
    np.testing.assert_allclose(st["positions"], pts + [7e-3, -3e-3], atol=1e-5)


# tests/solid-mpm.test.mjs:27-35
def test_stretched_panel_contracts_and_tracks_the_oracle(gpu):
    pts, _ = panel(-18, -18, 18, 18)
    ref = width(pts)
    x = pts * [1.25, 1]
    F = np.tile([1.25, 0, 0, 1], len(pts))
    e = Solid2D(len(pts))
    e.upload(x, F=F)
    sc = o.Scene2(x)
    sc.F[:] = F
    stretched = width(x)
    low, top = math.inf, 0.0
    for frame in range(120):
        n, adv = e.advance(1 / 60, "none")
        if frame < 30:
            no, ao = sc.advance(1 / 60, gravity=0)
            assert n == no and adv == pytest.approx(ao, rel=1e-12), frame
        st = e.download()
        low = min(low, width(st["positions"]))
        top = max(top, float(np.abs(st["vx"]).max()))
        if frame == 29:
            np.testing.assert_allclose(st["positions"], sc.positions(), atol=.003, rtol=.001)
            assert rel_l2(st["velocities"], sc.velocities()) < 1e-3
    assert low < ref * 1.1 and low < stretched * .9 and top > 1
    s = e.stats()
    assert s["min_J"] > 0 and s["grid_mass_q"] == s["particle_mass_q"]


def softbody_scene():
    """A fixed cantilever and a free panel in the softbody walls, as the
    reference's demo seeds them (src/simulation.js:66-67)."""
    a, pa = panel(-120, 70, -45, 98, fixed=True)
    b, pb = panel(30, 10, 90, 70)
    return np.concatenate([a, b]), np.concatenate([pa, pb])


def test_softbody_frames_match_the_oracle(gpu):
    # Centre gravity, walls, pinned columns: 20 frames of 1/60.
    pts, pins = softbody_scene()
    s = settings("softbody")
    e = Solid2D(len(pts))
    e.upload(pts, pinned=pins)
    sc = o.Scene2(pts, pinned=pins, walls="softbody")
    for frame in range(20):
        n, adv = e.advance(1 / 60, "center", s=s)
        no, ao = sc.advance(1 / 60, gravity=2)
        assert n == no, (frame, n, no)
    st = e.download()
    np.testing.assert_allclose(st["positions"], sc.positions(), atol=.003, rtol=.001)
    np.testing.assert_array_equal(st["positions"][pins == 1], pts[pins == 1].astype(np.float32))
    assert (np.abs(st["x"]) <= 170).all() and (np.abs(st["y"]) <= 145).all()
    g, want = e.stats(), sc.stats()
    assert g["min_J"] == pytest.approx(want["minJ"], abs=1e-4) and g["strain"] == pytest.approx(want["strain"],
                                                                                               abs=1e-4)
    assert abs(g["nodes"] - want["nodes"]) <= 2 and g["substeps"] == 4
    assert g["grid_mass_q"] == g["particle_mass_q"]
    assert e.stable_step(s) == pytest.approx(sc.stable_step(), rel=1e-5)


def test_single_steps_and_stable_step_match_the_oracle(gpu):
    rng = np.random.default_rng(8)
    pts, _ = panel(-30, -20, 30, 20)
    n = len(pts)
    v = rng.normal(0, 20, (n, 2))
    F = np.tile([1, 0, 0, 1], n) + rng.normal(0, .03, 4 * n)
    e = Solid2D(n)
    e.upload(pts, v, F=F)
    sc = o.Scene2(pts, v)
    sc.F[:] = F
    for k in range(10):
        dt = sc.stable_step()
        assert e.stable_step() == pytest.approx(dt, rel=1e-5), k
        e.step(dt)
        sc.step(dt)
    st = e.download()
    assert rel_l2(st["positions"] - pts, sc.positions() - pts) < 1e-4
    assert rel_l2(st["velocities"], sc.velocities()) < 1e-4
    np.testing.assert_allclose(st["F"], sc.F, atol=2e-5)
    P0 = (1.1 * v).sum(0)
    assert np.abs((np.float32(1.1) * st["velocities"]).sum(0) - P0).max() < 1e-5 * np.abs(1.1 * v).sum()


def test_bad_arguments_and_empty_engine(gpu):
    from mpm_amd import lib as L
    e = Solid2D(8)
    assert e.stats()["count"] == 0 and e.stats()["min_J"] == 1
    assert e.advance(1 / 60) == (0, pytest.approx(1 / 60))  # nothing to step: the frame passes
    with pytest.raises(L.MpmError):
        e.upload(np.zeros((9, 2)))  # above capacity
    with pytest.raises(L.MpmError):
        e.upload(np.zeros((2, 2)), mass=0.0)
    e.upload(np.zeros((2, 2)))
    with pytest.raises(L.MpmError):
        e.step(-1.0)
    with pytest.raises(L.MpmError):
        e.advance(1 / 60, gravity=5)
    bad = settings()
    bad.stiffness = 0
    with pytest.raises(L.MpmError):
        e.step(1e-3, s=bad)
    assert solid2d.WALLS["softbody"] == (170.0, -145.0, 145.0)
