"""Pins the oracle's frame wrapper (Simulation3.step / substep, all-solid branch,
src/three/simulation.js:305-388) and regularizeSolids
(src/three/solid-sampling.js:5-42) against closed forms. CPU only.

Reference units throughout: cell 8, particle volume 1 (the spacing rule's
unit), mass 1.1, E from mass and volume (parts with E = 0)."""
import math

import numpy as np
import pytest

from oracle import binding as o

DOM = dict(lower=(-64, -48, -64), upper=(64, 72, 64), h=8)


def scene(points, v=(0, 0, 0), part=0, pinned=None, volume=1.0):
    return o.Scene(np.asarray(points, float), v, 1.1, volume, part=part, pinned=pinned, **DOM)


def regularize(sc, dt, lo=(-1e9,) * 3, hi=(1e9,) * 3):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    o.lib().orc_regularize(o.C.byref(sc.s), dt, o._p(lo), o._p(hi))


def test_regularize_pushes_a_close_pair_apart_symmetrically():
    dt = 1 / 240
    r = 1.0  # target spacing 2.8 cbrt(1 * det I) = 2.8 for both
    sc = scene([[0, 0, 0], [r, 0, 0]])
    regularize(sc, dt)
    gain = 1 - math.exp(-40 * dt)
    step = (2.8 - r) * gain / 2  # each, equal masses, well inside the 8 dt budget?
    budget = 8 * dt
    scale = max(1, step / budget)
    assert sc.x[0] == pytest.approx(-step / scale, rel=1e-12)
    assert sc.x[1] == pytest.approx(r + step / scale, rel=1e-12)
    assert sc.y[0] == sc.y[1] == 0 and sc.z[0] == sc.z[1] == 0


def test_regularize_budget_scales_every_displacement():
    dt = 1 / 2000  # budget 8 dt = 0.004 is below the push
    sc = scene([[0, 0, 0], [0.5, 0, 0], [20, 0, 0], [21.0, 0, 0]])
    regularize(sc, dt)
    gain = 1 - math.exp(-40 * dt)
    push = [(2.8 - 0.5) * gain / 2, (2.8 - 1.0) * gain / 2]
    scale = max(push) / (8 * dt)
    assert scale > 1
    assert sc.x[1] - 0.5 == pytest.approx(push[0] / scale, rel=1e-12)
    assert sc.x[3] - 21.0 == pytest.approx(push[1] / scale, rel=1e-12)


def test_regularize_ignores_other_parts_and_respects_pins():
    dt = 1 / 240
    gain = 1 - math.exp(-40 * dt)
    sc = scene([[0, 0, 0], [1, 0, 0]], part=[0, 1])
    regularize(sc, dt)
    assert sc.x[0] == 0 and sc.x[1] == 1
    sc = scene([[0, 0, 0], [1, 0, 0]], pinned=[1, 0])
    regularize(sc, dt)
    assert sc.x[0] == 0
    push = (2.8 - 1) * gain  # the free particle takes the whole push, within the 8 dt budget
    assert sc.x[1] == pytest.approx(1 + push / max(1, push / (8 * dt)), rel=1e-12)


def test_regularize_coincident_pair_uses_the_hashed_direction():
    dt = 1 / 240
    gain = 1 - math.exp(-40 * dt)
    sc = scene([[0, 0, 0], [0, 0, 0]])
    regularize(sc, dt)
    h = ((1 * 73856093) ^ (2 * 19349663)) & 0xffffffff
    d = np.array([(h & 1023) / 511.5 - 1, ((h >> 10) & 1023) / 511.5 - 1, ((h >> 20) & 1023) / 511.5 - 1])
    d *= 2.8 * gain / (np.linalg.norm(d) * 2)  # target * gain / (r * total) * invMass
    got0 = np.array([sc.x[0], sc.y[0], sc.z[0]])
    got1 = np.array([sc.x[1], sc.y[1], sc.z[1]])
    scale = max(1, np.linalg.norm(d) / (8 * dt))
    np.testing.assert_allclose(got0, d / scale, rtol=1e-12)
    np.testing.assert_allclose(got1, -d / scale, rtol=1e-12)


def test_spacing_follows_volume_and_jacobian():
    # target = mean of 2.8 cbrt(clamp(V J, 1/64, 8)): a compressed particle has
    # a smaller spacing, so a pair at r = 2.5 is pushed only if J is large.
    dt = 1 / 240
    sc = scene([[0, 0, 0], [2.5, 0, 0]], volume=0.5)  # spacing 2.8 cbrt(.5) = 2.22 < 2.5
    regularize(sc, dt)
    assert sc.x[1] == 2.5
    sc = scene([[0, 0, 0], [2.5, 0, 0]], volume=0.5)
    sc.F[:] = np.tile(np.diag([2.0, 1.0, 1.0]).reshape(9), 2)  # J = 2: spacing 2.8
    regularize(sc, dt)
    assert sc.x[1] > 2.5


def test_frame_free_fall_is_exact_and_substeps_follow_the_stable_step():
    sc = scene([[0, 20, 0]])
    dt = 1 / 60
    hstable = sc.stable_step()  # h/8/240 for a resting particle
    n, adv = sc.advance_frame(dt, anti_clump=False)
    assert n == math.ceil(dt / hstable - 1e-9) and adv == pytest.approx(dt, rel=1e-12)
    assert sc.vy[0] == pytest.approx(-24 * dt, rel=1e-9)  # uniform gravity 24 per unit strength
    assert sc.vx[0] == pytest.approx(0, abs=1e-12)


def test_frame_attract_gravity_and_limits():
    c = np.array([0.0, 0.0, 0.0])
    sc = scene([[10, 0, 0]])
    dt = 1e-4
    sc.advance_frame(dt, gravity_mode=2, gravity_center=c, anti_clump=False)
    a = -10 * 24000 / (10 ** 2 + 8 ** 2) ** 1.5
    assert sc.vx[0] == pytest.approx(a * dt, rel=1e-6)
    # velocity limit: a particle at 400 units/s keeps moving at 400 inside the
    # step but leaves the substep with v clamped to 160
    sc = scene([[0, 0, 0]], v=(400, 0, 0))
    n, _ = sc.advance_frame(1e-4, gravity_mode=0, anti_clump=False)
    assert sc.vx[0] == pytest.approx(160)
    # bounds: a particle driven through the lower y bound is clamped to it
    sc = scene([[0, 0.01, 0]], v=(0, -50, 0))
    sc.advance_frame(1e-3, gravity_mode=0, anti_clump=False, bounds=((-60, 0, -60), (60, 60, 60)))
    assert sc.y[0] == 0.0 and sc.vy[0] <= 0
