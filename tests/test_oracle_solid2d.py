"""Pins the oracle's 2D SolidMPM restatement (src/solid-mpm.js) and the
all-solid branch of the 2D driver (src/simulation.js:240-300): the
reference's own tests (tests/solid-mpm.test.mjs) restated, and closed forms.
CPU only."""
import math

import numpy as np
import pytest

from mpm_amd import solid2d
from oracle import binding as o


def width(sc):
    return float(sc.x.max() - sc.x.min())


# tests/solid-mpm.test.mjs:7-12
def test_solid_stress_rotation_free_and_shear_restoring():
    th = .7
    c, s = math.cos(th), math.sin(th)
    for tau in (o.solid_stress2(1, 0, 0, 1, 400, 400), o.solid_stress2(c, -s, s, c, 400, 400)):
        assert np.abs(tau).max() < 1e-10
    shear = o.solid_stress2(1, .3, 0, 1, 400, 400)
    assert shear[1] > 0 and abs(shear[1] - shear[2]) < 1e-10
    assert o.solid_stress2(1.2, 0, 0, 1, 400, 400)[0] > 0


def test_solid_stress_equals_the_svd_polar_form():
    # 2 mu (F - R) F^T + lambda (J - 1) J I with R = U V^T from the SVD.
    rng = np.random.default_rng(2)
    for _ in range(200):
        F = np.eye(2) + rng.normal(0, .3, (2, 2))
        if np.linalg.det(F) <= .05:
            continue
        U, _, Vt = np.linalg.svd(F)
        R = U @ Vt
        if np.linalg.det(R) < 0:
            continue
        J = np.linalg.det(F)
        want = 2 * 3.0 * (F - R) @ F.T + 5.0 * (J - 1) * J * np.eye(2)
        got = o.solid_stress2(F[0, 0], F[0, 1], F[1, 0], F[1, 1], 3.0, 5.0)
        np.testing.assert_allclose(got, want.reshape(4), rtol=1e-10, atol=1e-12)


# tests/solid-mpm.test.mjs:20-26
def test_uniform_translation_is_kept():
    pts, _ = solid2d.panel(-15, -15, 15, 15)
    sc = o.Scene2(pts, (7, -3))
    sc.step(.001)
    assert np.abs(sc.vx - 7).max() < 1e-6 and np.abs(sc.vy + 3).max() < 1e-6
    assert sc.stats()["strain"] < 1e-10


# tests/solid-mpm.test.mjs:27-35
def test_stretched_panel_contracts_after_release():
    pts, _ = solid2d.panel(-18, -18, 18, 18)
    sc = o.Scene2(pts)
    reference = width(sc)
    sc.x[:] *= 1.25
    sc.F[0::4] = 1.25
    stretched = width(sc)
    low, top_speed = math.inf, 0.0
    for _ in range(120):
        sc.advance(1 / 60, gravity=0)
        low = min(low, width(sc))
        top_speed = max(top_speed, float(np.abs(sc.vx).max()))
    assert low < reference * 1.1 and low < stretched * .9 and top_speed > 1
    assert sc.stats()["minJ"] > 0


def test_panel_lattice_follows_create():
    # nx = floor(width / 3.6) + 1 per axis; a fixed panel pins its first two columns
    pts, pins = solid2d.panel(-60, 0, 25, 25, fixed=True)
    assert len(pts) == (int(85 // 3.6) + 1) * (int(25 // 3.6) + 1)
    assert pins.sum() == 2 * (int(25 // 3.6) + 1)
    assert solid2d.panel(0, 0, 5, 20) is None  # narrower than 7
    beam, _ = solid2d.panel(0, 0, 30, 40, shape="beam")
    assert len(beam) == (int(50 // 3.6) + 1) * (int(22 // 3.6) + 1)


def test_momentum_is_conserved_by_the_transfers():
    # No walls, no pins, no gravity: sum m v is kept up to float32 storage.
    rng = np.random.default_rng(3)
    pts, _ = solid2d.panel(-20, -10, 20, 10)
    sc = o.Scene2(pts, rng.normal(0, 5, (len(pts), 2)))
    sc.F[:] = (np.tile([1, 0, 0, 1], len(pts)) + rng.normal(0, .05, 4 * len(pts)))
    P0 = (sc.mass[:, None] * sc.velocities()).sum(0)
    scale = (sc.mass[:, None] * np.abs(sc.velocities())).sum()
    for _ in range(20):
        sc.step(1 / 480)
    P = (sc.mass[:, None] * sc.velocities()).sum(0)
    assert np.abs(P - P0).max() < 1e-6 * scale


def test_stable_step_closed_forms():
    sc = o.Scene2([[0, 0], [10, 0]])
    assert sc.stable_step() == pytest.approx(1 / 240, rel=1e-15)  # |I|_F^2 = 2: wave 2 sqrt(19200) < .18*8*240
    sc.F[:4] = [2, 0, 0, 2]  # |F|_F^2 = 8: wave = 8 sqrt(19200)
    assert sc.stable_step() == pytest.approx(.18 * 8 / (8 * math.sqrt(19200)), rel=1e-12)
    sc.F[:4] = [1, 0, 0, 1]
    sc.vx[1] = 300
    assert sc.stable_step() == pytest.approx(.18 * 8 / (300 + 2 * math.sqrt(19200)), rel=1e-6)
    sc.vx[1] = 0
    sc.C[4:8] = [20, 0, 0, 30]  # |C|_1 = 50: .1 / 50
    assert sc.stable_step() == pytest.approx(.002, rel=1e-12)
    sc2 = o.Scene2([[0, 0]], stiffness=4.0)
    sc2.F[:] = [1.5, 0, 0, 1.5]
    assert sc2.stable_step() == pytest.approx(.18 * 8 / (4.5 * math.sqrt(4 * 19200)), rel=1e-12)


def test_stats_closed_form():
    sc = o.Scene2([[0, 0], [4, 0]])
    sc.F[:] = [1.1, 0, 0, 1, .9, .2, 0, 1]
    st = sc.stats()
    assert (st["count"], st["minJ"], st["maxJ"]) == (2, pytest.approx(.9), pytest.approx(1.1))
    # .5 |(a^2 + c^2 - 1, b^2 + d^2 - 1, sqrt2 (ab + cd))| of the second particle
    want = .5 * math.sqrt((.81 - 1) ** 2 + (.04 + 1 - 1) ** 2 + 2 * (.9 * .2) ** 2)
    assert st["strain"] == pytest.approx(max(want, .5 * (1.21 - 1)), rel=1e-12)
    assert o.Scene2(np.zeros((0, 2))).stats()["minJ"] == 1


def test_walls_pins_and_gravity():
    pts, pins = solid2d.panel(-40, -120, 0, -100, fixed=True)
    sc = o.Scene2(pts, (0, -60), pinned=pins, walls="softbody")
    x0 = sc.positions()
    for _ in range(30):
        sc.advance(1 / 60, gravity=1)
    x = sc.positions()
    np.testing.assert_array_equal(x[pins == 1], x0[pins == 1])  # pinned particles never move
    assert (x[:, 1] >= -145 + .1 - 1e-4).all() and np.isfinite(x).all()
    assert x[pins == 0, 1].mean() < x0[pins == 0, 1].mean()  # uniform gravity pulls down
    sc = o.Scene2([[30, 0]])
    n, adv = sc.advance(1 / 60, gravity=2)
    assert adv == pytest.approx(1 / 60) and n == 4 and sc.vx[0] < 0  # centre gravity pulls towards 0
    n, adv = sc.advance(1.0)
    assert adv == pytest.approx(.05) and n == 12  # dt clamped to .05
