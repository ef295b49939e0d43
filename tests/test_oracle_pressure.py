"""Pins the oracle's pressure-dependent returns (models 8 and 9,
src/three/pressure-plasticity.js:11-45) and tensile-strain damage
(src/three/fracture.js:11-18).

The reference's own tests are restated with their inputs and thresholds
(tests/pressure-plasticity.test.mjs:11-39, tests/course.test.mjs:49-55);
closed forms of the two returns add what they leave open."""
import math

import numpy as np
import pytest

from oracle import binding as o

P = o.PRESSURE_DEFAULTS


def diag(x, y=None, z=None):
    y = x if y is None else y
    z = x if z is None else z
    return [x, 0, 0, 0, y, 0, 0, 0, z]


def ret(F, model=8, p=None, jp=1.0, ep=0.0):
    pr = o.pressure_params(**(p or {}))
    return o.project_plastic(F, model, jp, 0, .025, 35, ep, .01, 20, o.make_part(pressure=pr))


def det(F):
    return float(np.linalg.det(np.reshape(F, (3, 3))))


def stress(r):
    J = det(r["F"]) * r["jp"]
    t = o.constitutive_stress(r["F"], 8, .4, .4)
    return dict(q=o.von_mises(t, J), mean=(t[0] + t[4] + t[8]) / (3 * J))


def close(a, b, tol=1e-7):
    assert abs(a - b) < tol, (a, b)


# tests/pressure-plasticity.test.mjs:11-24
def test_cohesive_dp_shear_and_tension_limits():
    for F in (diag(1.1), diag(.8), diag(1.3, .8, 1), [1, .4, 0, 0, 1, 0, 0, 0, 1]):
        r = ret(F)
        s = stress(r)
        assert s["q"] / math.sqrt(3) + P["slope"] * s["mean"] <= P["cohesion"] + 1e-7
        assert s["mean"] <= P["tension"] + 1e-7
        close(det(r["F"]) * r["jp"], det(F))
        assert r["plasticIncrement"] >= 0
    tension = ret(diag(1.1))
    close(stress(tension)["mean"], P["tension"])
    assert tension["jp"] > 1
    compressed = ret(diag(.8))
    close(compressed["plasticIncrement"], 0)
    close(compressed["jp"], 1)
    F = [1, .5, 0, 0, 1, 0, 0, 0, 1]
    close(ret(F, 8, dict(dilatancy=0))["jp"], 1)
    assert ret(F, 8, dict(dilatancy=.5))["jp"] > 1
    assert stress(ret(F, 8, dict(cohesion=.1)))["q"] > stress(ret(F))["q"]
    residual = o.project_plastic(np.eye(3).reshape(9), 8, tension["jp"], 0, .025, 35, tension["plasticIncrement"],
                                 .01, 20, o.make_part())
    close(residual["jp"], tension["jp"])
    close(residual["plasticIncrement"], 0)


# tests/pressure-plasticity.test.mjs:26-37
def test_foam_yields_hardens_and_keeps_compaction():
    for F in (diag(.85), diag(1.1), [1, .45, 0, 0, .95, 0, 0, 0, 1.1]):
        for jp in (.6, 1, 1.2):
            r = ret(F, 9, None, jp, .05)
            s = stress(r)
            hat = math.sqrt((s["q"] ** 2 + P["alpha"] ** 2 * s["mean"] ** 2) / (1 + P["alpha"] ** 2 / 9))
            close(det(r["F"]) * r["jp"], det(F) * jp)
            close(hat, o.foam_strength(.05 + r["plasticIncrement"]), 2e-7)
            assert .05 + r["plasticIncrement"] < P["epsD"]
    p = ret(diag(.85), 9)
    assert p["jp"] < 1 and p["plasticIncrement"] > 0
    close(stress(p)["q"], 0)
    unloaded = o.project_plastic(np.eye(3).reshape(9), 9, p["jp"], 0, .025, 35, p["plasticIncrement"], .01, 20,
                                 o.make_part())
    close(unloaded["jp"], p["jp"])
    close(unloaded["plasticIncrement"], 0)
    close(stress(unloaded)["mean"], 0)
    assert o.foam_strength(.65) > o.foam_strength(.1)
    assert abs(stress(ret(diag(.85), 9, None, 1, .3))["mean"]) > abs(stress(p)["mean"])


# tests/pressure-plasticity.test.mjs:39-47
def test_foam_flow_is_associated_and_work_conjugate():
    F = diag(.85, .97, 1.02)
    p = ret(F, 9)
    s = stress(p)
    a2 = P["alpha"] ** 2
    hat = math.sqrt((s["q"] ** 2 + a2 * s["mean"] ** 2) / (1 + a2 / 9))
    d = [math.log(F[k] / p["F"][k]) for k in (0, 4, 8)]
    tr = sum(d)
    dev = [x - tr / 3 for x in d]
    conjugate = math.sqrt((1 + a2 / 9) * (2 / 3 * sum(x * x for x in dev) + tr * tr / a2))
    close(conjugate, p["plasticIncrement"])
    tau = o.constitutive_stress(p["F"], 9, .4, .4)
    J = det(p["F"]) * p["jp"]
    work = sum(tau[4 * k] / J * d[k] for k in range(3))
    close(work, hat * p["plasticIncrement"])


def test_foam_strength_closed_form():
    # plateau + gamma r - densify ln(1 - r^beta), r = ep / epsD; locked at full densification
    for ep in (0, .1, .35, .69):
        r = ep / P["epsD"]
        want = P["plateau"] + P["gamma"] * r - P["densify"] * math.log(1 - r ** P["beta"])
        assert o.foam_strength(ep) == pytest.approx(want, rel=1e-14)
    assert o.foam_strength(.7) == 1e20 and o.foam_strength(2) == 1e20


def test_dp_return_lands_on_the_cone_with_the_dilatancy_ratio():
    # A sheared state inside the tension cut-off returns onto t + a p = c;
    # the volumetric part of the plastic strain is b times the multiplier and
    # the deviatoric part is sqrt(3) times it along the trial direction.
    G, K = .4, 2 / 3
    for shear in (.3, .5, .8):
        F = [1, shear, 0, 0, 1, 0, 0, 0, .98]
        sig, jp, inc = o.pressure_return(np.linalg.svd(np.reshape(F, (3, 3)), compute_uv=False), 8, 1.0, 0.0)
        e = np.log(np.linalg.svd(np.reshape(F, (3, 3)), compute_uv=False))
        ne = np.log(sig)
        tr, trn = e.sum(), ne.sum()
        Jt = math.exp(tr)
        mean_new = K / Jt * trn
        dev_t, dev_n = e - tr / 3, ne - trn / 3
        t_new = math.sqrt(6) * G / Jt * np.linalg.norm(dev_n) / math.sqrt(3)
        assert mean_new < P["tension"]
        assert t_new + P["slope"] * mean_new == pytest.approx(P["cohesion"], abs=1e-12)
        dl = (tr - trn) / P["dilatancy"]  # K b dl / K
        q_t = math.sqrt(6) * G / Jt * np.linalg.norm(dev_t)
        assert math.sqrt(3) * (q_t / math.sqrt(3) - G / Jt * dl) == pytest.approx(
            math.sqrt(6) * G / Jt * np.linalg.norm(dev_n), rel=1e-10)
        np.testing.assert_allclose(np.cross(dev_t, dev_n), 0, atol=1e-14)  # radial in the deviatoric plane
        assert jp == pytest.approx(math.exp(tr - trn), rel=1e-14)


def test_pressure_return_inside_the_surface_is_the_identity():
    for model in (8, 9):
        s = np.array([1.001, .9995, 1.0])
        sig, jp, inc = o.pressure_return(s, model, 1.0, 0.0)
        np.testing.assert_allclose(sig, s, rtol=1e-14)
        assert jp == pytest.approx(1.0, rel=1e-14) and inc == 0


def test_pressure_return_scales_with_the_part_moduli():
    # G, K follow mu/E and lambda/E: a stiffer part at the same strain is
    # further outside the surface and returns more.
    F = np.array([1.0, .94, 1.06])
    _, _, inc_ref = o.pressure_return(F, 8, 1.0, 0.0)
    _, _, inc_stiff = o.pressure_return(F, 8, 1.0, 0.0, G=.8, L=.8)
    assert inc_stiff > inc_ref > 0


# tests/course.test.mjs:49-55
def test_damage_is_objective_irreversible_and_fails_to_sand():
    part = o.make_part(fracture=o.fracture_params(enabled=1, onset=.04, failure=.1))
    sc = o.Scene(np.zeros((1, 3)), (2, 3, 0), 1.1, 1.0, lower=(-64, -48, -64), upper=(64, 72, 64), h=8,
                 parts=[part])
    m = sc.mass[0]

    def upd(F):
        sc.F[:] = F
        o.lib().orc_update_damage(o.C.byref(sc.s), 0)

    upd([0, -1, 0, 1, 0, 0, 0, 0, 1])
    assert sc.damage[0] == 0
    upd([math.exp(.07), 0, 0, 0, 1, 0, 0, 0, 1])
    assert abs(sc.damage[0] - .5) < 1e-8
    upd(np.eye(3).reshape(9))
    assert sc.damage[0] > .49
    upd([math.exp(.11), 0, 0, 0, 1, 0, 0, 0, 1])
    assert sc.damage[0] == 1 and sc.model[0] == 3 and sc.mass[0] == m and sc.vx[0] == 2 and sc.vy[0] == 3
    np.testing.assert_array_equal(sc.F, np.eye(3).reshape(9))
    assert sc.Jp[0] == 1 and sc.dilation[0] == 0 and sc.cacheValid[0] == 0
    sc.refresh_diagnostics()
    assert sc.stress[0] < 1e-8


def test_damage_softens_the_step_stress():
    # tau carries (1 - damage) below failure (src/three/mpm.js:38).
    part = o.make_part(fracture=o.fracture_params(enabled=1, onset=.04, failure=.1))
    sc = o.Scene(np.zeros((1, 3)), 0, 1.1, 1.0, lower=(-64, -48, -64), upper=(64, 72, 64), h=8, parts=[part])
    sc.F[:] = [1.02, .01, 0, 0, .99, 0, 0, 0, 1]
    sc.refresh_diagnostics()
    o.lib().orc_p2g(o.C.byref(sc.s), 1e-3)
    t0 = sc.tau.copy()
    sc.damage[0] = .3
    o.lib().orc_p2g(o.C.byref(sc.s), 1e-3)
    np.testing.assert_allclose(sc.tau, .7 * t0, rtol=1e-14)


def test_failed_particle_steps_as_sand():
    # A stretched block whose every particle fails in the first G2P steps on
    # as model 3 with F reset to the identity.
    pts = np.array([[x, y, z] for z in (-4, 0, 4) for y in (0, 4, 8) for x in (-4, 0, 4)], float)
    v = np.stack([pts[:, 0] * 40, np.zeros(27), np.zeros(27)], 1)  # strain rate 40/s in x
    part = o.make_part(fracture=o.fracture_params(enabled=1, onset=.005, failure=.01))
    sc = o.Scene(pts, v, 1.1, 1.0, lower=(-64, -48, -64), upper=(64, 72, 64), h=8, parts=[part])
    for _ in range(6):
        sc.step(1 / 240)
    assert (sc.damage > 0).all()
    failed = sc.damage >= 1
    assert failed.any()
    assert (sc.model[failed] == 3).all() and (sc.model[~failed] == 0).all()
    assert np.isfinite(sc.positions()).all()
