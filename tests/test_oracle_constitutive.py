"""Pins the oracle's constitutive restatement against the reference's own tests.

Each test restates one reference test (file:line) with its inputs and
thresholds, plus checks the mathematics fixes (LAPACK singular values, the
SVD polar factor, closed forms). CPU only.
"""
import math

import numpy as np
import pytest

from oracle import binding as o

close = lambda a, b, eps=1e-8: abs(a - b) < eps  # noqa: E731


def mm(a, b):
    return (np.asarray(a).reshape(3, 3) @ np.asarray(b).reshape(3, 3)).reshape(9)


def tr(a):
    return np.asarray(a).reshape(3, 3).T.reshape(9)


def det(a):
    return float(np.linalg.det(np.asarray(a, float).reshape(3, 3)))


# tests/constitutive.test.mjs:10-16
@pytest.mark.parametrize("F", [[1, .8, .2, 0, .7, .4, .1, 0, 1.2], [.1, 0, 0, 0, .12, 0, 0, 0, 2],
                               [0, -1, 0, 1, 0, 0, 0, 0, 1]])
def test_svd_reconstructs_with_orthogonal_factors(F):
    u, v, s = o.svd3(F)
    rec = o.compose(u, v, s)
    assert np.all(np.abs(rec - np.asarray(F)) < 1e-8)
    for m in (u, v):
        assert np.all(np.abs(mm(m, tr(m)) - np.eye(3).reshape(9)) < 1e-8)
    # singular values agree with LAPACK
    ref = np.linalg.svd(np.asarray(F, float).reshape(3, 3), compute_uv=False)
    assert np.allclose(np.sort(s), np.sort(ref), rtol=1e-12, atol=1e-14)


def test_svd_random_matrices_match_lapack():
    rng = np.random.default_rng(1)
    for _ in range(200):
        F = np.eye(3) + rng.normal(0, .3, (3, 3))
        u, v, s = o.svd3(F.reshape(9))
        ref = np.linalg.svd(F, compute_uv=False)
        assert np.allclose(np.sort(s), np.sort(ref), rtol=1e-10, atol=1e-12)
        assert np.allclose(o.compose(u, v, s), F.reshape(9), atol=1e-10)


# tests/three.test.mjs:8-13
def test_stress3_rotation_invariant_and_shear_restoring():
    c, s = math.cos(.8), math.sin(.8)
    r = [1, 0, 0, 0, c, -s, 0, s, c]
    assert np.all(np.abs(o.stress3(r, 400, 400)) < 1e-8)
    f = list(np.eye(3).reshape(9))
    f[2] = .2
    t = o.stress3(f, 400, 400)
    assert t[2] > 0 and abs(t[2] - t[6]) < 1e-8
    f[5] = .3
    assert o.stress3(f, 400, 400)[5] > 0


def test_stress3_closed_form_uniaxial():
    mu, lam = 40.0, 60.0
    for s in (.8, .97, 1.0, 1.05, 1.4):
        t = o.stress3([s, 0, 0, 0, 1, 0, 0, 0, 1], mu, lam).reshape(3, 3)
        J = s
        assert abs(t[0, 0] - (2 * mu * (s - 1) * s + lam * (J - 1) * J)) < 1e-10
        assert abs(t[1, 1] - lam * (J - 1) * J) < 1e-10 and abs(t[2, 2] - lam * (J - 1) * J) < 1e-10
        assert np.all(np.abs(t - np.diag(np.diag(t))) < 1e-12)


def test_polar_rotation_matches_svd_polar_factor():
    rng = np.random.default_rng(7)
    for _ in range(100):
        A = rng.normal(0, .4, (3, 3)) + np.eye(3)
        if np.linalg.det(A) <= 0.05:
            continue
        U, S, Vt = np.linalg.svd(A)
        R = U @ Vt
        assert np.allclose(o.rotation(A.reshape(9)), R.reshape(9), atol=1e-7)
    assert np.allclose(o.rotation(np.zeros(9)), np.eye(3).reshape(9))  # singular -> identity


# tests/constitutive.test.mjs:17-25
def test_all_laws_objective_and_stress_free_under_rotation():
    c, s = math.cos(.7), math.sin(.7)
    R = [c, 0, s, 0, 1, 0, -s, 0, c]
    F = [1.03, .02, 0, 0, .97, 0, 0, 0, .99]
    for model in range(12):
        assert np.all(np.abs(o.constitutive_stress(R, model, 400, 400)) < 1e-8), model
        t = o.constitutive_stress(F, model, 400, 400)
        rotated = o.constitutive_stress(mm(R, F), model, 400, 400)
        expected = mm(mm(R, t), tr(R))
        assert np.all(np.abs(rotated - expected) < 1e-8), model
    a = o.constitutive_stress([1.5, 0, 0, 0, 1, 0, 0, 0, 1], 0, 400, 400)[0]
    b = o.constitutive_stress([1.5, 0, 0, 0, 1, 0, 0, 0, 1], 1, 400, 400)[0]
    assert abs(a - b) > 100


def test_hencky_stress_closed_form():
    # Hencky: tau = 2 mu ln(s) + lambda tr(ln s) on the principal axes.
    mu, lam = 1.0, .6
    sig = np.array([1.1, .95, 1.02])
    t = o.constitutive_stress(np.diag(sig).reshape(9), 10, mu, lam).reshape(3, 3)
    e = np.log(sig)
    assert np.allclose(np.diag(t), 2 * mu * e + lam * e.sum(), atol=1e-12)


# tests/constitutive.test.mjs:26-43 (snow and sand branches)
def test_snow_and_sand_returns():
    F = [.8, 0, 0, 0, .9, 0, 0, 0, 1.2]
    p = o.project_plastic(F, 2)
    u, v, sing = o.svd3(p["F"])
    assert np.all((sing >= .975 - 1e-8) & (sing <= 1.0075 + 1e-8))
    assert close(det(p["F"]) * p["jp"], det(F)) and p["jp"] < 1
    soft = o.project_plastic([.96, 0, 0, 0, 1, 0, 0, 0, 1], 2, 1, 0, .01)
    strong = o.project_plastic([.96, 0, 0, 0, 1, 0, 0, 0, 1], 2, 1, 0, .06)
    assert soft["jp"] < strong["jp"]
    tensile = o.project_plastic([1.1, 0, 0, 0, 1.1, 0, 0, 0, 1.1], 3)
    assert np.all(np.abs(o.constitutive_stress(tensile["F"], 3, 400, 400)) < 1e-8) and tensile["dilation"] > 0
    F2 = [.8, 0, 0, 0, 1.05, 0, 0, 0, 1.05]
    low = o.project_plastic(F2, 3, 1, 0, .025, 15)
    high = o.project_plastic(F2, 3, 1, 0, .025, 50)

    def dev_norm(f):
        e = np.log(o.svd3(f)[2])
        return np.linalg.norm(e - e.mean())
    assert dev_norm(high["F"]) > dev_norm(low["F"])
    assert close(det(low["F"]), det(F2)) and close(det(high["F"]), det(F2))


# tests/thermal-plasticity.test.mjs:7-20
TF = [1, .2, 0, 0, 1, 0, 0, 0, 1]


def ret(model, T, ep=0.0, p=None, dt=.01, f=TF):
    return o.project_plastic(f, model, 1, 0, .025, 35, ep, dt, T, o.make_part(thermal=p or o.thermal_params()))


def vm_thermal(r):
    return o.von_mises(o.constitutive_stress(r["F"], 10, .4, .4), det(r["F"]))


def test_thermal_j2_hot_return_on_yield_surface():
    cold, hot = ret(10, 20), ret(10, 180)
    assert hot["plasticIncrement"] > cold["plasticIncrement"]
    assert vm_thermal(hot) < vm_thermal(cold)
    assert close(vm_thermal(hot), o.thermal_strength(hot["plasticIncrement"], 180), 1e-7)
    assert close(det(hot["F"]), det(TF), 1e-7)
    cooled = ret(10, 20, hot["plasticIncrement"], None, .01, hot["F"])
    assert close(cooled["plasticIncrement"], 0, 1e-7)
    assert close(vm_thermal(cooled), vm_thermal(hot), 1e-7)
    assert o.thermal_strength(hot["plasticIncrement"], 20) > o.thermal_strength(hot["plasticIncrement"], 180)
    no_soft = o.thermal_params(hotRatio=1.0)
    assert close(ret(10, 20, 0, no_soft)["plasticIncrement"], ret(10, 1000, 0, no_soft)["plasticIncrement"], 1e-7)
    for T in (-273.15, 20, 180, 1e6):
        r = ret(10, T)
        assert r["plasticIncrement"] >= 0 and close(det(r["F"]), 1, 1e-7)


# tests/thermal-plasticity.test.mjs:22-33
def test_creep_temperature_stress_time_and_closed_form():
    cold, hot, long = ret(11, 20), ret(11, 180), ret(11, 180, 0, None, .1)
    assert hot["plasticIncrement"] > cold["plasticIncrement"] * 5
    assert long["plasticIncrement"] > hot["plasticIncrement"]
    assert close(hot["plasticIncrement"], .01 * math.exp(o.creep_log_rate(vm_thermal(hot), 180)), 1e-8)
    assert close(det(hot["F"]), 1, 1e-7)
    assert ret(11, -100)["plasticIncrement"] < cold["plasticIncrement"]
    assert close(ret(11, 180, 0, o.thermal_params(rate=0.0))["plasticIncrement"], 0, 1e-7)
    assert close(ret(11, 180, 0, None, 0)["plasticIncrement"], 0, 1e-7)
    assert close(ret(11, 180, 0, None, .01, list(np.eye(3).reshape(9)))["plasticIncrement"], 0, 1e-7)
    p = o.thermal_params(n=1.0, Q=0.0, rate=.1)
    sigma = [math.exp(.1), math.exp(-.1), 1]
    q = math.sqrt(6) * .4 * math.sqrt(.02)
    dt = .02
    _, inc = o.thermal_return(sigma, 11, 0, dt, 20, p)
    expected = dt * p.rate * q / p.stress / (1 + dt * p.rate * 1.2 / p.stress)
    assert close(inc, expected, 1e-9)
    for T in (-273.15, 1e6):
        assert np.all(np.isfinite(ret(11, T)["F"]))


def test_thermal_return_generalised_shear_modulus():
    # With G = mu/E != 0.4 the return still lands on the yield surface:
    # q(after) = sqrt(6) G |dev| / J = strength(ep + inc, T).
    sigma = np.array([1.004, .998, .999])
    p = o.thermal_params(A=1.25e-3, H=0.0, Tref=300, Thot=1200, hotRatio=.1, kelvin_offset=0.0)
    for G in (.3846, .4, .25):
        out, inc = o.thermal_return(sigma, 10, 0, 1e-4, 600, p, G)
        e = np.log(out)
        dev = e - e.mean()
        q = math.sqrt(6) * G * np.linalg.norm(dev) / math.exp(e.sum())
        assert inc > 0
        assert abs(q - o.thermal_strength(inc, 600, p)) < 1e-12
        assert abs(np.log(out).sum() - np.log(sigma).sum()) < 1e-14  # isochoric


# tests/metal-plasticity.test.mjs:7-42
MF = [1.3, .1, 0, 0, .8, 0, 0, 0, 1 / 1.04]


def q_metal(f):
    return o.von_mises(o.constitutive_stress(f, 4, .4, .4), det(f))


def project(model, ep=0.0, dt=.01, T=20.0, p=None):
    return o.project_plastic(MF, model, 1, 0, .025, 35, ep, dt, T, o.make_part(plastic=p or o.plastic_params()))


def test_j2_radial_return_yield_consistency():
    for model in (4, 5):
        p = project(model)
        assert p["plasticIncrement"] > 0 and close(det(p["F"]), det(MF), 1e-7)
        assert close(q_metal(p["F"]), o.flow_stress(model, p["plasticIncrement"], 0, 20), 1e-7)
        un = o.project_plastic(list(np.eye(3).reshape(9)), model, 1, 0, .025, 35, p["plasticIncrement"])
        assert close(un["plasticIncrement"], 0, 1e-7) and close(q_metal(un["F"]), 0, 1e-7)
        hydro = o.project_plastic([.9, 0, 0, 0, .9, 0, 0, 0, .9], model)
        assert close(hydro["plasticIncrement"], 0, 1e-7) and close(det(hydro["F"]), .9 ** 3, 1e-7)
    assert project(5)["plasticIncrement"] < project(4)["plasticIncrement"]
    assert project(5, .5)["plasticIncrement"] < project(5, 0)["plasticIncrement"]


def test_johnson_cook_implicit_rate_and_temperature():
    for model in (6, 7):
        for T in (-20, 20, 650, 1000, 1200):
            for dt in (.0001, .01, .1):
                p = project(model, .12, dt, T)
                assert close(det(p["F"]), det(MF), 1e-7) and p["plasticIncrement"] >= 0
                assert close(q_metal(p["F"]), o.flow_stress(model, .12 + p["plasticIncrement"],
                                                            p["plasticIncrement"] / dt, T), 2e-7)
    assert q_metal(project(6, 0, .0001)["F"]) > q_metal(project(6, 0, .1)["F"])
    assert q_metal(project(6, 0, .01, 650)["F"]) < q_metal(project(6, 0, .01, 20)["F"])
    assert close(q_metal(project(6, 0, .01, 1000)["F"]), 0, 1e-7)
    assert close(q_metal(project(7, 0, .01, 650)["F"]), q_metal(project(7, 0, .01, 20)["F"]), 1e-7)
    assert close(o.flow_stress(6, 0, 0, 20), o.PLASTICITY_DEFAULTS["A"])
    no_effects = o.plastic_params(B=0.0, C=0.0)
    assert close(q_metal(project(6, 0, .01, 20, no_effects)["F"]), q_metal(project(4)["F"]), 1e-7)
    assert close(q_metal(project(7, 0, .01, 650, o.plastic_params(n=1.0, C=0.0))["F"]), q_metal(project(5)["F"]), 1e-7)


def test_cached_return_stress_equals_fresh_stress_of_returned_F():
    # principalStress on the returned decomposition (src/three/constitutive.js:81-86)
    # equals the Hencky stress recomputed from the returned F.
    for model, T in ((10, 180), (4, 20), (6, 650)):
        part = o.make_part()
        r = o.project_plastic(MF if model < 10 else TF, model, 1, 0, .025, 35, 0.0, .01, T, part)
        d = r["decomposition"]
        cached = o.principal_stress(d["u"], d["sigma"], model)
        fresh = o.constitutive_stress(r["F"], model, 1, 1)
        assert np.allclose(cached, fresh, atol=1e-10)
