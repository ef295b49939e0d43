"""GPU parity of the Neo-Hookean (1), snow (2), sand (3), cohesive
Drucker-Prager (8) and crushable-foam (9) models and of fracture damage
against the FP64 oracle, through the C ABI.

  * return maps on batches of trial states (mpm_return_map_batch, Jp = 1 and
    dilation = 0 in): F atol 4e-6 rtol 5e-5, plastic increment 5e-6, stress
    atol 3e-6 rtol 5e-4 in units of E -- the bars of the reference's return
    map parity tests (gpu/test_thermal_plasticity.py:23);
  * Neo-Hookean step stress: within 2e-5 of the largest component;
  * two blocks colliding head-on at 120 units/s in the reference's units
    (cell 8, mass 1.1, particle volume 1, E = m/(64 V) 16000, mu = lambda =
    0.4 E), 40 steps of 1/240 through the collision: relative L2 of the
    displacement, velocity, Jp - 1 (snow) and dilation (sand) within 5e-3;
    grid mass equal to particle mass exactly; the stable step (snow
    hardening, Neo-Hookean compression factor) within 1e-3 relative.
"""
import math

import numpy as np
import pytest

from mpm_amd import Engine, Material
from mpm_amd.engine import return_map_batch, return_map_state_batch, stress_batch, sym6_to_full
from oracle import binding as o

pytestmark = pytest.mark.gpu

LOWER, UPPER, H = (-64, -48, -64), (64, 72, 64), 8.0
E_REF = 1.1 / 64 * 16000


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def trial_states(seed, n=512, spread=.06):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        A = np.eye(3) + rng.normal(0, spread, (3, 3))
        th = rng.uniform(0, 2 * np.pi)
        R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
        out.append((R @ A).reshape(9))
    return np.array(out)


def check_return(model, Fs, ys=.025, friction=35.0, oracle_ys=None, oracle_friction=None):
    mat = Material(model, 1.0, .4, .4, yield_strain=ys, friction=friction)
    Fo, epo, tau = return_map_batch(mat, Fs, 0.0, 20.0, 1 / 240)
    oys = ys if oracle_ys is None else oracle_ys
    ofr = friction if oracle_friction is None else oracle_friction
    for k in range(len(Fs)):
        r = o.project_plastic(Fs[k], model, 1, 0, oys, ofr, 0, 1 / 240, 20)
        np.testing.assert_allclose(Fo[k], r["F"], atol=4e-6, rtol=5e-5, err_msg=f"case {k}")
        assert abs(epo[k] - r["plasticIncrement"]) < 5e-6, k
        want = o.constitutive_stress(r["F"], model, .4, .4, r["jp"])
        np.testing.assert_allclose(sym6_to_full(tau[k]), want, atol=3e-6, rtol=5e-4, err_msg=f"case {k}")


def test_snow_return_map_matches_oracle(gpu):
    # The reference's snow cases (tests/constitutive.test.mjs:26-33) plus
    # rotated random trial states, at two yield strains.
    ref = np.array([[.8, 0, 0, 0, .9, 0, 0, 0, 1.2], [.96, 0, 0, 0, 1, 0, 0, 0, 1]])
    check_return(2, ref)
    check_return(2, ref, ys=.01)
    check_return(2, ref, ys=.06)
    check_return(2, trial_states(21))
    check_return(2, trial_states(22), ys=.05)


def test_snow_material_default_and_clamp(gpu):
    # yield_strain 0 is the reference default .025; values are clamped to
    # [.005, .08] as the reference clamps them when it adds a particle
    # (src/three/simulation.js:40).
    Fs = trial_states(23, 64)
    check_return(2, Fs, ys=0.0, oracle_ys=.025)
    check_return(2, Fs, ys=.2, oracle_ys=.08)


def test_sand_return_map_matches_oracle(gpu):
    # The reference's sand cases (tests/constitutive.test.mjs:34-43): tension
    # returns to the stress-free state with a dilation; compression-shear
    # returns onto the Drucker-Prager cone, closer for higher friction.
    ref = np.array([[1.1, 0, 0, 0, 1.1, 0, 0, 0, 1.1], [.8, 0, 0, 0, 1.05, 0, 0, 0, 1.05]])
    for fr in (15.0, 35.0, 50.0):
        check_return(3, ref, friction=fr)
    check_return(3, trial_states(31))
    check_return(3, trial_states(32, spread=.1), friction=20.0)
    Fo, _, tau = return_map_batch(Material(3, 1.0, .4, .4), ref[:1], 0.0, 20.0, 1 / 240)
    assert np.abs(tau).max() < 1e-6
    np.testing.assert_allclose(Fo[0], np.eye(3).reshape(9), atol=1e-6)


def test_neo_hookean_stress_matches_oracle(gpu):
    Fs = np.concatenate([trial_states(41, 256, .08), trial_states(42, 64, .25)])
    mu, lam = .4 * E_REF, .4 * E_REF
    tau = stress_batch(Material(1, E_REF, mu, lam), Fs)
    for k in range(len(Fs)):
        want = o.constitutive_stress(Fs[k], 1, mu, lam)
        assert np.abs(sym6_to_full(tau[k]) - want).max() <= 2e-5 * np.abs(want).max() + 1e-4, k


def colliding_blocks(nb=5, sp=4.0, gap=6.0, speed=60.0):
    """Two nb^3 lattice blocks at spacing sp moving toward each other."""
    pts, vel = [], []
    w = (nb - 1) * sp
    for s in (-1, 1):
        x0 = s * (gap / 2 + w / 2)
        for z in range(nb):
            for y in range(nb):
                for x in range(nb):
                    pts.append([x0 - w / 2 + sp * x, -w / 2 + sp * y + 8, -w / 2 + sp * z])
                    vel.append([-s * speed, 0, 0])
    return np.array(pts), np.array(vel)


@pytest.mark.parametrize("model", [1, 2, 3])
def test_colliding_blocks_match_oracle(gpu, model):
    pts, vel = colliding_blocks()
    n = len(pts)
    eng = Engine(LOWER, UPPER, H, n, strict_domain=True)
    eng.set_materials([Material(model, E_REF, .4 * E_REF, .4 * E_REF)])
    eng.upload(pts, vel, 1.1, 64.0)
    orc = o.Scene(pts, vel, 1.1, 1.0, lower=LOWER, upper=UPPER, h=H, model=model)
    xg0 = eng.download()["positions"]
    dt = 1 / 240
    for k in range(40):
        eng.step(dt, 1)
        orc.step(dt)
        if k in (9, 39):
            s_g, s_o = eng.stable_step(1.0), orc.stable_step(1.0)  # uncapped: the wave-speed bound decides
            assert abs(s_g - s_o) <= 1e-3 * s_o, (k, s_g, s_o)
    st = eng.download()
    e_dx = rel_l2(st["positions"] - xg0, orc.positions() - pts)
    e_v = rel_l2(st["velocities"], orc.velocities())
    msg = f"model {model} rel L2 dx {e_dx:.2e} v {e_v:.2e}"
    if model == 2:
        e_jp = rel_l2(st["jp"] - 1, orc.Jp - 1)
        msg += f" Jp-1 {e_jp:.2e} (oracle Jp min {orc.Jp.min():.3f})"
        assert orc.Jp.min() < .9 and e_jp < 5e-3, msg
    if model == 3:
        e_dil = rel_l2(st["dilation"], orc.dilation)
        msg += f" dilation {e_dil:.2e} (oracle max {orc.dilation.max():.3f})"
        assert orc.dilation.max() > .05 and e_dil < 5e-3, msg
    print(msg)
    assert e_dx < 5e-3 and e_v < 5e-3, msg
    c = eng.conservation()
    assert c["grid_mass_q"] == c["particle_mass_q"] and c["nonfinite"] == 0


# ------------------------------------------------ pressure models (8, 9)
def check_pressure(model, Fs, jp, ep, pressure=None):
    pr = dict(o.PRESSURE_DEFAULTS)
    pr.update(pressure or {})
    mat = Material(model, 1.0, .4, .4, pressure=pr)
    Fo, epo, jpo, dilo, tau = return_map_state_batch(mat, Fs, ep, 20.0, jp, 0.0, .01)
    part = o.make_part(pressure=o.pressure_params(**pr))
    for k in range(len(Fs)):
        r = o.project_plastic(Fs[k], model, jp, 0, .025, 35, ep, .01, 20, part)
        np.testing.assert_allclose(Fo[k], r["F"], atol=4e-6, rtol=5e-5, err_msg=f"model {model} case {k}")
        assert abs(jpo[k] - r["jp"]) <= 5e-5 * r["jp"], (model, k, jpo[k], r["jp"])
        assert abs(epo[k] - (ep + r["plasticIncrement"])) < 5e-6, (model, k)
        want = o.constitutive_stress(r["F"], model, .4, .4)
        np.testing.assert_allclose(sym6_to_full(tau[k]), want, atol=3e-6, rtol=5e-4, err_msg=f"model {model} {k}")


def diag(x, y=None, z=None):
    return [x, 0, 0, 0, x if y is None else y, 0, 0, 0, x if z is None else z]


def test_cohesive_dp_return_map_matches_oracle(gpu):
    # The reference's cases (tests/pressure-plasticity.test.mjs:11-24):
    # tension cut-off, compression inside the cone, shear with and without
    # dilatancy, a stronger cohesion; then rotated random trial states.
    ref = np.array([diag(1.1), diag(.8), diag(1.3, .8, 1), [1, .4, 0, 0, 1, 0, 0, 0, 1], [1, .5, 0, 0, 1, 0, 0, 0, 1]])
    check_pressure(8, ref, 1.0, 0.0)
    check_pressure(8, ref, 1.0, 0.0, dict(dilatancy=0.0))
    check_pressure(8, ref, 1.0, 0.0, dict(dilatancy=.5))
    check_pressure(8, ref, 1.0, 0.0, dict(cohesion=.1))
    check_pressure(8, trial_states(81, 512, .08), 1.0, 0.0)
    check_pressure(8, trial_states(82, 256, .08), 1.3, .02)


def test_foam_return_map_matches_oracle(gpu):
    # tests/pressure-plasticity.test.mjs:26-47: hydrostatic and mixed
    # compression at Jp .6, 1, 1.2 with ep .05, the associated-flow case,
    # then random states and other shapes of the yield ellipse.
    ref = np.array([diag(.85), diag(1.1), [1, .45, 0, 0, .95, 0, 0, 0, 1.1], diag(.85, .97, 1.02)])
    for jp in (.6, 1.0, 1.2):
        check_pressure(9, ref, jp, .05)
    check_pressure(9, ref, 1.0, .3)
    check_pressure(9, trial_states(91, 512, .08), 1.0, 0.0)
    check_pressure(9, trial_states(92, 256, .08), .8, .1, dict(alpha=.5))
    check_pressure(9, trial_states(93, 256, .08), 1.0, .2, dict(alpha=2.0, beta=3.0))


@pytest.mark.parametrize("model", [8, 9])
def test_pressure_model_blocks_match_oracle(gpu, model):
    pts, vel = colliding_blocks()
    n = len(pts)
    eng = Engine(LOWER, UPPER, H, n, strict_domain=True)
    eng.set_materials([Material(model, E_REF, .4 * E_REF, .4 * E_REF)])
    eng.upload(pts, vel, 1.1, 64.0)
    orc = o.Scene(pts, vel, 1.1, 1.0, lower=LOWER, upper=UPPER, h=H, model=model)
    xg0 = eng.download()["positions"]
    for k in range(40):
        eng.step(1 / 240, 1)
        orc.step(1 / 240)
        if k in (9, 39):
            s_g, s_o = eng.stable_step(1.0), orc.stable_step(1.0)  # J Jp compression factor
            assert abs(s_g - s_o) <= 1e-3 * s_o, (k, s_g, s_o)
    st = eng.download()
    e_dx = rel_l2(st["positions"] - xg0, orc.positions() - pts)
    e_v = rel_l2(st["velocities"], orc.velocities())
    e_jp = rel_l2(st["jp"] - 1, orc.Jp - 1)
    e_ep = rel_l2(st["plastic"], orc.plastic)
    msg = (f"model {model} rel L2 dx {e_dx:.2e} v {e_v:.2e} Jp-1 {e_jp:.2e} ep {e_ep:.2e} "
           f"(oracle Jp {orc.Jp.min():.3f}..{orc.Jp.max():.3f}, ep max {orc.plastic.max():.3f})")
    print(msg)
    assert np.abs(orc.Jp - 1).max() > 1e-3 and orc.plastic.max() > 1e-3, msg
    assert e_dx < 5e-3 and e_v < 5e-3 and e_jp < 5e-3 and e_ep < 5e-3, msg
    vm = eng.refresh_diagnostics()  # von Mises over J Jp, in units of E
    orc.refresh_diagnostics()
    np.testing.assert_allclose(vm, orc.stress, rtol=5e-3, atol=1e-4 * orc.stress.max())
    c = eng.conservation()
    assert c["grid_mass_q"] == c["particle_mass_q"] and c["nonfinite"] == 0


# ------------------------------------------------------- fracture damage
def test_fracture_block_matches_oracle(gpu):
    # A 5^3 block stretched at 6/s along x with damage onset .02 and failure
    # .06 (src/three/fracture.js:11-18): damage grows for three steps, most
    # particles fail in the fourth and all by the fifth (the damage rises
    # about .4 per step, so failure is never marginal), then the block steps
    # on as sand. Bars: the same failed set after every step, damage within
    # 1e-3, relative L2 of displacement and velocity within 5e-3 after 12
    # steps.
    nb, sp = 5, 4.0
    w = (nb - 1) * sp
    ax = np.arange(nb) * sp - w / 2
    pts = np.array([[x, y + 8, z] for z in ax for y in ax for x in ax])
    v = np.stack([pts[:, 0] * 6.0, np.zeros(len(pts)), np.zeros(len(pts))], 1)
    fr = dict(enabled=True, onset=.02, failure=.06)
    eng = Engine(LOWER, UPPER, H, len(pts), strict_domain=True)
    eng.set_materials([Material(0, E_REF, .4 * E_REF, .4 * E_REF, fracture=fr)])
    eng.upload(pts, v, 1.1, 64.0)
    part = o.make_part(fracture=o.fracture_params(enabled=1, onset=.02, failure=.06))
    orc = o.Scene(pts, v, 1.1, 1.0, lower=LOWER, upper=UPPER, h=H, parts=[part])
    xg0 = eng.download()["positions"]
    for k in range(12):
        eng.step(1 / 240, 1)
        orc.step(1 / 240)
        st = eng.download()
        np.testing.assert_array_equal(st["damage"] >= 1, orc.damage >= 1, err_msg=f"step {k}")
        np.testing.assert_allclose(st["damage"], orc.damage, atol=1e-3, err_msg=f"step {k}")
    assert (orc.damage >= 1).all() and (orc.model == 3).all()
    np.testing.assert_allclose(st["F"].reshape(-1, 9), orc.F.reshape(-1, 9), atol=1e-4)
    e_dx = rel_l2(st["positions"] - xg0, orc.positions() - pts)
    e_v = rel_l2(st["velocities"], orc.velocities())
    print(f"fracture block rel L2 dx {e_dx:.2e} v {e_v:.2e}; dilation max {orc.dilation.max():.3f}")
    assert e_dx < 5e-3 and e_v < 5e-3
    c = eng.conservation()
    assert c["grid_mass_q"] == c["particle_mass_q"] and c["nonfinite"] == 0
