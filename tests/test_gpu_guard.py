"""The state guard (status codes 7 and 8 of the reference GPU service,
gpu/mixed_solver.py:381-388) and the FP64 mechanical diagnostics of the
device, through the C ABI.

Guard bars: the CFL flag agrees with the stable step of the same state
(mpm_stable_dt, itself checked against the oracle) on either side of a 2%
margin; the invalid-state flag fires for det F < 1e-6 and |v| > 1e5 and not
for any state of a normal run; findings are cleared once reported.
Diagnostic bars (FP32 state vs the FP64 oracle): angular momentum with the
APIC spin term, kinetic and affine energy within 1e-3 relative of the
oracle's particleMechanics restatement; in free space the device's angular
momentum stays within 2e-5 of its start over 50 steps.
"""
import numpy as np
import pytest

from mpm_amd import Batch, Engine, Material, scenes
from mpm_amd import lib as L
from mpm_amd.engine import make_config
from oracle import binding as o

pytestmark = pytest.mark.gpu


def cube_engine(sc, **kw):
    eng = Engine(sc.lower, sc.upper, sc.h, sc.n, wall_mode=sc.wall_mode, gravity=sc.gravity,
                 gravity_on_particles=sc.gravity_on_particles, **kw)
    eng.set_materials(sc.materials)
    return eng


def test_cfl_guard_agrees_with_the_stable_step(gpu):
    sc = scenes.elastic_cube()
    eng = cube_engine(sc)
    eng.upload(sc.x, sc.v, sc.mass, sc.volume)
    # the stable step without the cap term, at the reference's Courant number
    stable = eng.stable_step(cap=1e9, cfl=1.4 / 8)
    orc = o.Scene(sc.x, sc.v, sc.mass, sc.volume, lower=sc.lower, upper=sc.upper, h=sc.h, vol_scale=1.0,
                  parts=[o.make_part(E=sc.materials[0].E, G=sc.materials[0].mu / sc.materials[0].E,
                                     L=sc.materials[0].lam / sc.materials[0].E)])
    assert stable == pytest.approx(orc.stable_step(1e9, 1.4 / 8), rel=1e-5)
    eng.step(0.98 * stable, 1)
    eng.synchronize()  # no finding
    eng.upload(sc.x, sc.v, sc.mass, sc.volume)
    eng.step(1.02 * stable, 1)
    with pytest.raises(L.MpmError) as err:
        eng.synchronize()
    assert err.value.code == L.MPM_ERR_CFL
    eng.synchronize()  # reported once, then cleared
    # the guard switched off reports nothing
    off = cube_engine(sc, guard=False)
    off.upload(sc.x, sc.v, sc.mass, sc.volume)
    off.step(1.02 * stable, 1)
    off.synchronize()


def test_dt_no_state_can_take_is_refused_at_once(gpu):
    sc = scenes.elastic_cube()
    eng = cube_engine(sc)
    eng.upload(sc.x, sc.v, sc.mass, sc.volume)
    with pytest.raises(L.MpmError) as err:
        eng.step(0.2 * sc.h, 1)  # the wave term's floor: 1.4/8 h / 1
    assert err.value.code == L.MPM_ERR_CFL
    capped = cube_engine(sc, cfl_cap=5e-5)
    capped.upload(sc.x, sc.v, sc.mass, sc.volume)
    with pytest.raises(L.MpmError) as err:
        capped.step(1e-4, 1)
    assert err.value.code == L.MPM_ERR_CFL
    capped.step(4e-5, 3)
    capped.synchronize()
    st = capped.download()
    assert np.all(np.isfinite(st["positions"]))


def test_inverted_particle_is_an_invalid_state(gpu):
    sc = scenes.elastic_cube()
    eng = cube_engine(sc)
    F = np.tile(np.eye(3).reshape(9), (sc.n, 1))
    F[17] = np.diag([-1.0, 1.0, 1.0]).reshape(9)  # det F = -1: finite, inside the domain
    eng.upload(sc.x, sc.v, sc.mass, sc.volume, F=F)
    eng.step(sc.dt, 1)
    with pytest.raises(L.MpmError) as err:
        eng.synchronize()
    assert err.value.code == L.MPM_ERR_INVALID_STATE
    eng.upload(sc.x, sc.v, sc.mass, sc.volume)
    eng.step(sc.dt, 5)
    eng.synchronize()


def test_runaway_velocity_is_an_invalid_state(gpu):
    # One isolated particle at 2e5 m/s: its wave speed breaks the CFL limit
    # (7) and after the step |v| > 1e5 (8); 8 wins, as in check_state.
    x = np.array([[0.3, 0.5, 0.5], [0.7, 0.5, 0.5]])
    v = np.array([[2e5, 0.0, 0.0], [0.0, 0.0, 0.0]])
    eng = Engine((0, 0, 0), (1, 1, 1), 1 / 32, 2)
    eng.set_materials([Material.from_E_nu(0, 1e5, 0.3)])
    eng.upload(x, v, np.full(2, 1e-3), np.full(2, 1e-6))
    eng.step(1e-6, 1)
    with pytest.raises(L.MpmError) as err:
        eng.synchronize()
    assert err.value.code == L.MPM_ERR_INVALID_STATE
    st = eng.download()
    assert np.abs(st["velocities"][0, 0]) > 1e5 and st["positions"][0, 0] < 1.0


def test_batch_reports_the_scene_that_went_bad(gpu):
    mats = [Material.from_E_nu(0, 1e5, 0.3)]
    sc = scenes.elastic_cube()
    cfgs = [make_config(sc.lower, sc.upper, sc.h, sc.n, wall_mode=sc.wall_mode, gravity=sc.gravity)
            for _ in range(3)]
    b = Batch(cfgs, [sc.n] * 3, sort_every=4)
    b.set_materials(mats)
    for s in range(3):
        F = np.tile(np.eye(3).reshape(9), (sc.n, 1))
        if s == 1:
            F[5] = np.diag([1.0, 1e-7, 1.0]).reshape(9)  # det F = 1e-7 < 1e-6
        b.upload(s, sc.x, sc.v, sc.mass, sc.volume, F=F)
    b.set_dt([sc.dt] * 3)
    b.step(6)
    with pytest.raises(L.MpmError) as err:
        b.synchronize()
    assert err.value.code == L.MPM_ERR_INVALID_STATE and "scene 1" in str(err.value)
    b.synchronize()
    with pytest.raises(L.MpmError) as err:
        b.set_dt([sc.dt, 0.2 * sc.h, sc.dt])
    assert err.value.code == L.MPM_ERR_CFL


def diag_of(orc):
    m = orc.mechanics()
    return m["angular"], m["kinetic"], m["affine"], m["momentum"]


def test_diagnostics_match_oracle_on_the_elastic_cube(gpu):
    # Elastic cube (32^3 cells, ~7k particles, fixed-corotated, nu = 0.25,
    # gravity on the particles), 200 steps: the device's FP64 sums of the FP32
    # state against the oracle's particleMechanics restatement.
    sc = scenes.elastic_cube(nu=0.25, sticky_floor=False)
    eng = sc.engine()
    mat = sc.materials[0]
    orc = o.Scene(sc.x, sc.v, sc.mass, sc.volume, lower=sc.lower, upper=sc.upper, h=sc.h, vol_scale=1.0,
                  parts=[o.make_part(E=mat.E, G=mat.mu / mat.E, L=mat.lam / mat.E)], temperature=sc.temperature)
    for k in range(4):
        eng.step(sc.dt, 50)
        for _ in range(50):
            orc.step(sc.dt, particle_gravity=sc.gravity)
        c = eng.conservation()
        ang, kin, aff, mom = diag_of(orc)
        xs = np.abs(orc.positions()).max()
        ang_scale = xs * np.abs(mom).sum() + np.abs(ang).max()
        assert np.abs(c["angular"] - ang).max() <= 1e-3 * ang_scale, (k, c["angular"], ang)
        assert c["kinetic"] == pytest.approx(kin, rel=1e-3)
        assert c["affine"] == pytest.approx(aff, rel=2e-3, abs=1e-3 * kin)
        assert np.abs(c["momentum"] - mom).max() <= 1e-3 * np.abs(mom).max()
    eng.synchronize()


def test_spinning_block_keeps_its_angular_momentum(gpu):
    # A stretched, spinning block in free space (no gravity, walls out of
    # reach): APIC conserves the total angular momentum including the spin
    # term; the device agrees with the oracle on it and on the energies.
    h = 1 / 32
    g = np.arange(-0.1, 0.1 + 1e-9, h / 2)
    pts = np.array([[0.5 + x, 0.5 + y, 0.5 + z] for x in g for y in g for z in g if abs(z) <= 0.05])
    n = len(pts)
    vol = (h / 2) ** 3
    omega = np.array([0.3, -0.2, 2.0])
    v = np.cross(omega, pts - pts.mean(0)) + np.array([0.05, 0.0, -0.03])
    F = np.tile(np.diag([1.03, 0.98, 1.0]).reshape(9), (n, 1))
    mat = Material.from_E_nu(0, 2e4, 0.3)
    eng = Engine((0, 0, 0), (1, 1, 1), h, n, mass_eps=0.0)
    eng.set_materials([mat])
    eng.upload(pts, v, np.full(n, 1000 * vol), np.full(n, vol), F=F)
    orc = o.Scene(pts, v, 1000 * vol, vol, lower=(0, 0, 0), upper=(1, 1, 1), h=h, vol_scale=1.0, mass_eps=0.0,
                  parts=[o.make_part(E=mat.E, G=mat.mu / mat.E, L=mat.lam / mat.E)])
    orc.F[:] = F.reshape(-1)
    L0 = eng.conservation()["angular"]
    assert np.allclose(L0, orc.mechanics()["angular"], rtol=1e-5)
    for _ in range(5):
        eng.step(2e-4, 10)
        for _ in range(10):
            orc.step(2e-4)
        c = eng.conservation()
        ang, kin, aff, _ = diag_of(orc)
        scale = np.linalg.norm(L0)
        assert np.linalg.norm(c["angular"] - L0) <= 2e-5 * scale
        assert np.linalg.norm(c["angular"] - ang) <= 2e-5 * scale
        assert c["kinetic"] == pytest.approx(kin, rel=1e-3)
        assert c["affine"] == pytest.approx(aff, rel=1e-3)
    eng.synchronize()
