"""GPU parity of the HIP engine against the FP64 oracle, through the C ABI.

Tolerances (FP32 device state vs the FP64 oracle):
  * binning keys and the sort permutation: bit-exact;
  * grid mass: fixed-point total equal to the particle total, exactly, every
    step; per-node mass vs oracle within 2e-6 of the largest node mass;
  * one P2G momentum: within 2e-5 of the largest node momentum;
  * return map F: atol 4e-6, rtol 5e-5 (thermal) / 4e-5 (J2, Johnson-Cook),
    plastic strain to 5 places, stress/E atol 3e-6 rtol 5e-4 -- the bars of
    the reference's own GPU tests (gpu/test_thermal_plasticity.py:23,
    gpu/test_metal_plasticity.py:31-36);
  * uniform translation: velocity atol 2e-5 (gpu/test_solver.py:74);
  * multi-step trajectories: relative L2 of displacement and of velocity
    within the per-test bound stated below.
"""
import math

import numpy as np
import pytest

from mpm_amd import Engine, Material, scenes, sort_pairs
from mpm_amd.engine import return_map_batch, stress_batch, sym6_to_full
from oracle import binding as o

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


# ----------------------------------------------------------------- sorting
@pytest.mark.parametrize("n,bits", [(1, 8), (7, 8), (2047, 12), (2048, 16), (2049, 16), (100_000, 22),
                                    ((1 << 20) + 13, 32)])
def test_radix_sort_equals_stable_argsort(gpu, n, bits):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
    ko, perm = sort_pairs(keys, bits)
    want = np.argsort(keys, kind="stable")
    assert np.array_equal(perm, want)
    assert np.array_equal(ko, keys[want])


def test_radix_sort_edge_cases(gpu):
    for keys, bits in [(np.zeros(5000, np.uint32), 8), (np.arange(9000, 0, -1, dtype=np.uint32), 16),
                       (np.full(3000, 0xFFFFFFFF, np.uint32), 32), (np.arange(4096, dtype=np.uint32) % 3, 8),
                       (np.zeros(0, np.uint32), 8)]:
        ko, perm = sort_pairs(keys, bits)
        want = np.argsort(keys, kind="stable")
        assert np.array_equal(perm, want) and np.array_equal(ko, keys[want])


def cpu_keys(x, lower, upper, h):
    origin, nodes = scenes.domain(lower, upper, h)
    N = int(nodes.max())
    x32 = np.asarray(x).astype(np.float32)
    O = origin.astype(np.float32)
    q = (x32 - O) * np.float32(1.0 / h)
    b = np.clip(np.floor(q - np.float32(0.5)).astype(np.int64), 0, N - 3)
    # Cells in 4x4x4 blocks (blocks in z, y, x order), x fastest inside.
    nb = (N + 3) // 4
    blk = ((b[:, 2] >> 2) * nb + (b[:, 1] >> 2)) * nb + (b[:, 0] >> 2)
    return (blk * 64 + (b[:, 2] & 3) * 16 + (b[:, 1] & 3) * 4 + (b[:, 0] & 3)).astype(np.uint32)


def test_binning_keys_and_permutation_bit_exact(gpu):
    sc = scenes.elastic_cube()
    eng = sc.engine()
    for _ in range(2):
        x = eng.download()["positions"]
        want = cpu_keys(x, sc.lower, sc.upper, sc.h)
        kin, kout, perm = eng.debug_bin()
        assert np.array_equal(kin, want)
        assert np.array_equal(perm, np.argsort(want, kind="stable"))
        assert np.array_equal(kout, want[perm])
        eng.step(sc.dt, 30)  # particles move; ties still resolve by id


# --------------------------------------------------------------- transfers
def oracle_for(sc, parity):
    mat = sc.materials[0]
    parts = [o.make_part(E=mat.E, G=mat.mu / mat.E, L=mat.lam / mat.E,
                         thermal=o.thermal_params(**{k: v for k, v in mat.thermal.items()}) if mat.thermal else None)]
    grid_g = (0, 0, 0) if sc.gravity_on_particles else sc.gravity
    return o.Scene(sc.x, sc.v, sc.mass, sc.volume, lower=sc.lower, upper=sc.upper, h=sc.h, vol_scale=1.0,
                   parts=parts, model=mat.model, temperature=sc.temperature, wall_mode=sc.wall_mode,
                   grid_gravity=grid_g)


def test_single_p2g_matches_oracle_grid(gpu):
    sc = scenes.elastic_cube(nu=0.25, sticky_floor=False)
    rng = np.random.default_rng(4)
    n = sc.n
    F = (np.eye(3) + rng.normal(0, .03, (n, 3, 3))).reshape(n, 9)
    Cm = rng.normal(0, 1.0, (n, 9))
    v = rng.normal(0, .05, (n, 3))
    eng = Engine(sc.lower, sc.upper, sc.h, n)
    eng.set_materials(sc.materials)
    eng.upload(sc.x, v, sc.mass, sc.volume, F=F, Cmat=Cm)
    gm, gp, N = eng.debug_p2g(sc.dt)
    orc = oracle_for(sc, True)
    orc.vx[:], orc.vy[:], orc.vz[:] = v[:, 0], v[:, 1], v[:, 2]
    orc.F[:] = F.reshape(-1)
    orc.C[:] = Cm.reshape(-1)
    orc.p2g(sc.dt)
    assert N == orc.N
    assert np.abs(gm - orc.gmass).max() <= 2e-6 * orc.gmass.max()
    op = orc.gv.reshape(-1, 3)
    assert np.abs(gp - op).max() <= 2e-5 * np.abs(op).max()


def test_uniform_translation(gpu):
    # tests/three.test.mjs:14-21 in reference units: 125 particles at spacing 4,
    # mass 1.1, V = 64 * 1, E = m/(64V) 16000, mu = lambda = 0.4 E.
    lo, hi = np.array([-8.] * 3), np.array([8.] * 3)
    g = np.arange(5) * 4.0 - 8
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    E = 1.1 / 64 * 16000
    eng = Engine((-64, -48, -64), (64, 72, 64), 8.0, len(x), strict_domain=True)
    eng.set_materials([Material(0, E, .4 * E, .4 * E)])
    eng.upload(x, (3, -2, 7), 1.1, 64.0)
    eng.step(.001, 1)
    v = eng.download()["velocities"]
    assert np.abs(v - np.array([3, -2, 7])).max() < 2e-5


# ------------------------------------------------------------- return maps
def test_thermal_return_map_matches_oracle(gpu):
    # Cases of the reference's thermal parity test (gpu/thermal_reference.mjs:4-8).
    for model in (10, 11):
        for T in (-100, 20, 180, 800):
            for F in ([1, .2, 0, 0, 1, 0, 0, 0, 1], [.9, 0, 0, 0, 1.05, 0, 0, 0, 1]):
                mat = Material(model, 1.0, .4, .4)
                Fo, epo, tau = return_map_batch(mat, np.array(F), .05, T, .01)
                r = o.project_plastic(F, model, 1, 0, .025, 35, .05, .01, T, o.make_part())
                np.testing.assert_allclose(Fo[0], r["F"], atol=4e-6, rtol=5e-5)
                assert abs(epo[0] - (.05 + r["plasticIncrement"])) < 5e-6
                want = o.constitutive_stress(r["F"], model, .4, .4)
                np.testing.assert_allclose(sym6_to_full(tau[0]), want, atol=3e-6, rtol=5e-4)


def test_metal_return_map_matches_oracle(gpu):
    # Cases of the reference's metal parity test (gpu/metal_reference.mjs:4-8).
    F = [1.3, .1, 0, 0, .8, 0, 0, 0, 1 / 1.04]
    for model in (4, 5, 6, 7):
        for T in (20, 650, 1100):
            for dt in (.001, .02):
                mat = Material(model, 1.0, .4, .4, plastic=dict(A=.03, B=.12))
                Fo, epo, tau = return_map_batch(mat, np.array(F), .15, T, dt)
                r = o.project_plastic(F, model, 1, 0, .025, 35, .15, dt, T, o.make_part(plastic=o.plastic_params(
                    A=.03, B=.12)))
                np.testing.assert_allclose(Fo[0], r["F"], atol=4e-6, rtol=4e-5)
                assert abs(epo[0] - (.15 + r["plasticIncrement"])) < 5e-6
                assert abs(np.linalg.det(Fo[0].reshape(3, 3).astype(np.float64)) - np.linalg.det(
                    np.reshape(F, (3, 3)))) < 5e-6
                want = o.constitutive_stress(r["F"], model, .4, .4)
                np.testing.assert_allclose(sym6_to_full(tau[0]), want, atol=3e-6, rtol=5e-4)


def test_fixed_corotated_stress_matches_oracle(gpu):
    rng = np.random.default_rng(9)
    Fs = []
    for _ in range(256):
        A = np.eye(3) + rng.normal(0, .08, (3, 3))
        th = rng.uniform(0, 2 * np.pi)
        R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
        Fs.append((R @ A).reshape(9))
    Fs = np.array(Fs)
    mu, lam = 38461.5, 57692.3
    tau = stress_batch(Material(0, 1e5, mu, lam), Fs)
    for k in range(len(Fs)):
        want = o.stress3(Fs[k], mu, lam)
        scale = np.abs(want).max()
        assert np.abs(sym6_to_full(tau[k]) - want).max() <= 2e-5 * scale + 1e-3


# ---------------------------------------------------------- trajectories
def run_pair(sc, steps, F0=None, per_particle_v=None):
    eng = sc.engine()
    orc = oracle_for(sc, sc.gravity_on_particles)
    if per_particle_v is not None:
        eng.upload(sc.x, per_particle_v, sc.mass, sc.volume, temperature=sc.temperature, material=sc.material)
        orc.vx[:], orc.vy[:], orc.vz[:] = per_particle_v.T
    x0 = orc.positions().copy()
    xg0 = eng.download()["positions"]  # the device's own (FP32-rounded) start
    eng.step(sc.dt, steps)
    pg = sc.gravity if sc.gravity_on_particles else None
    for _ in range(steps):
        orc.step(sc.dt, particle_gravity=pg)
    return eng, orc, x0, xg0


@pytest.mark.parametrize("parity", [True, False])
def test_elastic_cube_200_steps_matches_oracle(gpu, parity):
    # Bound: relative L2 of displacement and velocity <= 1e-3 after 200 steps;
    # max position error <= 1.2e-5 m (0.00075 particle spacings, the
    # reference's atol .003 at spacing 4 scaled to spacing 1/64).
    sc = scenes.elastic_cube(nu=0.25, sticky_floor=False) if parity else scenes.elastic_cube()
    eng, orc, x0, xg0 = run_pair(sc, sc.steps)
    st = eng.download()
    dx = st["positions"] - xg0
    dxo = orc.positions() - x0
    e_dx, e_v = rel_l2(dx, dxo), rel_l2(st["velocities"], orc.velocities())
    e_max = np.abs(st["positions"] - orc.positions()).max()
    print(f"parity={parity} rel L2 dx {e_dx:.2e} v {e_v:.2e} max |x| err {e_max:.2e}")
    assert e_dx < 1e-3 and e_v < 1e-3 and e_max < 1.2e-5
    c = eng.conservation()
    assert c["grid_mass_q"] == c["particle_mass_q"]


def test_thermal_j2_block_matches_oracle(gpu):
    # A block with random per-particle velocities yields within a few steps.
    # Bound: relative L2 of displacement, velocity and plastic strain <= 5e-3
    # after 40 steps of plastic flow.
    rng = np.random.default_rng(1)
    h = 1 / 32
    pts, vp = scenes.geometry((.5, .5, .5), (.2, .2, .2), "box", h / 2, 3, .2)
    n = len(pts)
    mat = Material.from_E_nu(10, 2e11, .3, thermal=dict(scenes.STEEL_THERMAL))
    sc = scenes.Scene("j2_block", (0, 0, 0), (1, 1, 1), h, pts, np.zeros((n, 3)), np.full(n, 2.5e7 * vp),
                      np.full(n, vp), rng.uniform(300, 900, n), np.zeros(n, np.uint8), [mat], 2e-5, 40)
    v = rng.normal(0, .5, (n, 3))
    eng, orc, x0, xg0 = run_pair(sc, sc.steps, per_particle_v=v)
    st = eng.download()
    e_dx = rel_l2(st["positions"] - xg0, orc.positions() - x0)
    e_v = rel_l2(st["velocities"], orc.velocities())
    e_ep = rel_l2(st["plastic"], orc.plastic)
    print(f"j2 rel L2 dx {e_dx:.2e} v {e_v:.2e} ep {e_ep:.2e} ep max {orc.plastic.max():.3e}")
    assert orc.plastic.max() > 1e-4
    assert e_dx < 5e-3 and e_v < 5e-3 and e_ep < 5e-3


def test_momentum_and_mass_conserved_in_free_flight(gpu):
    # No walls in reach, no gravity: sum m v is constant up to FP32 rounding
    # (bound 1e-5 of sum m|v|); grid mass equals particle mass exactly.
    rng = np.random.default_rng(3)
    sc = scenes.elastic_cube(nu=0.3, sticky_floor=False)
    n = sc.n
    x = sc.x - np.array([0, .19 - .5, 0])
    v = rng.normal(0, .05, (n, 3))
    eng = Engine(sc.lower, sc.upper, sc.h, n)
    eng.set_materials(sc.materials)
    eng.upload(x, v, sc.mass, sc.volume)
    P0 = (sc.mass[:, None] * v).sum(0)
    scale = (sc.mass[:, None] * np.abs(v)).sum()
    for _ in range(10):
        eng.step(1e-4, 10)
        c = eng.conservation()
        assert c["grid_mass_q"] == c["particle_mass_q"]
        assert np.abs(c["momentum"] - P0).max() <= 1e-5 * scale
        assert np.abs(c["grid_momentum"] - P0).max() <= 1e-5 * scale
        assert c["nonfinite"] == 0


def test_thermal_boxes_full_size_invariants(gpu):
    # The 128^3 / ~2.1M-particle workload: exact mass every step, finite state,
    # momentum = M g t (no wall contact within 20 steps), deterministic binning.
    sc = scenes.thermal_boxes()
    eng = sc.engine()
    M = sc.mass.sum()
    P0 = (sc.mass[:, None] * sc.v).sum(0)
    g = np.array(sc.gravity)
    _, _, perm_a = eng.debug_bin()
    for k in range(1, 5):
        eng.step(sc.dt, 5)
        c = eng.conservation()
        assert c["grid_mass_q"] == c["particle_mass_q"]
        assert c["nonfinite"] == 0 and c["min_J"] > 0.9
        want = P0 + M * g * sc.dt * 5 * k
        assert np.abs(c["momentum"] - want).max() <= 1e-4 * (np.abs(P0).max() + M * 9.81 * sc.dt * 5 * k)
    eng2 = sc.engine()
    _, _, perm_b = eng2.debug_bin()
    assert np.array_equal(perm_a, perm_b)
