"""Independent pins of the oracle's 3D step terms (CPU only).

Each test fixes a term of the MPM3.step restatement by a property the
mathematics or the physics fixes, chosen so that a plausible slip (a sign, a
missing factor, a transposed index) fails it:

- the P2G stress-force term A = m C - dt V 4/h^2 tau (src/three/mpm.js:71): the
  first moment of one particle's grid momentum is -dt V tau, because the
  quadratic B-spline's second moment is h^2/4 (any other factor or sign fails);
- the G2P affine gather C = sum w v (x_i - x_p) 4/h^2 and F <- (I + dt C) F
  (src/three/mpm.js:113-121): a grid velocity linear in x returns exactly its
  gradient (a transposed C or a wrong factor fails);
- the APIC spin term of the angular momentum (src/simulation/mechanical-
  diagnostics.js:19-31) together with the P2G/G2P transfers: a spinning,
  stretched block in free space keeps its total angular momentum (APIC
  conserves it, Jiang et al. 2015) while its orbital part changes by orders
  of magnitude more than the tolerance;
- orc_mechanics: the spin of a particle in rigid rotation is m 2D omega, the
  moment of inertia of the B-spline's smeared mass about an axis;
- the 3D stableStep (src/three/mpm.js:43-55) on states where it has a closed
  form: the P-wave speed c = sqrt((lambda + 2 mu) / rho) from the moduli and
  density (19200 stiffness in the reference's units), the row/column factor of
  a stretch, Neo-Hookean compression, snow hardening, the |C| term and the cap.
"""
import numpy as np
import pytest

from oracle import binding as o

SI = dict(lower=(0, 0, 0), upper=(1, 1, 1), h=1 / 32)


def si_part(E=1e5, nu=0.3):
    mu = E / (2 * (1 + nu))
    lam = E * nu / ((1 + nu) * (1 - 2 * nu))
    return o.make_part(E=E, G=mu / E, L=lam / E), mu, lam


def node_positions(s):
    N = s.N
    idx = np.arange(N ** 3)
    i, j, k = idx % N, (idx // N) % N, idx // (N * N)
    return np.stack([s.origin[0] + i * s.H, s.origin[1] + j * s.H, s.origin[2] + k * s.H], 1)


@pytest.mark.parametrize("stretch", [1.04, 0.97])
@pytest.mark.parametrize("pos", [(0.5, 0.5, 0.5), (0.4321, 0.5117, 0.6093)])
def test_p2g_force_moment_is_minus_dt_V_tau(stretch, pos):
    part, mu, lam = si_part()
    vol, rho, dt = 1e-6, 1000.0, 1e-4
    s = o.Scene(np.array([pos]), (0, 0, 0), rho * vol, vol, parts=[part], vol_scale=1.0, **SI)
    s.F[:] = np.diag([stretch, 1.0, 1.0]).reshape(9)
    s.p2g(dt)
    # closed-form fixed-corotated stress of F = diag(s, 1, 1): R = I, J = s
    tau = np.diag([2 * mu * stretch * (stretch - 1), 0, 0]) + lam * (stretch - 1) * stretch * np.eye(3)
    p = s.gv.reshape(-1, 3)
    d = node_positions(s) - np.array(pos)
    moment = p.T @ d  # sum_n p_n (x_n - x_p)^T
    assert np.allclose(moment, -dt * vol * tau, rtol=1e-10, atol=1e-12 * np.abs(dt * vol * tau).max())
    # and no net momentum from an internal force
    assert np.allclose(p.sum(0), 0, atol=1e-15 * np.abs(dt * vol * tau).max() / s.H)


def test_compressed_block_expands():
    part, mu, lam = si_part()
    h = 1 / 32
    pts = np.array([[x, y, z] for x in np.arange(0.4, 0.6, h / 2) for y in np.arange(0.4, 0.6, h / 2)
                    for z in np.arange(0.4, 0.6, h / 2)])
    vol = (h / 2) ** 3
    s = o.Scene(pts, (0, 0, 0), 1000 * vol, vol, parts=[part], vol_scale=1.0, **SI)
    s.F[:] = np.tile((0.97 * np.eye(3)).reshape(9), len(pts))
    w0 = np.ptp(s.positions(), 0)
    for _ in range(10):
        s.step(1e-4)
    assert np.all(np.ptp(s.positions(), 0) > w0)
    # outward velocity on the +x face, inward-free: the stress pushes out
    face = s.x > s.x.max() - h / 4
    assert s.vx[face].mean() > 0


@pytest.mark.parametrize("pos", [(0.5, 0.5, 0.5), (0.4617, 0.5383, 0.5091)])
def test_g2p_returns_the_gradient_of_a_linear_grid_velocity(pos):
    part, _, _ = si_part()
    vol, dt = 1e-6, 1e-4
    s = o.Scene(np.array([pos]), (0, 0, 0), 1000 * vol, vol, parts=[part], vol_scale=1.0, **SI)
    F0 = np.array([[1.01, 0.02, 0.0], [-0.01, 0.99, 0.03], [0.0, 0.01, 1.0]])
    s.F[:] = F0.reshape(9)
    s.step_to_grid(dt)
    G = np.array([[0.3, -1.7, 0.4], [2.1, -0.2, 0.9], [-0.6, 1.3, 0.5]])
    v0 = np.array([0.25, -0.5, 0.75])
    xn = node_positions(s)
    s.gv[:] = (v0 + xn @ G.T).reshape(-1)
    s.step_from_grid(dt)
    C = s.C.reshape(3, 3)
    assert np.allclose(C, G, rtol=0, atol=1e-11)
    vp = v0 + G @ np.array(pos)
    assert np.allclose(s.velocities()[0], vp, rtol=0, atol=1e-12)
    assert np.allclose(s.F.reshape(3, 3), (np.eye(3) + dt * G) @ F0, rtol=0, atol=1e-13)
    assert np.allclose(s.positions()[0], np.array(pos) + dt * vp, rtol=0, atol=1e-15)


def test_apic_conserves_angular_momentum_with_spin_term():
    part, _, _ = si_part(E=2e4)
    h = 1 / 32
    g = np.arange(-0.1, 0.1 + 1e-9, h / 2)
    pts = np.array([[0.5 + x, 0.5 + y, 0.5 + z] for x in g for y in g for z in g if abs(z) <= 0.05])
    n = len(pts)
    vol = (h / 2) ** 3
    omega = np.array([0.3, -0.2, 2.0])
    c = pts.mean(0)
    v = np.cross(omega, pts - c) + np.array([0.05, 0.0, -0.03])
    s = o.Scene(pts, v, 1000 * vol, vol, parts=[part], vol_scale=1.0, mass_eps=0.0, **SI)
    s.F[:] = np.tile(np.diag([1.03, 0.98, 1.0]).reshape(9), n)  # stretched: the stress acts
    L0 = s.mechanics()["angular"]
    orbital0 = np.cross(s.positions(), s.velocities() * s.mass[:, None]).sum(0)
    for _ in range(50):
        s.step(2e-4)
    m = s.mechanics()
    orbital = np.cross(s.positions(), s.velocities() * s.mass[:, None]).sum(0)
    scale = np.linalg.norm(L0)
    assert np.linalg.norm(m["angular"] - L0) < 1e-10 * scale
    # the spin term carries a part of it that is far above the tolerance
    assert np.linalg.norm(orbital - m["angular"]) > 1e-4 * scale
    assert np.linalg.norm(orbital - orbital0) > 1e-6 * scale
    # linear momentum is conserved too
    assert np.allclose(m["momentum"], s.mass.sum() * np.array([0.05, 0.0, -0.03]), rtol=1e-10, atol=1e-12)


def test_mechanics_spin_of_a_rigid_rotation_is_its_smeared_inertia():
    h = 8.0
    s = o.Scene(np.array([[4.0, 8.0, 12.0]]), (0, 0, 0), 1.1, 1.0, **dict(lower=(-64, -48, -64), upper=(64, 72, 64),
                                                                         h=h))
    w = 1.7
    # v = omega x d with omega = w e_z: dv_x/dy = -w, dv_y/dx = w
    s.C[:] = np.array([[0, -w, 0], [w, 0, 0], [0, 0, 0]]).reshape(9)
    m = s.mechanics()
    D = h * h / 4
    # the B-spline smears the mass with second moment D per axis: I_z = m 2D
    assert np.allclose(m["angular"], [0, 0, 1.1 * 2 * D * w], rtol=1e-14)
    assert m["affine"] == pytest.approx(0.5 * 1.1 * D * 2 * w * w, rel=1e-14)
    assert m["kinetic"] == 0 and np.all(m["momentum"] == 0)
    s2 = o.Scene(np.array([[4.0, 8.0, 12.0], [-3.0, 1.0, 2.0]]), [(1, 2, 3), (-2, 0.5, 1)], [1.1, 2.2], 1.0,
                 lower=(-64, -48, -64), upper=(64, 72, 64), h=h)
    m2 = s2.mechanics()
    P = np.array([1.1 * np.array([1, 2, 3]), 2.2 * np.array([-2, 0.5, 1])])
    X = np.array([[4.0, 8.0, 12.0], [-3.0, 1.0, 2.0]])
    assert np.allclose(m2["momentum"], P.sum(0), rtol=1e-15)
    assert np.allclose(m2["angular"], np.cross(X, P).sum(0), rtol=1e-14)
    assert m2["kinetic"] == pytest.approx(0.5 * (P ** 2 / np.array([[1.1], [2.2]])).sum(), rel=1e-14)
    assert m2["mass"] == pytest.approx(3.3)


REF = dict(lower=(-64, -48, -64), upper=(64, 72, 64), h=8.0)


def ref_scene(model=0, v=(0, 0, 0), **kw):
    s = o.Scene(np.array([[4.0, 8.0, 12.0]]), np.array([v], float), 1.1, 1.0, model=model, **REF)
    for k, val in kw.items():
        getattr(s, k)[:] = val
    return s


def c_ref(stiffness=1.0):
    # P-wave speed sqrt((lambda + 2 mu) / rho) in the reference's units:
    # E = m/(64 V) 16000 stiffness, mu = lambda = 0.4 E, rho = m/(64 V).
    rho = 1.1 / 64
    E = rho * 16000 * stiffness
    return np.sqrt((0.4 * E + 2 * 0.4 * E) / rho)


BIG = 1e9


def test_stable_step_resting_moving_stretched():
    h = REF["h"]
    c = c_ref()
    assert ref_scene().stable_step(BIG) == pytest.approx(1.4 * h / 8 / c, rel=1e-14)
    s = ref_scene(v=(30.0, -40.0, 0.0))
    assert s.stable_step(BIG) == pytest.approx(1.4 * h / 8 / (50 + c), rel=1e-14)
    st = 1.5
    s = ref_scene(F=np.diag([st, 1, 1]).reshape(9))
    assert s.stable_step(BIG) == pytest.approx(1.4 * h / 8 / (c * st * st), rel=1e-14)
    # a shear: rows (1 + g, 1, 1), columns (1, 1 + g, 1)
    g = 0.4
    s = ref_scene(F=np.array([[1, g, 0], [0, 1, 0], [0, 0, 1]], float).reshape(9))
    assert s.stable_step(BIG) == pytest.approx(1.4 * h / 8 / (c * (1 + g) ** 2), rel=1e-14)
    # the default cap h/8/240 binds at rest
    assert ref_scene().stable_step() == pytest.approx(h / 8 / 240, rel=1e-14)
    # a custom Courant number scales the wave term
    assert ref_scene().stable_step(BIG, 0.3) == pytest.approx(0.3 * h / c, rel=1e-14)


def test_stable_step_gradient_compression_and_hardening():
    h = REF["h"]
    c = c_ref()
    Cm = np.array([[30, -20, 5], [0, 40, 0], [-10, 0, 25]], float)
    s = ref_scene(C=Cm.reshape(9))
    assert s.stable_step(BIG) == pytest.approx(min(1.4 * h / 8 / c, .08 / np.abs(Cm).sum()), rel=1e-14)
    # Neo-Hookean in compression: 1 / max(.05, J)
    J = 0.5
    s = ref_scene(model=1, F=np.diag([J, 1, 1]).reshape(9))
    assert s.stable_step(BIG) == pytest.approx(1.4 * h / 8 / (c / J), rel=1e-14)
    s = ref_scene(model=1, F=np.diag([0.01, 1, 1]).reshape(9))
    assert s.stable_step(BIG) == pytest.approx(1.4 * h / 8 / (c / .05), rel=1e-14)
    # snow hardening clamp(exp(10 (1 - Jp)), .25, 4) scales the modulus
    s = ref_scene(model=2, Jp=[0.95])
    assert s.stable_step(BIG) == pytest.approx(1.4 * h / 8 / (c * np.sqrt(np.exp(0.5))), rel=1e-13)
    s = ref_scene(model=2, Jp=[0.5])
    assert s.stable_step(BIG) == pytest.approx(1.4 * h / 8 / (c * 2), rel=1e-14)
    # the stiffness setting scales c^2
    s = o.Scene(np.array([[4.0, 8.0, 12.0]]), (0, 0, 0), 1.1, 1.0, stiffness=2.5, **REF)
    assert s.stable_step(BIG) == pytest.approx(1.4 * h / 8 / c_ref(2.5), rel=1e-14)


def test_stable_step_takes_the_slowest_particle():
    h = REF["h"]
    c = c_ref()
    s = o.Scene(np.array([[4.0, 8.0, 12.0], [-8.0, 0.0, 4.0], [0.0, 16.0, -4.0]]), [(0, 0, 0), (0, 0, 100), (5, 0, 0)],
                1.1, 1.0, **REF)
    assert s.stable_step(BIG) == pytest.approx(1.4 * h / 8 / (100 + c), rel=1e-14)


def test_stable_step_si_part_uses_its_moduli():
    part, mu, lam = si_part(E=1e5, nu=0.3)
    vol, rho = 1e-6, 1000.0
    s = o.Scene(np.array([[0.5, 0.5, 0.5]]), (0, 0, 0), rho * vol, vol, parts=[part], vol_scale=1.0, **SI)
    c = np.sqrt((lam + 2 * mu) / rho)
    assert s.stable_step(BIG) == pytest.approx(1.4 * SI["h"] / 8 / c, rel=1e-12)
