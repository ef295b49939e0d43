"""Pins the oracle's MPM3.step restatement (CPU only).

Reference invariants: exact uniform translation (tests/three.test.mjs:14-21,
tests/discretization.test.mjs:38-46), stable step bound
(tests/discretization.test.mjs:44), plus identities the mathematics fixes:
partition of unity, linear reproduction and the h^2/4 second moment of the
quadratic B-spline, grid mass = particle mass, momentum conservation away
from walls, equivalence of particle-side and grid-side gravity.
"""
import numpy as np
import pytest

from oracle import binding as o

DEFAULT = dict(lower=(-64, -48, -64), upper=(64, 72, 64), h=8)


def ref_box(a, b, v=(0, 0, 0)):
    """Simulation3.box lattice (src/three/simulation.js:95-101): spacing 4,
    mass 1.1 (material 11 density 1100 / 1000), volume 1."""
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    dims = (np.floor((hi - lo) / 4) + 1).astype(int)
    pts = [[lo[0] + 4 * x, lo[1] + 4 * y, lo[2] + 4 * z] for z in range(dims[2]) for y in range(dims[1])
           for x in range(dims[0])]
    return np.array(pts, float)


# tests/three.test.mjs:14-21
def test_uniform_translation_is_exact():
    x = ref_box([-8, -8, -8], [8, 8, 8])
    s = o.Scene(x, (3, -2, 7), 1.1, 1.0, **DEFAULT)
    s.step(.001)
    v = s.velocities()
    assert np.all(np.abs(v[:, 0] - 3) < 1e-8) and np.all(np.abs(v[:, 1] + 2) < 1e-8)
    assert np.all(np.abs(v[:, 2] - 7) < 1e-8)


# tests/discretization.test.mjs:38-46
@pytest.mark.parametrize("cell", [4, 8, 12, 16])
def test_expanded_domain_translation_and_stable_step(cell):
    spacing = cell / 4
    pts = []
    c, size = np.array([76., 40, -76]), np.array([8., 8, 8])
    lo, hi = c - size / 2, c + size / 2
    first = np.ceil(lo / spacing - .5 - 1e-10).astype(int)
    last = np.ceil(hi / spacing - .5 - 1e-10).astype(int) - 1
    for z in range(first[2], last[2] + 1):
        for y in range(first[1], last[1] + 1):
            for xx in range(first[0], last[0] + 1):
                pts.append([(xx + .5) * spacing, (y + .5) * spacing, (z + .5) * spacing])
    x = np.array(pts)
    n = len(x)
    vol = 512 / n / 64
    s = o.Scene(x, (3, -2, 7), 1.1 * vol, vol, lower=(-120,) * 3, upper=(120,) * 3, h=cell, strict_domain=True, ppa=4)
    x0 = s.positions().copy()
    s.step(.001)
    assert np.all(np.abs(s.x - x0[:, 0] - .003) < 1e-9)
    assert np.all(np.abs(s.vz - 7) < 1e-9)
    assert s.stable_step() <= cell / 8 / 240 + 1e-12


def jittered_block(n_side=6, h=1 / 16, seed=3):
    rng = np.random.default_rng(seed)
    sp = h / 2
    g = (np.arange(n_side) + .5) * sp
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1) + 0.4 + rng.uniform(-.2, .2, (n_side ** 3, 3)) * sp
    return x


def si_scene(x, v, **kw):
    n = len(x)
    E, nu, rho = 1e4, .3, 1000
    mu, lam = E / (2 * (1 + nu)), E * nu / ((1 + nu) * (1 - 2 * nu))
    vp = (1 / 32) ** 3
    kw.setdefault("parts", [o.make_part(E=E, G=mu / E, L=lam / E)])
    return o.Scene(x, v, rho * vp, vp, lower=(0, 0, 0), upper=(1, 1, 1), h=1 / 16, vol_scale=1.0, **kw)


def test_grid_mass_and_momentum_identities_after_p2g():
    rng = np.random.default_rng(0)
    x = jittered_block()
    v = rng.normal(0, .3, (len(x), 3))
    s = si_scene(x, v)
    s.F[:] = (np.eye(3) + rng.normal(0, .05, (len(x), 3, 3))).reshape(-1)
    s.C[:] = rng.normal(0, 2, s.C.shape)
    s.p2g(1e-3)
    m = s.mass.sum()
    assert abs(s.gmass.sum() - m) < 1e-12 * m
    # Internal forces and the affine term sum to zero over nodes (linear
    # reproduction of the B-spline), so the grid momentum is sum m v.
    P = (s.mass[:, None] * v).sum(0)
    assert np.allclose(s.gv.reshape(-1, 3).sum(0), P, rtol=1e-11, atol=1e-15)


def test_bspline_moments_single_particle():
    # One particle at random offsets: sum w = 1, sum w (x_i - x) = 0 and
    # sum w |x_i - x|^2 = 3 h^2 / 4 (the D = h^2/4 of APIC per axis).
    rng = np.random.default_rng(5)
    h = 1 / 16
    for _ in range(20):
        x = rng.uniform(.3, .7, (1, 3))
        s = si_scene(x, np.zeros((1, 3)))
        s.p2g(1e-3)
        N = s.N
        idx = np.nonzero(s.gmass)[0]
        w = s.gmass[idx] / s.mass[0]
        node = np.stack([idx % N, (idx // N) % N, idx // (N * N)], 1) * h + s.origin
        d = node - x[0]
        assert abs(w.sum() - 1) < 1e-14
        assert np.allclose((w[:, None] * d).sum(0), 0, atol=1e-15)
        assert abs((w * (d ** 2).sum(1)).sum() - 3 * h * h / 4) < 1e-15


def test_momentum_conserved_away_from_walls_and_rigid_rotation_is_stress_free():
    rng = np.random.default_rng(11)
    x = jittered_block(8)
    v = rng.normal(0, .2, (len(x), 3))
    s = si_scene(x, v)
    P0 = (s.mass[:, None] * s.velocities()).sum(0)
    for _ in range(20):
        s.step(2e-4)
    P1 = (s.mass[:, None] * s.velocities()).sum(0)
    assert np.allclose(P1, P0, rtol=1e-10, atol=1e-14)
    # A rigidly rotated block (F = R, zero velocity) has zero stress: no motion.
    c, sn = np.cos(.6), np.sin(.6)
    R = np.array([[c, -sn, 0], [sn, c, 0], [0, 0, 1]])
    s2 = si_scene(x, np.zeros_like(x))
    s2.F[:] = np.tile(R.reshape(9), len(x))
    s2.step(1e-3)
    assert np.abs(s2.velocities()).max() < 1e-10


def test_grid_gravity_equals_particle_gravity():
    x = jittered_block(6)
    g = (0, -9.81, 0)
    a = si_scene(x, np.zeros_like(x))
    b = si_scene(x, np.zeros_like(x), grid_gravity=g)
    for _ in range(10):
        a.step(1e-3, particle_gravity=g)
        b.step(1e-3)
    assert np.allclose(a.positions(), b.positions(), atol=1e-12)
    assert np.allclose(a.velocities(), b.velocities(), atol=1e-10)


def test_sticky_floor_stops_all_components():
    x = jittered_block(6)
    x[:, 1] -= 0.4 - 0.02  # base near the floor
    v = np.tile([.5, -.5, .3], (len(x), 1))
    s = si_scene(x, v, wall_mode=(0, 0, 1, 0, 0, 0))
    s.step(1e-4)
    N, H = s.N, s.H
    iy = (np.arange(N ** 3) // N) % N
    near = (s.origin[1] + iy * H < 0 + H) & (s.gmass > 1e-12)
    assert near.any()
    assert np.all(s.gv.reshape(-1, 3)[near] == 0)


def test_thermal_scene_yields_and_conserves_mass():
    rng = np.random.default_rng(2)
    x = jittered_block(6)
    v = rng.normal(0, .5, (len(x), 3))
    E = 1e6
    part = o.make_part(E=E, G=.3846, L=.5769,
                       thermal=o.thermal_params(A=1.25e-3, H=0.0, Tref=300, Thot=1200, hotRatio=.1,
                                                kelvin_offset=0.0))
    s = si_scene(x, v, parts=[part], model=10, temperature=rng.uniform(300, 900, len(x)))
    for _ in range(10):
        s.step(1e-4)
    assert s.plastic.max() > 0
    assert np.all(s.cacheValid == 1)
    assert abs(s.gmass.sum() - s.mass.sum()) < 1e-12 * s.mass.sum()
