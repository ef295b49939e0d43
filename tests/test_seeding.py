"""Seeding: the product's numpy restatement (mpm_amd.scenes) against the
oracle's C restatement, bit for bit, and the reference's sampling invariants
(tests/sampling.test.mjs:9-35, tests/discretization.test.mjs:7-23). CPU only.
"""
import numpy as np
import pytest

from mpm_amd import scenes
from oracle import binding as o


def test_hash_matches_oracle_bitwise():
    rng = np.random.default_rng(0)
    seeds = rng.integers(0, 2 ** 31 - 1, 50)
    for seed in seeds:
        idx = rng.integers(0, 5_000_000, 20)
        for attempt in range(8):
            for axis in range(3):
                got = scenes.random(int(seed), idx, attempt, axis)
                want = np.array([o.random(int(seed), int(i), attempt, axis) for i in idx])
                assert np.array_equal(got, want)
    v = scenes.random(7, np.arange(100000), 0, 0)
    assert v.min() >= -1 and v.max() < 1 and abs(v.mean()) < 0.01


def ref_box_points(a, b, shape):
    """Simulation3.box lattice with spacing 4 (src/three/simulation.js:95-101)."""
    lo, hi = np.minimum(a, b).astype(float), np.maximum(a, b).astype(float)
    dims = (np.floor((hi - lo) / 4) + 1).astype(int)
    pts = []
    for z in range(dims[2]):
        for y in range(dims[1]):
            for x in range(dims[0]):
                p = lo + 4 * np.array([x, y, z])
                q = (p - (lo + hi) / 2) / ((hi - lo) / 2)
                r = (q ** 2).sum() if shape == "ellipsoid" else q[0] ** 2 + q[2] ** 2
                if shape == "box" or r <= 1 + 1e-7:
                    pts.append(p)
    return np.array(pts), lo, hi


@pytest.mark.parametrize("shape", ["box", "ellipsoid", "cylinder"])
@pytest.mark.parametrize("fixed", [False, True])
def test_jitter_matches_oracle_and_reference_invariants(shape, fixed):
    pts, lo, hi = ref_box_points([-8, 0, -8], [8, 16, 8], shape)
    pinned = (pts[:, 0] == lo[0]) & fixed
    got = scenes.jitter(pts, lo, hi, shape, 123, .35, 4.0, pinned)
    want = np.array([o.jitter_point(p, k, lo, hi, scenes.SHAPES[shape], pinned[k], 1, .35, 123, 4.0)
                     for k, p in enumerate(pts)])
    assert np.array_equal(got, want)
    d = np.linalg.norm(got - pts, axis=1)
    assert d.max() <= 1.4 + 1e-10
    assert (d > 1e-6).sum() > .8 * len(pts)
    assert np.all(got >= lo - 1e-12) and np.all(got <= hi + 1e-12)
    if fixed:
        assert np.all(got[pinned, 0] == lo[0])
    q = (got - (lo + hi) / 2) / ((hi - lo) / 2)
    if shape == "ellipsoid":
        assert np.all((q ** 2).sum(1) <= 1 + 1e-10)
    if shape == "cylinder":
        assert np.all(q[:, 0] ** 2 + q[:, 2] ** 2 <= 1 + 1e-10)
    dist = np.linalg.norm(got[:, None] - got[None], axis=2)
    np.fill_diagonal(dist, np.inf)
    assert dist.min() >= 1.2 - 1e-10
    # repeatable and seed-dependent
    assert np.array_equal(got, scenes.jitter(pts, lo, hi, shape, 123, .35, 4.0, pinned))
    assert not np.array_equal(got, scenes.jitter(pts, lo, hi, shape, 124, .35, 4.0, pinned))
    assert np.array_equal(scenes.jitter(pts, lo, hi, shape, 123, 0.0, 4.0), pts)


@pytest.mark.parametrize("n", [2, 3, 4])
def test_full_cells_hold_n_cubed_particles_and_mass(n):
    spacing = 8 / n
    pts = scenes.lattice((0, 0, 0), (16, 16, 16), "box", spacing)
    assert len(pts) == 8 * n ** 3
    want = o.geometry_points((0, 0, 0), (16, 16, 16), 0, spacing)
    assert np.array_equal(pts, want)
    vol = 4096 / len(pts) / 64
    assert abs(1.1 * vol * len(pts) - 4096 / 64 * 1.1) < 1e-9


@pytest.mark.parametrize("shape", ["box", "ellipsoid", "cylinder"])
def test_lattice_matches_oracle(shape):
    c, s = (21.125, 5.5, -17.25), (20.5, 24, 16)
    got = scenes.lattice(c, s, shape, 4.0)
    want = o.geometry_points(c, s, scenes.SHAPES[shape], 4.0)
    assert np.array_equal(got, want)


def test_domain_matches_oracle_and_reference_rules():
    d = o.domain([-7, -7, -7], [9, 9, 9], 8, 2)
    assert list(d["cells"]) == [3, 3, 3]
    for cell in (4, 8, 12, 16):
        od = o.domain([-120] * 3, [120] * 3, cell, 4)
        origin, nodes = scenes.domain([-120] * 3, [120] * 3, cell)
        assert np.array_equal(origin, od["origin"]) and np.array_equal(nodes, od["nodes"])
        for a in range(3):
            for x in (-120, 120):
                base = np.floor((x - od["origin"][a]) / cell - .5)
                assert base >= 0 and base + 2 < od["nodes"][a] and od["nodes"][a] <= 67
    for bad in (dict(h=0), dict(h=np.inf), dict(h=16, n=2), dict(lower=[-2401, 0, 0]), dict(h=4, upper=[240] * 3)):
        kw = dict(lower=[-64, -48, -64], upper=[64, 72, 64], h=8, n=2)
        kw.update(bad)
        with pytest.raises(ValueError):
            o.domain(kw["lower"], kw["upper"], kw["h"], kw["n"], strict=True)


def test_workload_scenes():
    c = scenes.elastic_cube()
    assert 6859 <= c.n <= 8000 and np.allclose(c.mass.sum(), 1000 * .027)
    b = scenes.thermal_boxes(cells=32, ppa=1)
    assert len(b.info["boxes"]) == 6 and b.temperature.min() >= 300 and b.temperature.max() < 900
