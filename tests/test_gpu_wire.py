"""GPU frame snapshot (mpm_snapshot, the service's pack kernel,
gpu/solver.py:986-1032) against the picks of the oracle's hash and the
downloaded device state, through the C ABI.

Bars: the preview rows equal the downloaded FP32 state of the picked
particles bit for bit (x, y, z, T, pinned, plastic, code) and the stress
column equals the device diagnostic (mpm_refresh_diagnostics) bit for bit;
counts are exact; the FP64 mass total within 1e-12 relative; min det F
within 1e-6 of the FP64 determinant of the FP32 F."""
import numpy as np
import pytest

from mpm_amd import Engine, Material, scenes, wire
from oracle import binding as o

pytestmark = pytest.mark.gpu

LOWER, UPPER, H = (-64, -48, -64), (64, 72, 64), 8.0
E_REF = 1.1 / 64 * 16000


def picks(n, stride):
    rows = (n + stride - 1) // stride
    ids = np.array([r * stride + o.lib().orc_preview_pick(r, stride, n) for r in range(rows)], np.int64)
    assert (ids < n).all() and (ids // stride == np.arange(rows)).all()
    return ids


def check_frame(eng, st, data, codes):
    s = eng.download()
    n = eng.count
    assert st["count"] == n and data.shape == (st["render_count"], 8)
    ids = picks(n, st["sample_stride"]) if n <= 200_000 else None
    if ids is None:  # vectorised pick for large n (the same hash in numpy uint32)
        rows = np.arange(st["render_count"], dtype=np.uint32)
        with np.errstate(over="ignore"):
            h = rows * np.uint32(747796405) + np.uint32(2891336453)
            h = ((h >> ((h >> np.uint32(28)) + np.uint32(4))) ^ h) * np.uint32(277803737)
            h = (h >> np.uint32(22)) ^ h
        span = np.minimum(st["sample_stride"], n - rows.astype(np.int64) * st["sample_stride"])
        ids = rows.astype(np.int64) * st["sample_stride"] + (h.astype(np.int64) % span)
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32)  # noqa: E731
    np.testing.assert_array_equal(data[:, 0], f32(s["x"][ids]))
    np.testing.assert_array_equal(data[:, 1], f32(s["y"][ids]))
    np.testing.assert_array_equal(data[:, 2], f32(s["z"][ids]))
    np.testing.assert_array_equal(data[:, 4], f32(s["temperature"][ids]))
    np.testing.assert_array_equal(data[:, 5], s["pinned"][ids].astype(np.float32))
    np.testing.assert_array_equal(data[:, 7], f32(s["plastic"][ids]))
    diag = eng.refresh_diagnostics()
    np.testing.assert_array_equal(data[:, 6], f32(diag[ids]))
    return s, ids, diag


def test_snapshot_of_a_mixed_scene_matches_the_state():
    # Three parts: fixed-corotated, snow and foam (codes 11, 11, 5), some
    # pinned particles, after 20 steps of two colliding pairs of blocks.
    rng = np.random.default_rng(5)
    blocks = []
    for k, (x0, z0) in enumerate(((-16, -12), (16, -12), (-16, 12), (16, 12))):
        a = np.arange(5) * 3.0 - 6
        blocks.append(np.array([[x0 + x, y + 8, z0 + z] for z in a for y in a for x in a]))
    pts = np.concatenate(blocks)
    n = len(pts)
    mat = np.repeat([0, 1, 2, 0], 125).astype(np.uint8)
    v = np.repeat([[40, 0, 0], [-40, 0, 0], [40, 0, 0], [-40, 0, 0]], 125, axis=0) + rng.normal(0, 1, (n, 3))
    pinned = (rng.uniform(size=n) < .02).astype(np.uint8)
    eng = Engine(LOWER, UPPER, H, n, strict_domain=True)
    eng.set_materials([Material(0, E_REF, .4 * E_REF, .4 * E_REF), Material(2, E_REF, .4 * E_REF, .4 * E_REF),
                       Material(9, E_REF, .4 * E_REF, .4 * E_REF)])
    eng.upload(pts, v, 1.1, 64.0, material=mat, pinned=pinned, temperature=rng.uniform(0, 100, n))
    eng.step(1 / 240, 20)
    codes = [11, 11, 5]
    st, data = eng.snapshot(100_000, codes)
    s, ids, diag = check_frame(eng, st, data, codes)
    want_code = np.where(mat == 0, 11, np.where(mat == 1, 13, 5)).astype(np.float32)
    np.testing.assert_array_equal(data[:, 3], want_code[ids])
    assert st["models"][0] == 250 and st["models"][2] == 125 and st["models"][9] == 125
    assert st["solids"] == n and st["liquids"] == 0 and st["nonfinite"] == 0
    assert st["mass"] == pytest.approx(np.asarray(s["mass"], np.float32).astype(np.float64).sum(), rel=1e-12)
    T = np.asarray(s["temperature"], np.float32)
    assert (st["min_temperature"], st["max_temperature"]) == (T.min(), T.max())
    J = np.linalg.det(np.asarray(s["F"], np.float32).astype(np.float64).reshape(-1, 3, 3))
    assert st["min_J"] == pytest.approx(min(1.0, J.min()), rel=1e-6)
    assert st["max_stress"] == np.float32(diag.max()) and st["max_plastic"] == np.float32(s["plastic"]).max()
    # a smaller preview limit samples one particle per block
    st2, data2 = eng.snapshot(100, codes)
    assert st2["sample_stride"] == 5 and st2["render_count"] == 100
    check_frame(eng, st2, data2, codes)
    meta, rows = wire.frame(eng, 100, codes, time=20 / 240)
    back, got = wire.decode(wire.packet(meta, rows))
    assert back["renderCount"] == 100 and back["models"][9] == 125 and np.array_equal(got, data2)


def test_snapshot_of_the_full_size_workload():
    # 2,118,656 particles previewed in at most 100,000 rows.
    sc = scenes.thermal_boxes()
    eng = sc.engine(sort_every=16)
    eng.step(sc.dt, 20)
    st, data = eng.snapshot()
    assert st["sample_stride"] == 22 and st["render_count"] == 96_303
    check_frame(eng, st, data, None)
    assert st["models"][10] == sc.n and (data[:, 3] == 21).all()


def test_snapshot_marks_failed_particles_as_sand():
    a = np.arange(5) * 4.0 - 8
    pts = np.array([[x, y + 8, z] for z in a for y in a for x in a])
    v = np.stack([pts[:, 0] * 6.0, np.zeros(125), np.zeros(125)], 1)
    eng = Engine(LOWER, UPPER, H, 125, strict_domain=True)
    eng.set_materials([Material(0, E_REF, .4 * E_REF, .4 * E_REF,
                                fracture=dict(enabled=True, onset=.02, failure=.06))])
    eng.upload(pts, v, 1.1, 64.0)
    eng.step(1 / 240, 12)
    st, data = eng.snapshot()
    assert st["models"][3] == 125 and st["models"][0] == 0
    assert (data[:, 3] == 14).all()


def test_empty_engine_and_bad_arguments():
    from mpm_amd import lib as L
    eng = Engine(LOWER, UPPER, H, 16, strict_domain=True)
    st, data = eng.snapshot()
    assert st["count"] == 0 and data.shape == (0, 8) and st["min_J"] == 1 and st["min_temperature"] == 0
    eng.upload(np.zeros((4, 3)), 0, 1.1, 64.0)
    with pytest.raises(L.MpmError):
        eng.snapshot(0)
    eng.set_materials([Material(0, E_REF, .4 * E_REF, .4 * E_REF)] * 2)
    with pytest.raises(L.MpmError):
        eng.snapshot(10, [11])  # fewer codes than materials
