"""Pins the oracle's frame snapshot (the service's pack kernel,
gpu/solver.py:986-1032) and the packet layout (gpu/server.py:17-20) that
the browser worker decodes (src/three/gpu-worker.js:36-42). CPU only."""
import json
import struct

import numpy as np
import pytest

from oracle import binding as o

DOM = dict(lower=(-64, -48, -64), upper=(64, 72, 64), h=8)


def pcg_output(row):
    """The PCG RXS-M-XS 32-bit output permutation (O'Neill 2014) applied to
    one LCG step of the block index, in numpy uint32 arithmetic."""
    with np.errstate(over="ignore"):
        h = np.uint32(row) * np.uint32(747796405) + np.uint32(2891336453)
        h = ((h >> ((h >> np.uint32(28)) + np.uint32(4))) ^ h) * np.uint32(277803737)
        return (h >> np.uint32(22)) ^ h


def test_pick_is_the_pcg_output_modulo_the_block_span():
    for n, stride in ((1_000_000, 10), (100_001, 2), (37, 5), (7, 1)):
        for row in list(range(0, (n + stride - 1) // stride, max(1, n // stride // 50))) + [(n - 1) // stride]:
            span = min(stride, n - row * stride)
            assert o.lib().orc_preview_pick(row, stride, n) == int(pcg_output(row)) % span


def test_every_block_has_exactly_one_sample():
    for n, stride in ((1000, 7), (1001, 10), (64, 64), (65, 64), (5, 1)):
        rows = (n + stride - 1) // stride
        hits = np.zeros(rows, int)
        for i in range(n):
            if i % stride == o.lib().orc_preview_pick(i // stride, stride, n):
                hits[i // stride] += 1
        assert (hits == 1).all(), (n, stride)


def lattice(count_per_axis, model):
    a = np.arange(count_per_axis) * 4.0 - 4
    pts = np.array([[x, y + 4, z] for z in a for y in a for x in a])
    return pts, np.full(len(pts), model)


@pytest.mark.parametrize("n,limit,stride,rows", [(54, 100_000, 1, 54), (1_000_000, 100_000, 10, 100_000),
                                                  (100_001, 100_000, 2, 50_001), (7, 3, 3, 3)])
def test_stride_and_rows(n, limit, stride, rows):
    # gpu/test_planet.py:82-84: a million particles preview at stride 10 in 100,000 rows
    pts = np.zeros((n, 3))
    sc = o.Scene(pts, 0, 1.0, 1.0, **DOM)
    st = o.OrcSnapshot()
    o.lib().orc_snapshot(o.C.byref(sc.s), limit, None, None, o.C.byref(st))
    assert (st.sample_stride, st.render_count, st.count) == (stride, rows, n)


def test_codes_counts_and_totals():
    # gpu/test_pressure_plasticity.py:28-37: a model-9 block previews with code 11 + 9 = 20
    # and the frame counts the particles of each model.
    p8, m8 = lattice(3, 8)
    p9, m9 = lattice(3, 9)
    pts = np.concatenate([p9, p8 + [20, 0, 0]])
    models = np.concatenate([m9, m8[1:], [0]])
    temps = np.linspace(10, 90, len(pts))
    sc = o.Scene(pts, (1, 2, 3), 1.1, 1.0, model=models, temperature=temps, part=[0] * 27 + [1] * 27, **DOM)
    sc.plastic[:] = np.linspace(0, .2, len(pts))
    st, data = sc.snapshot(codes=[11, 7])  # part 1 is not the MPM solid material: its code is its own
    assert st["models"][8:10] == [26, 27] and st["models"][0] == 1  # solids count by model whatever the code
    assert data[0, 3] == 20 and (data[27:, 3] == 7).all()
    np.testing.assert_array_equal(data[:, :3], pts.astype(np.float32))
    np.testing.assert_array_equal(data[:, 4], temps.astype(np.float32))
    np.testing.assert_array_equal(data[:, 7], sc.plastic.astype(np.float32))
    assert st["mass"] == pytest.approx(1.1 * 54) and st["solids"] == 54 and st["liquids"] == 0
    assert (st["min_temperature"], st["max_temperature"]) == (10, 90)
    assert st["min_J"] == 1 and st["max_plastic"] == pytest.approx(.2) and st["max_stress"] == 0
    st, data = sc.snapshot()
    np.testing.assert_array_equal(data[:, 3], 11 + models)


def test_stress_column_is_the_diagnostic():
    pts, models = lattice(2, 0)
    sc = o.Scene(pts, 0, 1.1, 1.0, **DOM)
    sc.F[:9] = [.95, .02, 0, 0, .97, 0, 0, 0, 1]
    st, data = sc.snapshot()
    sc.refresh_diagnostics()
    np.testing.assert_array_equal(data[:, 6], sc.stress.astype(np.float32))
    assert data[0, 6] > 0 and st["max_stress"] == sc.stress.max()
    assert st["min_J"] == pytest.approx(.95 * .97)  # from 1: a stretched particle alone would leave 1


def test_nonfinite_particles_are_counted():
    # gpu/test_planet.py:88-90: a NaN position makes the frame invalid
    pts, _ = lattice(2, 0)
    sc = o.Scene(pts, 0, 1.0, 1.0, **DOM)
    sc.x[3] = np.nan
    sc.vy[5] = 2e10
    st, _ = sc.snapshot()
    assert st["nonfinite"] == 2
