"""GPU tests of the batched multi-scene engine (mpm_batch_*), through the C ABI.

Bars:
  * every scene of a batch equals the same scene stepped alone by the
    single-scene engine: relative L2 of displacement and velocity <= 1e-5
    (both run the same per-particle arithmetic on the same 256-particle
    chunks; only the order of the FP32 node atomics differs);
  * graph replay of whole re-bin periods equals eager launches, also for an
    odd number of replayed periods (buffer swap tracking): same bar;
  * every particles-per-workgroup choice (64, 128, 256, equal or mixed
    across size buckets) gives the same trajectories: same bar;
  * fixed-point grid mass equals the particle mass of every scene exactly;
  * the autotuner returns a valid configuration and the batch keeps stepping;
  * the batch's binning (sorted per scene, scene-local keys) is bit-exact: each
    scene's keys equal the host's blocked keys of its FP32 positions, and its
    permutation equals a stable argsort of them (ties by id), before and after
    steps that re-bin.
"""
import numpy as np
import pytest

from mpm_amd import lib as L
from mpm_amd import scenes

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.fixture(scope="module")
def small():
    # 6 scenes, 2k-20k particles, grids 32/64/96 cells, both materials
    return scenes.batched_scenes(nscenes=6, counts=(2_000, 20_000), seed=5)


def assert_same(sa, sb, x0, bar=1e-5):
    assert rel_l2(sa["positions"] - x0, sb["positions"] - x0) <= bar
    assert rel_l2(sa["velocities"], sb["velocities"]) <= bar


def test_batch_scenes_match_single_engines(gpu, small):
    steps = 40
    b = small.batch(sort_every=8)
    x0 = [b.download(i)["positions"] for i in range(len(small.scenes))]
    b.step(steps)
    b.synchronize()
    cells = set()
    for i, sc in enumerate(small.scenes):
        eng = sc.engine(sort_every=8)  # the scene's material table is the batch's
        eng.step(sc.dt, steps)
        assert_same(b.download(i), eng.download(), x0[i])
        c = b.conservation(i)
        assert c["grid_mass_q"] == c["particle_mass_q"] and c["nonfinite"] == 0, i
        cells.add(sc.info["cells"])
    assert len(cells) >= 2  # more than one size bucket took part


def test_graph_replay_equals_eager(gpu, small):
    steps = 40  # 5 periods of 8: graphs 0, 1, 0, 1, 0 (odd: the buffers end swapped)
    eager, graph = small.batch(sort_every=8), small.batch(sort_every=8, use_graph=True)
    x0 = [eager.download(i)["positions"] for i in range(len(small.scenes))]
    eager.step(steps)
    graph.step(steps)
    for i in range(len(small.scenes)):
        assert_same(graph.download(i), eager.download(i), x0[i])
    # a partial period after graph replays, then more replays
    eager.step(13)
    graph.step(13)
    eager.step(16)
    graph.step(16)
    for i in range(len(small.scenes)):
        assert_same(graph.download(i), eager.download(i), x0[i])
        c = graph.conservation(i)
        assert c["grid_mass_q"] == c["particle_mass_q"]


@pytest.mark.parametrize("fb", [[64, 64, 64], [128, 128, 128], [64, 128, 256]])
def test_particles_per_group_variants_agree(gpu, small, fb):
    # Equal sizes take one particle launch over all buckets, mixed sizes one
    # launch per bucket; the reference batch (256 everywhere) takes one.
    ref, var = small.batch(sort_every=8), small.batch(sort_every=8)
    var.set_tuning(particles_per_group=fb)
    assert var.tuning()["particles_per_group"] == fb
    x0 = [ref.download(i)["positions"] for i in range(len(small.scenes))]
    ref.step(24)
    var.step(24)
    for i in range(len(small.scenes)):
        assert_same(var.download(i), ref.download(i), x0[i])


def test_autotuner_picks_a_valid_configuration(gpu, small):
    b = small.batch(sort_every=8, use_graph=True)
    t = b.tune(trial_steps=4)
    assert t["buckets"] == 3 and sum(t["bucket_scenes"]) == len(small.scenes)
    for k in range(3):
        assert t["particles_per_group"][k] in (64, 128, 256)
        if t["bucket_scenes"][k]:
            assert t["bucket_ms"][k] > 0
    assert t["sort_every"] in (4, 8, 16, 32) and t["ms_per_step"] > 0
    b.step(2 * t["sort_every"] + 3)
    for i in range(len(small.scenes)):
        c = b.conservation(i)
        assert c["grid_mass_q"] == c["particle_mass_q"] and c["nonfinite"] == 0


def test_batch_rejects_bad_use(gpu, small):
    b = small.batch(sort_every=8)
    sc = small.scenes[0]
    with pytest.raises(L.MpmError):
        b.upload(0, sc.x[:-1], sc.v[:-1], sc.mass[:-1], sc.volume[:-1])  # count differs from the declared one
    with pytest.raises(L.MpmError):
        b.set_dt(0.0)
    with pytest.raises(L.MpmError):
        b.set_tuning(particles_per_group=[96, 256, 256])
    with pytest.raises(L.MpmError):
        b.set_tuning(sort_every=0)
    # a batch whose scenes have no step size yet refuses to step
    from mpm_amd.engine import Batch, make_config
    fresh = Batch([make_config(s.lower, s.upper, s.h) for s in small.scenes], small.counts)
    fresh.set_materials(small.materials)
    for i, s in enumerate(small.scenes):
        fresh.upload(i, s.x, s.v, s.mass, s.volume, temperature=s.temperature, material=s.material)
    with pytest.raises(L.MpmError) as ei:
        fresh.step(1)
    assert ei.value.code == L.MPM_ERR_NOT_READY


def test_batch_binning_is_a_stable_sort_per_scene(gpu):
    from mpm_amd import Material
    from mpm_amd.engine import Batch, make_config
    from test_gpu_engine import cpu_keys
    rng = np.random.default_rng(23)
    # (particles, cells per axis): 3, 1, 2 and 1 sort tiles of 4096 keys
    specs = [(9_000, 32), (100, 64), (5_000, 96), (4_096, 48)]
    cfgs, xs = [], []
    for n, cells in specs:
        h = 1.0 / cells
        cfgs.append(make_config((0, 0, 0), (1, 1, 1), h, n))
        # clustered around a few centres, so cells hold many ties
        centres = rng.uniform(0.2, 0.8, (6, 3))
        xs.append(np.clip(centres[rng.integers(0, 6, n)] + rng.normal(0, 2 * h, (n, 3)), 0.05, 0.95))
    b = Batch(cfgs, [n for n, _ in specs], sort_every=2)
    b.set_materials([Material.from_E_nu(0, 1e5, 0.3)])
    for i, (n, _) in enumerate(specs):
        b.upload(i, xs[i], rng.normal(0, 0.3, (n, 3)), np.full(n, 1e-3), np.full(n, 1e-6))
    b.set_dt([1e-4] * len(specs))
    for _ in range(2):
        kin, kout, perm = b.debug_bin()
        p0 = 0
        for i, (n, cells) in enumerate(specs):
            want = cpu_keys(b.download(i)["positions"], (0, 0, 0), (1, 1, 1), 1.0 / cells)
            order = np.argsort(want, kind="stable")
            assert np.array_equal(kin[p0:p0 + n], want), i
            assert np.array_equal(perm[p0:p0 + n], p0 + order), i
            assert np.array_equal(kout[p0:p0 + n], want[order]), i
            p0 += n
        b.step(7)  # re-bins inside the call move particles between slots
        b.synchronize()
    b.close()
