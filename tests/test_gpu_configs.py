"""The benchmark's large configurations under test, through the C ABI.

Large single scene (256^3 cells, 16.5M particles: a thermal-J2 steel block on
the floor and 32 falling fixed-corotated spheres, T 300-1200 K, dt 1e-5):
  * 20 steps at full size: fixed-point grid mass equal to the particle mass
    after every step, no non-finite particle, min J > 0, the guard silent;
    no horizontal momentum appears (|P_x|, |P_z| within 1e-5 of M g t plus
    1e-5 of the start momentum, the FP32 rounding of the velocities); the
    vertical momentum lies between free fall of everything and free fall of
    the spheres alone (the floor only pushes up and the spheres are still in
    the air) within the same tolerance; binning is deterministic (two
    engines from the same scene give the same permutation).
  * the same scene at 64^3 cells (220k particles) with the spheres lowered
    onto the block at 2 m/s, 45 steps of impact against the FP64 oracle:
    positions within the reference's CPU-GPU bar (gpu/test_solver.py:42, atol
    .003 at particle spacing 4, scaled to this spacing, rtol .001), relative
    L2 of displacement and of plastic strain <= 5e-3.
Batched scenes (256 scenes, 11.6M particles, grids 32^3/64^3/96^3, mixed
materials):
  * 300 steps: every scene keeps its fixed-point mass exactly, stays finite,
    min J > 0 (the soft elastic scenes hitting a wall at 1 m/s compress to
    J of about 0.4), the guard is silent;
  * three sampled scenes (the smallest of each material and a mid-sized one)
    stepped 60 steps in the batch against the FP64 oracle stepping the same
    scene alone: relative L2 of displacement and velocity <= 5e-3 and
    positions within the reference's bar.
"""
import numpy as np
import pytest

from mpm_amd import scenes
from oracle import binding as o

pytestmark = pytest.mark.gpu

G = 9.81


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def oracle_scene(sc, x=None, v=None):
    """The scene stepped by the FP64 oracle: one oracle part per material."""
    parts = []
    for m in sc.materials:
        th = o.thermal_params(**m.thermal) if m.thermal else None
        parts.append(o.make_part(E=m.E, G=m.mu / m.E, L=m.lam / m.E, thermal=th))
    models = np.array([sc.materials[k].model for k in sc.material], np.int32)
    return o.Scene(sc.x if x is None else x, sc.v if v is None else v, sc.mass, sc.volume, lower=sc.lower,
                   upper=sc.upper, h=sc.h, vol_scale=1.0, parts=parts, model=models,
                   part=sc.material.astype(np.int32), temperature=sc.temperature, wall_mode=sc.wall_mode,
                   grid_gravity=(0, 0, 0) if sc.gravity_on_particles else sc.gravity)


def test_large_scene_full_size_invariants(gpu):
    sc = scenes.large_scene(steps=20)
    spheres = sc.material == 1
    M, Ms = sc.mass.sum(), sc.mass[spheres].sum()
    P0 = (sc.mass[:, None] * sc.v).sum(0)
    eng = sc.engine(sort_every=8)
    print(f"\nlarge scene: {sc.n} particles, {spheres.sum()} in spheres")
    for k in range(1, sc.steps + 1):
        eng.step(sc.dt, 1)
        c = eng.conservation()
        assert c["grid_mass_q"] == c["particle_mass_q"], k
        assert c["nonfinite"] == 0 and c["min_J"] > 0, k
        t = k * sc.dt
        tol = 1e-5 * M * G * t + 1e-5 * np.abs(P0).sum()
        assert abs(c["momentum"][0] - P0[0]) <= tol and abs(c["momentum"][2] - P0[2]) <= tol, (k, c["momentum"])
        assert P0[1] - M * G * t - tol <= c["momentum"][1] <= P0[1] - Ms * G * t + tol, (k, c["momentum"])
    eng.synchronize()
    _, _, perm_a = eng.debug_bin()
    del eng
    eng2 = sc.engine(sort_every=8)
    eng2.step(sc.dt, sc.steps)
    _, _, perm_b = eng2.debug_bin()
    assert np.array_equal(perm_a, perm_b)


def lowered_spheres(sc, gap_cells=1.0):
    """Positions with the spheres moved down, lowest first, so that each
    lowest point is gap_cells cells above the block's top; a sphere that
    would then come within a cell of one already moved stays where it is.
    Returns (positions, number of spheres moved)."""
    x = sc.x.copy()
    top = sc.x[sc.material == 0, 1].max()
    placed = []
    for c, r in sorted(sc.info["spheres"], key=lambda s: s[0][1] - s[1]):
        c = np.asarray(c)
        new = c.copy()
        new[1] = top + gap_cells * sc.h + r
        if any(np.linalg.norm(new - c2) < r + r2 + sc.h for c2, r2 in placed):
            continue
        mine = (sc.material == 1) & (np.linalg.norm(sc.x - c, axis=1) <= r * 1.001 + sc.h)
        x[mine, 1] += new[1] - c[1]
        placed.append((new, r))
    return x, len(placed)


def test_large_scene_reduced_matches_oracle(gpu):
    sc = scenes.large_scene(cells=64, nspheres=6, drop=2.0, dt=4e-5, steps=45)
    x, moved = lowered_spheres(sc)
    assert moved >= 3
    eng = sc.engine(sort_every=8)
    eng.upload(x, sc.v, sc.mass, sc.volume, temperature=sc.temperature, material=sc.material)
    xg0 = eng.download()["positions"]
    orc = oracle_scene(sc, x=x)
    x0 = orc.positions().copy()
    for k in range(3):
        eng.step(sc.dt, sc.steps // 3)
        c = eng.conservation()
        assert c["grid_mass_q"] == c["particle_mass_q"]
    eng.synchronize()
    for _ in range(sc.steps):
        orc.step(sc.dt)
    st = eng.download()
    spacing = sc.info["spacing"]
    e_dx = rel_l2(st["positions"] - xg0, orc.positions() - x0)
    e_ep = rel_l2(st["plastic"], orc.plastic)
    print(f"\nreduced large scene ({sc.n} particles, {sc.steps} steps): rel L2 dx {e_dx:.2e} ep {e_ep:.2e}, "
          f"ep max {orc.plastic.max():.2e}")
    assert orc.plastic.max() > 1e-3  # the impact yields the steel
    np.testing.assert_allclose(st["positions"], orc.positions(), atol=.003 * spacing / 4, rtol=.001)
    assert e_dx <= 5e-3 and e_ep <= 5e-3


@pytest.fixture(scope="module")
def batch256():
    return scenes.batched_scenes()


def test_batch_256_scenes_invariants(gpu, batch256):
    bs = batch256
    b = bs.batch(sort_every=4, use_graph=True)
    b.step(bs.steps)
    b.synchronize()
    for i in range(len(bs.scenes)):
        c = b.conservation(i)
        assert c["grid_mass_q"] == c["particle_mass_q"], i
        assert c["nonfinite"] == 0 and c["min_J"] > 0, (i, c["min_J"])


def test_batch_sampled_scenes_match_oracle(gpu, batch256):
    bs = batch256
    counts = np.array(bs.counts)
    mats = np.array([sc.info["material"] for sc in bs.scenes])
    picks = [int(np.flatnonzero(mats == m)[np.argmin(counts[mats == m])]) for m in (0, 1)]
    order = np.argsort(counts)
    picks.append(int(order[len(order) // 2]))
    steps = 60
    b = bs.batch(sort_every=4)
    x0 = {i: b.download(i)["positions"] for i in picks}
    b.step(steps)
    b.synchronize()
    for i in picks:
        sc = bs.scenes[i]
        orc = oracle_scene(sc)
        xo0 = orc.positions().copy()
        for _ in range(steps):
            orc.step(sc.dt)
        st = b.download(i)
        e_dx = rel_l2(st["positions"] - x0[i], orc.positions() - xo0)
        e_v = rel_l2(st["velocities"], orc.velocities())
        spacing = sc.h / sc.info["per_axis"]
        print(f"\nscene {i}: {sc.n} particles, {sc.info['cells']}^3 cells, material {sc.info['material']}: "
              f"rel L2 dx {e_dx:.2e} v {e_v:.2e}")
        np.testing.assert_allclose(st["positions"], orc.positions(), atol=.003 * spacing / 4, rtol=.001)
        assert e_dx <= 5e-3 and e_v <= 5e-3
