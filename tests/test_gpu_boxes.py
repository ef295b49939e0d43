"""GPU parity of the P2G node boxes on every path, against the FP64 oracle
through the C ABI.

The fused P2G adds each run sum into a node box in LDS (src/three/mpm.js:66-85
is the scatter it replaces) and takes one of four paths per 256-particle
block, decided from the base cells of its particles in sorted order:
  1. the block's box fits the buffer (640 nodes);
  2. else each wave's own box fits its quarter (160 nodes);
  3. else the wave adds its run sums to the grid with global atomics.
The scene below (a dense block of 8 particles per cell and about one
particle per cell scattered over a slab that spans whole rows of 4x4x4 cell
blocks) reaches all three; `paths` restates the decision on the CPU to check
that it does.

Bars: one P2G from a state with random F, C and v against the oracle's P2G,
node mass within 2e-6 and momentum within 2e-5 of the largest (as the
single-P2G test of the dense cube); 20 steps against the oracle stepping the
same scene, positions within the reference's CPU-GPU bar
(gpu/test_solver.py:42: atol .003 at spacing 4, scaled to the spacing here,
rtol .001), grid mass equal to particle mass exactly.
"""
import numpy as np
import pytest

from mpm_amd import Engine, Material, scenes
from oracle import binding as o
from test_gpu_engine import cpu_keys

LOWER, UPPER, H = (0, 0, 0), (1, 1, 1), 1 / 32


def box_scene(seed=5):
    rng = np.random.default_rng(seed)
    g = (np.arange(16) + 0.5) * (H / 2)
    dense = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + 0.2
    sparse = np.column_stack([rng.uniform(0.05, 0.95, 6000), rng.uniform(0.55, 0.9, 6000),
                              rng.uniform(0.55, 0.62, 6000)])
    return np.concatenate([dense, sparse])


def paths(x, fb=256, block_nodes=640, wave_nodes=160):
    """Particles per path (block box, wave box, global atomics) for a
    freshly sorted scene."""
    origin, nodes = scenes.domain(LOWER, UPPER, H)
    N = int(nodes.max())
    q = (np.asarray(x).astype(np.float32) - origin.astype(np.float32)) * np.float32(1 / H)
    b = np.clip(np.floor(q - np.float32(0.5)).astype(np.int64), 0, N - 3)
    b = b[np.argsort(cpu_keys(x, LOWER, UPPER, H), kind="stable")]

    def nodes_of(c):
        return int(np.prod(c.max(0) - c.min(0) + 3)) if len(c) else 0

    out = np.zeros(3, int)
    for f in range(0, len(b), fb):
        c = b[f:f + fb]
        if nodes_of(c) <= block_nodes:
            out[0] += len(c)
            continue
        for w in range(0, len(c), 64):
            cw = c[w:w + 64]
            out[1 if nodes_of(cw) <= wave_nodes else 2] += len(cw)
    return out


def oracle_scene(x, v, mass, vol, mat):
    return o.Scene(x, v, mass, vol, lower=LOWER, upper=UPPER, h=H, vol_scale=1.0,
                   parts=[o.make_part(E=mat.E, G=mat.mu / mat.E, L=mat.lam / mat.E)], grid_gravity=(0, 0, 0))


def test_scene_reaches_every_box_path():
    counts = paths(box_scene())
    print(f"\nparticles per path (block box, wave box, global atomics): {counts.tolist()}")
    assert (counts > 0).all()


@pytest.mark.gpu
def test_p2g_on_every_box_path_matches_oracle(gpu):
    x = box_scene()
    n = len(x)
    rng = np.random.default_rng(9)
    F = (np.eye(3) + rng.normal(0, .03, (n, 3, 3))).reshape(n, 9)
    Cm = rng.normal(0, 1.0, (n, 9))
    v = rng.normal(0, .05, (n, 3))
    vol = (H / 2) ** 3
    mat = Material.from_E_nu(0, 1e5, 0.3)
    eng = Engine(LOWER, UPPER, H, n, gravity=(0, 0, 0))
    eng.set_materials([mat])
    eng.upload(x, v, 1000 * vol, vol, F=F, Cmat=Cm)
    gm, gp, N = eng.debug_p2g(1e-4)
    orc = oracle_scene(x, v, 1000 * vol, vol, mat)
    orc.F[:] = F.reshape(-1)
    orc.C[:] = Cm.reshape(-1)
    orc.p2g(1e-4)
    assert N == orc.N
    assert np.abs(gm - orc.gmass).max() <= 2e-6 * orc.gmass.max()
    op = orc.gv.reshape(-1, 3)
    assert np.abs(gp - op).max() <= 2e-5 * np.abs(op).max()


@pytest.mark.gpu
def test_steps_on_every_box_path_match_oracle(gpu):
    x = box_scene()
    n = len(x)
    rng = np.random.default_rng(10)
    v = rng.normal(0, .2, (n, 3))
    vol = (H / 2) ** 3
    mat = Material.from_E_nu(0, 1e5, 0.3)
    eng = Engine(LOWER, UPPER, H, n, gravity=(0, 0, 0), sort_every=64)
    eng.set_materials([mat])
    eng.upload(x, v, 1000 * vol, vol)
    orc = oracle_scene(x, v, 1000 * vol, vol, mat)
    dt = 2e-4
    eng.step(dt, 20)
    for _ in range(20):
        orc.step(dt)
    c = eng.conservation()
    assert c["grid_mass_q"] == c["particle_mass_q"]
    st = eng.download()
    np.testing.assert_allclose(st["positions"], orc.positions(), atol=.003 * (H / 2) / 4, rtol=.001)
