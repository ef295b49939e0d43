"""GPU parity of the split step (mpm_step_to_grid / mpm_grid_download /
mpm_grid_upload_velocity / mpm_step_from_grid): the reference's grid hooks
inside MPM3.step, externalGridContact (src/three/mpm.js:104) and
contactGridDeltas (:114), against the oracle's split step, through the C ABI.

Bars: the grid view after P2G and the update -- node mass within 2e-6 of the
largest, node momentum m v within 2e-5 of the largest (the single-P2G bar;
a node with a tail weight has a tiny mass whose fixed-point rounding shows
in v alone), pinned flags equal;
trajectories with a hook acting every step -- positions atol .003 rtol .001
(gpu/test_solver.py:42) after 30 steps; grid mass equal to particle mass."""
import numpy as np
import pytest

from mpm_amd import Engine, Material
from oracle import binding as o

pytestmark = pytest.mark.gpu

LOWER, UPPER, H = (-64, -48, -64), (64, 72, 64), 8.0
E_REF = 1.1 / 64 * 16000


def two_blocks():
    a = np.arange(5) * 3.0 - 6
    pts = np.array([[x0 + x, y + 10, z] for x0 in (-14, 14) for z in a for y in a for x in a])
    part = np.repeat([0, 1], 125)
    return pts, part


def pair(pinned=None):
    pts, part = two_blocks()
    n = len(pts)
    v = np.where(part[:, None] == 0, [20.0, -10.0, 0.0], [-20.0, -10.0, 5.0])
    eng = Engine(LOWER, UPPER, H, n, strict_domain=True)
    eng.set_materials([Material(0, E_REF, .4 * E_REF, .4 * E_REF), Material(10, E_REF, .4 * E_REF, .4 * E_REF)])
    eng.upload(pts, v, 1.1, 64.0, material=part.astype(np.uint8), pinned=pinned, temperature=20.0)
    orc = o.Scene(pts, v, 1.1, 1.0, lower=LOWER, upper=UPPER, h=H, part=part, model=np.where(part == 0, 0, 10),
                  pinned=pinned)
    return eng, orc, pts


def test_grid_view_matches_the_oracle_grid(gpu):
    pinned = np.zeros(250, np.uint8)
    pinned[[3, 130]] = 1
    eng, orc, _ = pair(pinned)
    dt = 1 / 240
    for _ in range(3):
        eng.step(dt)
        orc.step(dt)
    eng.step_to_grid(dt)
    orc.step_to_grid(dt)
    m, v, f = eng.grid()
    mo = orc.gmass
    vo = orc.gv.reshape(-1, 3)
    active = mo > 1e-12
    assert np.abs(m - mo).max() <= 2e-6 * mo.max()
    P, Po = m[:, None] * v, (mo[:, None] * vo) * active[:, None]
    assert np.abs(P - Po).max() <= 2e-5 * np.abs(Po).max()
    heavy = mo > 1e-6 * mo.max()
    np.testing.assert_array_equal(f[heavy], orc.gfixed[heavy])
    assert f.sum() > 0 and (v[f == 1] == 0).all()
    assert (m[~active] == 0).all()
    eng.step_from_grid(dt)
    orc.step_from_grid(dt)
    np.testing.assert_allclose(eng.download()["positions"], orc.positions(), atol=.003, rtol=.001)


def floor_plate(m, v, y_plate):
    """A rigid plate below y_plate: nodes with mass stop moving into it."""
    N = round(len(m) ** (1 / 3))
    iy = (np.arange(len(m)) // N) % N
    y = -48 - 16 + iy * H  # domain origin floor(lower/h) h - 2h
    hit = (m > 1e-10) & (y < y_plate) & (v[:, 1] < 0)
    v = v.copy()
    v[hit, 1] = 0
    return v


def test_external_grid_contact_hook_matches_the_oracle(gpu):
    eng, orc, pts = pair()
    dt = 1 / 240
    for k in range(30):
        eng.step_to_grid(dt)
        orc.step_to_grid(dt)
        m, v, _ = eng.grid()
        eng.set_grid_velocity(floor_plate(m, v, 6.0))
        vo = floor_plate(orc.gmass, orc.gv.reshape(-1, 3), 6.0)
        orc.gv[:] = vo.reshape(-1)
        eng.step_from_grid(dt)
        orc.step_from_grid(dt)
    x = eng.download()["positions"]
    np.testing.assert_allclose(x, orc.positions(), atol=.003, rtol=.001)
    assert orc.positions()[:, 1].min() > pts[:, 1].min() - 10  # the plate held the blocks up
    c = eng.conservation()
    assert c["grid_mass_q"] == c["particle_mass_q"]


def test_contact_grid_deltas_match_the_oracle(gpu):
    eng, orc, _ = pair()
    N = eng.grid_nodes()
    rng = np.random.default_rng(4)
    delta = np.tile([.5, 0, -.25], (N ** 3, 1)) + rng.normal(0, .05, (N ** 3, 3))
    dt = 1 / 240
    for k in range(30):
        eng.step_to_grid(dt)
        orc.step_to_grid(dt)
        eng.step_from_grid(dt, {1: delta})  # material 1 is part 1
        orc.step_from_grid(dt, {1: delta})
    x = eng.download()["positions"]
    np.testing.assert_allclose(x, orc.positions(), atol=.003, rtol=.001)
    # the delta moved part 1 relative to a run without it
    _, orc2, _ = pair()
    for _ in range(30):
        orc2.step(dt)
    assert np.abs(orc.positions()[125:, 0] - orc2.positions()[125:, 0]).mean() > .05


def test_split_step_ordering_is_enforced(gpu):
    from mpm_amd import lib as L
    eng, _, _ = pair()
    with pytest.raises(L.MpmError):
        eng.grid()  # nothing pending
    eng.step_to_grid(1e-3)
    with pytest.raises(L.MpmError):
        eng.step(1e-3)
    with pytest.raises(L.MpmError):
        eng.step_to_grid(1e-3)
    with pytest.raises(L.MpmError):
        eng.step_from_grid(2e-3)  # another dt
    eng.step_from_grid(1e-3)
    eng.step(1e-3)
