"""GPU parity of the re-bin inside a call, through the C ABI.

A re-bin that falls due inside one mpm_step call sorts the particles between
a grid update and the fused step after it, and moves only the planes a
particle carries from one step to the next: x, F, m, V, T, ep and, where a
material keeps them, Jp, dilation and damage; a pinned particle moves every
plane (it skips G2P and keeps its stored v, C and stress). A re-bin at the
start of a call moves every plane.

Scene: two blocks colliding head-on at 120 units/s in the reference's units
(cell 8, mass 1.1, particle volume 1, E = m/(64 V) 16000, mu = lambda =
0.4 E), snow (model 2, Jp) on the left and sand (model 3, dilation) on the
right, with four pinned particles; 40 steps of 1/240 through the collision.

Each re-bin inside the call is a sort and a fused step that reads the
particles through the new order and writes them in it; re-binning every step
chains those steps back to back.

Bars: one call of 40 steps re-binned every step or every 3 steps against the
FP64 oracle
-- relative L2 of the displacement, velocity, Jp - 1 and dilation within
5e-3 (the colliding-blocks bar of the models' tests), pinned particles in
place, grid mass equal to particle mass exactly; the same steps as 40 calls
of one step (every re-bin at a call start, every plane moved) within 1e-4
relative L2 of the one-call run, whose re-bins fall inside the call.
"""
import numpy as np
import pytest

from mpm_amd import Engine, Material
from oracle import binding as o

pytestmark = pytest.mark.gpu

LOWER, UPPER, H = (-64, -48, -64), (64, 72, 64), 8.0
E_REF = 1.1 / 64 * 16000
DT, STEPS = 1 / 240, 40


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def scene():
    nb, sp, gap, speed = 5, 4.0, 6.0, 60.0
    pts, vel, part = [], [], []
    w = (nb - 1) * sp
    for k, s in enumerate((-1, 1)):
        x0 = s * (gap / 2 + w / 2)
        for z in range(nb):
            for y in range(nb):
                for x in range(nb):
                    pts.append([x0 - w / 2 + sp * x, -w / 2 + sp * y + 8, -w / 2 + sp * z])
                    vel.append([-s * speed, 0, 0])
                    part.append(k)
    pts, vel, part = np.array(pts), np.array(vel), np.array(part)
    pinned = np.zeros(len(pts), np.uint8)
    pinned[[0, 62, 130, 249]] = 1  # corners and centres of both blocks
    return pts, vel, part, pinned


def engine(pts, vel, part, pinned, sort_every):
    eng = Engine(LOWER, UPPER, H, len(pts), strict_domain=True, sort_every=sort_every)
    eng.set_materials([Material(2, E_REF, .4 * E_REF, .4 * E_REF), Material(3, E_REF, .4 * E_REF, .4 * E_REF)])
    eng.upload(pts, vel, 1.1, 64.0, material=part.astype(np.uint8), pinned=pinned)
    return eng


@pytest.mark.parametrize("sort_every", [1, 3])
def test_rebin_inside_a_call_matches_oracle_with_history_and_pins(gpu, sort_every):
    pts, vel, part, pinned = scene()
    one = engine(pts, vel, part, pinned, sort_every)
    one.step(DT, STEPS)  # re-bins after every sort_every steps fall inside the call
    many = engine(pts, vel, part, pinned, sort_every)
    for _ in range(STEPS):
        many.step(DT, 1)  # every re-bin at a call start
    orc = o.Scene(pts, vel, 1.1, 1.0, lower=LOWER, upper=UPPER, h=H, part=part, model=np.where(part == 0, 2, 3),
                  pinned=pinned)
    for _ in range(STEPS):
        orc.step(DT)

    a, b = one.download(), many.download()
    xo = orc.positions()
    e_dx = rel_l2(a["positions"] - pts, xo - pts)
    e_v = rel_l2(a["velocities"], orc.velocities())
    e_jp = rel_l2(a["jp"][part == 0] - 1, orc.Jp[part == 0] - 1)
    e_dil = rel_l2(a["dilation"][part == 1], orc.dilation[part == 1])
    msg = f"rel L2 dx {e_dx:.2e} v {e_v:.2e} Jp-1 {e_jp:.2e} dilation {e_dil:.2e}"
    print(msg)
    assert orc.Jp[part == 0].min() < .95 and orc.dilation[part == 1].max() > .01, msg  # both histories move
    assert e_dx < 5e-3 and e_v < 5e-3 and e_jp < 5e-3 and e_dil < 5e-3, msg
    pin = pinned == 1
    np.testing.assert_allclose(a["positions"][pin], pts[pin], atol=1e-4)
    np.testing.assert_allclose(xo[pin], pts[pin], atol=1e-12)

    for key in ("positions", "velocities", "jp", "dilation"):
        ref = b[key] - (pts if key == "positions" else 1.0 if key == "jp" else 0.0)
        got = a[key] - (pts if key == "positions" else 1.0 if key == "jp" else 0.0)
        assert rel_l2(got, ref) < 1e-4, (key, rel_l2(got, ref))
    for eng in (one, many):
        c = eng.conservation()
        assert c["grid_mass_q"] == c["particle_mass_q"] and c["nonfinite"] == 0
        eng.close()
