"""2D elastic MLS-MPM (the reference's SolidMPM, src/solid-mpm.js) on the
device, through the library's mpm2_* C ABI, and the panel seeding of
SolidMPM.create (src/solid-mpm.js:93-105)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import lib as L

DX, ORIGIN, NODES, AREA, SPACING = 8.0, -592.0, 149, 12.96, 3.6  # src/solid-mpm.js:4, :99
# Wall boxes (bound, floor, top) of the 2D driver: the softbody demo, the
# basin, and every other scene (src/solid-mpm.js:69-70).
WALLS = dict(softbody=(170.0, -145.0, 145.0), basin=(245.0, -175.0, 245.0), default=(555.0, -555.0, 555.0))
GRAVITY = dict(none=0, uniform=1, center=2)


def panel(x1, y1, x2, y2, shape="panel", fixed=False):
    """Lattice of SolidMPM.create: a panel between two corners, or a 22-wide
    beam along the segment, at spacing 3.6 after clamping to +-530. Returns
    (points [n, 2], pinned [n] uint8), or None where the reference refuses."""
    if not all(math.isfinite(v) for v in (x1, y1, x2, y2)):
        return None
    cl = lambda v: max(-530.0, min(530.0, v))  # noqa: E731
    x1, x2, y1, y2 = cl(x1), cl(x2), cl(y1), cl(y2)
    length = math.hypot(x2 - x1, y2 - y1)
    is_panel = shape == "panel"
    width = abs(x2 - x1) if is_panel else length
    height = abs(y2 - y1) if is_panel else 22.0
    if width < 7 or height < 7:
        return None
    nx, ny = int(width // SPACING) + 1, int(height // SPACING) + 1
    ux, uy = (x2 - x1) / length, (y2 - y1) / length
    pts, pins = [], []
    sx, sy = math.copysign(1, x2 - x1) if x2 != x1 else 0.0, math.copysign(1, y2 - y1) if y2 != y1 else 0.0
    for row in range(ny):
        for col in range(nx):
            if is_panel:
                x, y = x1 + sx * col * SPACING, y1 + sy * row * SPACING
            else:
                x = x1 + ux * col * SPACING - uy * (row * SPACING - height / 2)
                y = y1 + uy * col * SPACING + ux * (row * SPACING - height / 2)
            pts.append((x, y))
            pins.append(1 if fixed and col < 2 else 0)
    return np.array(pts, np.float64), np.array(pins, np.uint8)
