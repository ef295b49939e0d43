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


def settings(scene=None, stiffness=None, walls=None):
    """mpm2_settings of a scene (softbody, basin, other), optionally with
    another stiffness or walls on/off (the reference's settings.boundary)."""
    s = L.Mpm2Settings()
    L.check(L.load().mpm2_settings_defaults(None if scene is None else scene.encode(), C.byref(s)))
    if stiffness is not None:
        s.stiffness = stiffness
    if walls is not None:
        s.walls = int(bool(walls))
    return s


class Solid2D:
    """The reference's SolidMPM (src/solid-mpm.js) on the device: particles
    in the caller's order, F and C as 4 values per particle (00, 01, 10,
    11)."""

    def __init__(self, capacity):
        self._lib = L.load()
        h = C.c_void_p()
        L.check(self._lib.mpm2_create(int(capacity), C.byref(h)))
        self.h = h
        self.capacity = int(capacity)
        self.count = 0

    def close(self):
        if self.h:
            self._lib.mpm2_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != L.MPM_OK:
            raise L.MpmError(rc, self._lib.mpm2_last_error(self.h).decode())

    def upload(self, x, v=(0, 0), mass=1.1, volume=1.0, F=None, Cmat=None, pinned=None):
        x = np.asarray(x, np.float64).reshape(-1, 2)
        n = len(x)
        col = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float64), (n,)))  # noqa: E731
        v = np.broadcast_to(np.asarray(v, np.float64), (n, 2))
        keep = dict(x=col(x[:, 0]), y=col(x[:, 1]), vx=col(v[:, 0]), vy=col(v[:, 1]),
                    F=np.ascontiguousarray(np.tile([1.0, 0, 0, 1], n) if F is None else np.asarray(F, np.float64)
                                           .reshape(4 * n)),
                    C=np.ascontiguousarray(np.zeros(4 * n) if Cmat is None else np.asarray(Cmat, np.float64)
                                           .reshape(4 * n)),
                    mass=col(mass), volume=col(volume),
                    pinned=np.ascontiguousarray(np.broadcast_to(np.asarray(0 if pinned is None else pinned, np.uint8),
                                                                (n,))))
        p = L.Mpm2Particles()
        p.count, p.dtype = n, L.MPM_F64
        for k, a in keep.items():
            setattr(p, k, a.ctypes.data_as(C.c_void_p))
        self._check(self._lib.mpm2_upload(self.h, C.byref(p)))
        self.count = n

    def download(self):
        n = self.count
        out = dict(x=np.zeros(n), y=np.zeros(n), vx=np.zeros(n), vy=np.zeros(n), F=np.zeros(4 * n),
                   C=np.zeros(4 * n), mass=np.zeros(n), volume=np.zeros(n))
        p = L.Mpm2Particles()
        p.count, p.dtype = n, L.MPM_F64
        for k, a in out.items():
            setattr(p, k, a.ctypes.data_as(C.c_void_p))
        self._check(self._lib.mpm2_download(self.h, C.byref(p)))
        out["positions"] = np.stack([out["x"], out["y"]], 1)
        out["velocities"] = np.stack([out["vx"], out["vy"]], 1)
        return out

    def step(self, dt, nsteps=1, s=None):
        self._check(self._lib.mpm2_step(self.h, C.byref(s or settings(walls=False)), float(dt), int(nsteps)))

    def stable_step(self, s=None):
        out = C.c_double()
        self._check(self._lib.mpm2_stable_dt(self.h, C.byref(s or settings(walls=False)), C.byref(out)))
        return out.value

    def advance(self, dt=1 / 60, gravity="none", strength=1.0, s=None, max_substeps=96):
        n, adv = C.c_int32(), C.c_double()
        g = GRAVITY[gravity] if isinstance(gravity, str) else int(gravity)
        self._check(self._lib.mpm2_advance(self.h, C.byref(s or settings(walls=False)), float(dt), g, float(strength),
                                           int(max_substeps), C.byref(n), C.byref(adv)))
        return n.value, adv.value

    def stats(self):
        st = L.Mpm2Stats()
        self._check(self._lib.mpm2_statistics(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in L.Mpm2Stats._fields_}
