"""Scene seeding and the benchmark workloads.

Seeding restates the reference's lattice and jitter rules, vectorised:
  random(seed, index, attempt, axis)   src/three/sampling.js:8-12
  jitterPoint                          src/three/sampling.js:13-29
  geometry lattice (unrotated)         src/three/discretization.js:33-48
  domain grid                          src/three/discretization.js:4-15
The workloads are the MLS-MPM configurations the engine is measured on:
  elastic_cube    unit domain, 32^3 cells, 0.3 cube, fixed-corotated
  thermal_boxes   128^3 cells, ~2M particles in 6 seeded boxes, thermal J2
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .engine import Material

_M32 = 0xFFFFFFFF


def random(seed, index, attempt, axis):
    """Seeded uint32 hash mapped to [-1, 1); vectorised over numpy inputs."""
    # uint32 arithmetic wraps modulo 2^32 as the reference's Math.imul does.
    u32 = np.uint32
    seed = np.asarray(np.asarray(seed, np.int64) & _M32, u32)
    index = np.asarray(index, np.int64)
    attempt = np.asarray(attempt, np.int64)
    axis = np.asarray(axis, np.int64)
    with np.errstate(over="ignore"):
        a = ((index + 1) & _M32).astype(u32) * u32(0x9E3779B9)
        b = ((attempt * 3 + axis + 1) & _M32).astype(u32) * u32(0x85EBCA6B)
        h = seed + a + b
        h = (h ^ (h >> u32(16))) * u32(0x7FEB352D)
        h = (h ^ (h >> u32(15))) * u32(0x846CA68B)
        h = h ^ (h >> u32(16))
    return h.astype(np.float64) / 4294967296.0 * 2 - 1


def uniform(seed, index, attempt, axis, lo, hi):
    """Uniform draw in [lo, hi) from the seeded hash."""
    return lo + (hi - lo) * (random(seed, index, attempt, axis) + 1) * 0.5


def normal(seed, index, attempt, axis):
    """Standard normal draw (Box-Muller on two hash draws)."""
    u1 = (random(seed, index, attempt, 2 * axis) + 1) * 0.5
    u2 = (random(seed, index, attempt, 2 * axis + 1) + 1) * 0.5
    u1 = np.maximum(u1, 1e-300)
    return np.sqrt(-2 * np.log(u1)) * np.cos(2 * np.pi * u2)


def domain(lower, upper, h):
    """Background grid: origin = floor(lower/h) h - 2h, nodes = ceil((upper -
    origin)/h) + 3 per axis (src/three/discretization.js:13)."""
    lower, upper = np.asarray(lower, float), np.asarray(upper, float)
    origin = np.floor(lower / h) * h - 2 * h
    nodes = (np.ceil((upper - origin) / h) + 3).astype(np.int64)
    return origin, nodes


SHAPES = {"box": 0, "ellipsoid": 1, "cylinder": 2}


def lattice(center, size, shape, spacing):
    """Cell-centred lattice points (k + 1/2) spacing inside an unrotated shape,
    z outermost, x fastest (src/three/discretization.js:33-48)."""
    center, size = np.asarray(center, float), np.asarray(size, float)
    lo, hi = center - size * .5, center + size * .5
    first = np.ceil(lo / spacing - .5 - 1e-10).astype(np.int64)
    last = np.ceil(hi / spacing - .5 - 1e-10).astype(np.int64) - 1
    zs, ys, xs = [np.arange(first[a], last[a] + 1) for a in (2, 1, 0)]
    Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
    p = np.stack([(X.ravel() + .5) * spacing, (Y.ravel() + .5) * spacing, (Z.ravel() + .5) * spacing], 1)
    q = 2 * (p - center) / size
    inside = np.all(np.abs(q) <= 1 + 1e-10, axis=1)
    if shape == "ellipsoid":
        inside &= (q ** 2).sum(1) <= 1 + 1e-10
    elif shape == "cylinder":
        inside &= q[:, 0] ** 2 + q[:, 2] ** 2 <= 1 + 1e-10
    return p[inside]


def jitter(points, lo, hi, shape, seed, strength, spacing, fixed=None):
    """Bounded seeded jitter of points in local coordinates with reflection at
    the faces and up to 8 retries (src/three/sampling.js:13-29). Point k uses
    hash index k."""
    p = np.asarray(points, float)
    n = len(p)
    if strength == 0 or n == 0:
        return p.copy()
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    fixed = np.zeros(n, bool) if fixed is None else np.asarray(fixed, bool)
    amplitude = spacing * strength
    out = p.copy()
    idx = np.arange(n)  # points still without an accepted attempt
    for attempt in range(8):
        if idx.size == 0:
            break
        d = np.stack([random(seed, idx, attempt, a) for a in range(3)], 1)
        d[fixed[idx], 0] = 0
        scale = amplitude / np.maximum(1.0, np.sqrt(d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]))
        c = p[idx] + d * scale[:, None]
        c = np.where(c < lo, 2 * lo - c, c)
        c = np.where(c > hi, 2 * hi - c, c)
        q = (c - (lo + hi) / 2) / ((hi - lo) / 2)
        if shape == "box":
            ok = np.ones(idx.size, bool)
        elif shape == "ellipsoid":
            ok = q[:, 0] ** 2 + q[:, 1] ** 2 + q[:, 2] ** 2 <= 1 + 1e-12
        else:
            ok = q[:, 0] ** 2 + q[:, 2] ** 2 <= 1 + 1e-12
        out[idx[ok]] = c[ok]
        idx = idx[~ok]
    return out


def geometry(center, size, shape, spacing, seed=42, strength=0.0):
    """Seed one geometry: lattice, then jitter in local coordinates
    (src/three/simulation.js:66-71 with rotation 0). Returns (points, V_p)."""
    center, size = np.asarray(center, float), np.asarray(size, float)
    pts = lattice(center, size, shape, spacing)
    local = pts - center
    if strength > 0:
        local = jitter(local, -size * .5, size * .5, shape, seed, strength, spacing)
    vol = float(np.prod(size)) * (np.pi / 6 if shape == "ellipsoid" else np.pi / 4 if shape == "cylinder" else 1.0)
    return local + center, vol / max(len(pts), 1)


@dataclass
class Scene:
    """A workload: particle arrays plus engine configuration."""
    name: str
    lower: tuple
    upper: tuple
    h: float
    x: np.ndarray
    v: np.ndarray
    mass: np.ndarray
    volume: np.ndarray
    temperature: np.ndarray
    material: np.ndarray
    materials: list
    dt: float
    steps: int
    gravity: tuple = (0.0, -9.81, 0.0)
    wall_mode: tuple = (0, 0, 0, 0, 0, 0)
    gravity_on_particles: bool = False
    pinned: np.ndarray = None
    info: dict = field(default_factory=dict)
    # Courant number and gradient bound of the state guard
    # (mpm_config.cfl_number, grad_number); 0: the reference's 1.4/8 and .08.
    courant: float = 0.0
    gradient: float = 0.0

    @property
    def n(self):
        return len(self.x)

    def engine(self, capacity=None, sort_every=1):
        from .engine import Engine
        e = Engine(self.lower, self.upper, self.h, capacity or self.n, wall_mode=self.wall_mode,
                   gravity=self.gravity, gravity_on_particles=self.gravity_on_particles, sort_every=sort_every,
                   cfl_number=self.courant, grad_number=self.gradient)
        e.set_materials(self.materials)
        e.upload(self.x, self.v, self.mass, self.volume, temperature=self.temperature, material=self.material,
                 pinned=self.pinned)
        return e


def elastic_cube(nu=0.3, E=1e5, rho=1000.0, sticky_floor=True, seed=42, strength=0.2):
    """Elastic cube in the unit domain: h = 1/32, 2 particles per axis per cell
    (spacing 1/64), side 0.3 centred in x/z with its base at y = 0.04, seeded
    jitter 0.2; fixed-corotated E = 1e5 Pa, nu = 0.3, rho = 1000, gravity
    -9.81, sticky floor, dt = 1e-4, 200 steps. nu = 0.25 with separating walls
    and gravity on the particles is the reference's own convention
    (mu = lambda = 0.4 E, src/three/mpm.js:37; src/three/simulation.js:337-344)."""
    h = 1 / 32
    pts, vp = geometry((0.5, 0.19, 0.5), (0.3, 0.3, 0.3), "box", h / 2, seed, strength)
    n = len(pts)
    mat = Material.from_E_nu(0, E, nu)
    walls = (0, 0, 1, 0, 0, 0) if sticky_floor else (0,) * 6
    return Scene("elastic_cube", (0, 0, 0), (1, 1, 1), h, pts, np.zeros((n, 3)), np.full(n, rho * vp),
                 np.full(n, vp), np.full(n, 20.0), np.zeros(n, np.uint8), [mat], 1e-4, 200, wall_mode=walls,
                 gravity_on_particles=not sticky_floor,
                 info=dict(E=E, nu=nu, rho=rho, spacing=h / 2, particle_volume=vp))


# Thermal J2 of the surrogate scenes: yield A = 0.02 E with linear softening to
# 10% between 300 K and 1200 K, perfectly plastic (H = 0), so that the trial
# strains of the elastic-cube scene fall in the surrogate's training range.
SURROGATE_THERMAL = dict(Tref=300.0, Thot=1200.0, hotRatio=0.1, A=0.02, H=0.0, kelvin_offset=0.0)


def j2_cube(nu=0.3, E=1e5, rho=1000.0, seed=42, strength=0.2, T=(300.0, 900.0), thermal=None, steps=200,
            squeeze=0.5):
    """The elastic-cube scene with a thermal-J2 material (SURROGATE_THERMAL),
    per-particle temperature uniform in [300, 900] K and a vertical squeeze:
    start velocity v_y = -squeeze (y - y_c) / (side / 2), a uniaxial strain
    rate of 3.3/s at squeeze = 0.5 m/s, so that the 200 steps reach about
    6.7% strain against a yield strain of 0.8-2%: the scene of the learned
    constitutive surrogate."""
    sc = elastic_cube(nu=nu, E=E, rho=rho, seed=seed, strength=strength)
    n = sc.n
    sc.name = "j2_cube"
    sc.v = np.zeros((n, 3))
    sc.v[:, 1] = -squeeze * (sc.x[:, 1] - 0.19) / 0.15
    sc.materials = [Material.from_E_nu(10, E, nu, thermal=dict(thermal or SURROGATE_THERMAL))]
    sc.temperature = uniform(seed + 3, np.arange(n), 0, 0, T[0], T[1])
    sc.steps = steps
    sc.info.update(T=T, thermal=dict(thermal or SURROGATE_THERMAL), squeeze=squeeze)
    return sc


# Thermal-J2 steel in scaled units: E = 2e11 Pa, nu = 0.3, yield 2.5e8 Pa
# (A = 1.25e-3 E) with linear softening to 10% between 300 K and 1200 K,
# perfectly plastic (H = 0). The density is scaled to 2.5e7 kg/m^3 so that the
# P-wave CFL number c_p dt / h stays 0.27 at h = 1/128, dt = 2e-5.
STEEL_THERMAL = dict(Tref=300.0, Thot=1200.0, hotRatio=0.1, A=1.25e-3, H=0.0, kelvin_offset=0.0)
STEEL_RHO_SCALED = 2.5e7
# Guard constants of the workloads that fix dt at a P-wave CFL number of
# 0.25-0.27: the reference's Courant number 1.4/8 is below that, and its wave
# measure max(1, |F|row |F|col) grows up to 3 under rotation alone; its bound
# .08 on dt sum |C| is passed where the batched scenes hit the walls at
# 1 m/s. These scenes flag a wave that crosses a whole cell in one step and a
# step that changes F by 100% (summed over the components).
FIXED_DT_COURANT = 1.0
FIXED_DT_GRADIENT = 1.0


def thermal_boxes(cells=128, nboxes=6, ppa=4, seed=6, side=(0.05, 0.2), T=(300.0, 900.0), vsigma=0.5,
                  dt=2e-5, steps=500, rho=STEEL_RHO_SCALED, per_particle=False, squeeze=0.0):
    """Thermal-J2 workload: 6 seeded non-overlapping boxes (sides uniform in
    [0.05, 0.2]) in the unit domain at 128^3 cells, 4 particles per axis per
    cell (~2M particles), per-box velocity N(0, 0.5 m/s) per component,
    per-particle temperature uniform in [300, 900] K, gravity -9.81,
    separating walls, dt = 2e-5, 500 steps. With per_particle, every particle
    draws its own velocity N(0, vsigma) per component (seeded hash normals):
    neighbours then move apart and together; the grid transfer averages most
    of that noise away, so only a few particles yield. With squeeze > 0 every
    box also gets the vertical squeeze v_y = -squeeze (y - y_c) / (side/2)
    (a uniaxial strain rate of 2 squeeze / side, 5-20/s at 0.5 m/s): the steel
    yields within a few steps and flows plastically for the rest (the
    flowing variant)."""
    h = 1.0 / cells
    spacing = h / ppa
    boxes = []
    attempt = 0
    while len(boxes) < nboxes and attempt < 1000:
        k = len(boxes)
        s = uniform(seed, k, attempt, 0, side[0], side[1])
        c = np.array([uniform(seed, k, attempt, 1 + a, 0.1 + s / 2, 0.9 - s / 2) for a in range(3)])
        # Separated along at least one axis by a gap of two cells.
        ok = all(np.any(np.abs(c - c2) > (s + s2) / 2 + 2 * h) for c2, s2 in boxes)
        if ok:
            boxes.append((c, float(s)))
        attempt += 1
    xs, vs, ms, vols = [], [], [], []
    for k, (c, s) in enumerate(boxes):
        pts, vp = geometry(c, (s, s, s), "box", spacing, seed + k, 0.2)
        vel = np.array([vsigma * normal(seed, 10_000 + k, 0, a) for a in range(3)], float)
        v = np.array(np.broadcast_to(vel, pts.shape))
        if squeeze:
            v[:, 1] -= squeeze * (pts[:, 1] - c[1]) / (s / 2)
        xs.append(pts)
        vs.append(v)
        ms.append(np.full(len(pts), rho * vp))
        vols.append(np.full(len(pts), vp))
    x = np.concatenate(xs)
    n = len(x)
    temp = uniform(seed + 7, np.arange(n), 0, 0, T[0], T[1])
    mat = Material.from_E_nu(10, 2e11, 0.3, thermal=dict(STEEL_THERMAL))
    if per_particle:
        ids = np.arange(n)
        vs = [np.stack([vsigma * normal(seed + 11, ids, 0, a) for a in range(3)], 1)]
    name = "thermal_boxes" + ("_particle_v" if per_particle else "") + ("_squeezed" if squeeze else "")
    return Scene(name, (0, 0, 0), (1, 1, 1), h, x,
                 np.concatenate(vs), np.concatenate(ms),
                 np.concatenate(vols), temp, np.zeros(n, np.uint8), [mat], dt, steps,
                 info=dict(boxes=[(c.tolist(), s) for c, s in boxes], rho=rho, spacing=spacing),
                 courant=FIXED_DT_COURANT, gradient=FIXED_DT_GRADIENT)


def large_scene(cells=256, ppa=2, seed=11, nspheres=32, radius=(0.03, 0.08), T=(300.0, 1200.0), dt=1e-5,
                steps=1000, rho=STEEL_RHO_SCALED, block=(0.84, 0.14, 0.84), drop=1.0, gap=2.0):
    """Large single scene: 256^3 cells, 2 particles per axis per cell; a
    seeded dense thermal-J2 block (STEEL_THERMAL, material 0) on the floor
    plus `nspheres` fixed-corotated elastic spheres (material 1, radii
    uniform in `radius`, non-overlapping, at least `gap` cells above the
    block, falling at `drop` m/s) above it;
    temperature from T[0] at x = 0 to T[1] at x = 1 with seeded noise of 30 K
    (clipped to T); gravity -9.81, separating walls, dt = 1e-5, 1000 steps.
    The density is scaled as in thermal_boxes (P-wave CFL 0.27 at h = 1/256,
    dt = 1e-5). About 16M particles."""
    h = 1.0 / cells
    spacing = h / ppa
    xs, vs, ms, vols, mats = [], [], [], [], []
    bc = (0.5, 2 * h + block[1] / 2, 0.5)
    pts, vp = geometry(bc, block, "box", spacing, seed, 0.2)
    xs.append(pts)
    vs.append(np.zeros_like(pts))
    ms.append(np.full(len(pts), rho * vp))
    vols.append(np.full(len(pts), vp))
    mats.append(np.zeros(len(pts), np.uint8))
    top = 2 * h + block[1]
    spheres = []
    attempt = 0
    while len(spheres) < nspheres and attempt < 10000:
        k = len(spheres)
        r = uniform(seed, k, attempt, 0, radius[0], radius[1])
        c = np.array([uniform(seed, k, attempt, 1, 0.05 + r, 0.95 - r),
                      uniform(seed, k, attempt, 2, top + gap * h + r, 0.95 - r),
                      uniform(seed, k, attempt, 3, 0.05 + r, 0.95 - r)])
        if all(np.linalg.norm(c - c2) > r + r2 + 2 * h for c2, r2 in spheres):
            spheres.append((c, float(r)))
        attempt += 1
    for k, (c, r) in enumerate(spheres):
        pts, vp = geometry(c, (2 * r, 2 * r, 2 * r), "ellipsoid", spacing, seed + 1 + k, 0.2)
        xs.append(pts)
        vs.append(np.tile(np.array([0.0, -drop, 0.0]), (len(pts), 1)))
        ms.append(np.full(len(pts), rho * vp))
        vols.append(np.full(len(pts), vp))
        mats.append(np.ones(len(pts), np.uint8))
    x = np.concatenate(xs)
    n = len(x)
    temp = T[0] + (T[1] - T[0]) * x[:, 0] + 30.0 * (uniform(seed + 5, np.arange(n), 0, 0, -1.0, 1.0))
    temp = np.clip(temp, T[0], T[1])
    materials = [Material.from_E_nu(10, 2e11, 0.3, thermal=dict(STEEL_THERMAL)),
                 Material.from_E_nu(0, 2e11, 0.3)]
    return Scene("large_scene", (0, 0, 0), (1, 1, 1), h, x, np.concatenate(vs), np.concatenate(ms),
                 np.concatenate(vols), temp, np.concatenate(mats), materials, dt, steps,
                 info=dict(spheres=[(c.tolist(), r) for c, r in spheres], rho=rho, spacing=spacing, block=block),
                 courant=FIXED_DT_COURANT, gradient=FIXED_DT_GRADIENT)


@dataclass
class BatchScenes:
    """Many scenes for one batch: per-scene configs (mpm_config), counts, step
    sizes and particle arrays, plus the shared material table."""
    scenes: list
    materials: list
    steps: int
    info: dict = field(default_factory=dict)

    @property
    def counts(self):
        return [sc.n for sc in self.scenes]

    @property
    def total(self):
        return sum(self.counts)

    def batch(self, sort_every=16, use_graph=False):
        from .engine import Batch, make_config
        cfgs = [make_config(sc.lower, sc.upper, sc.h, sc.n, wall_mode=sc.wall_mode, gravity=sc.gravity,
                            gravity_on_particles=sc.gravity_on_particles, cfl_number=sc.courant,
                            grad_number=sc.gradient)
                for sc in self.scenes]
        b = Batch(cfgs, self.counts, sort_every=sort_every, use_graph=use_graph)
        b.set_materials(self.materials)
        for i, sc in enumerate(self.scenes):
            b.upload(i, sc.x, sc.v, sc.mass, sc.volume, temperature=sc.temperature, material=sc.material)
        b.set_dt([sc.dt for sc in self.scenes])
        return b


def batched_scenes(nscenes=256, seed=21, counts=(2_000, 200_000), grids=(32, 64, 96), steps=300, E=1e5, nu=0.3,
                   rho=1000.0, vsigma=0.5, cfl=0.25):
    """Independent scenes of uneven size: scene k has a particle count
    log-uniform in `counts`, a unit domain with a grid edge drawn from
    `grids` cells, one seeded box of particles (2 or more per axis per cell,
    side so that the count fits, centred at random), a per-scene velocity
    N(0, vsigma) per component, and material 0 (fixed-corotated) or 1 (thermal
    J2 with SURROGATE_THERMAL, temperature uniform in [300, 900] K) with equal
    odds; E = 1e5 Pa, nu = 0.3, rho = 1000, gravity -9.81, separating walls;
    dt = cfl h / c_p per scene. 300 steps."""
    materials = [Material.from_E_nu(0, E, nu), Material.from_E_nu(10, E, nu, thermal=dict(SURROGATE_THERMAL))]
    lam, mu = materials[0].lam, materials[0].mu
    cp = float(np.sqrt((lam + 2 * mu) / rho))
    out = []
    for k in range(nscenes):
        n_target = float(np.exp(uniform(seed, k, 0, 0, np.log(counts[0]), np.log(counts[1]))))
        cells = int(grids[min(int(uniform(seed, k, 0, 1, 0, len(grids))), len(grids) - 1)])
        h = 1.0 / cells
        per_axis = max(2, int(np.ceil(n_target ** (1 / 3) / (0.6 * cells))))
        spacing = h / per_axis
        side = n_target ** (1 / 3) * spacing
        c = np.array([uniform(seed, k, 0, 2 + a, 0.1 + side / 2, 0.9 - side / 2) for a in range(3)])
        pts, vp = geometry(c, (side, side, side), "box", spacing, seed + 1000 + k, 0.2)
        n = len(pts)
        mat = 1 if uniform(seed, k, 0, 5, 0, 1) < 0.5 else 0
        vel = np.array([vsigma * normal(seed, 20_000 + k, 0, a) for a in range(3)], float)
        temp = uniform(seed + 9, np.arange(n) + 1_000_000 * k, 0, 0, 300.0, 900.0) if mat == 1 else np.full(n, 300.0)
        sc = Scene(f"batch_{k}", (0, 0, 0), (1, 1, 1), h, pts, np.tile(vel, (n, 1)), np.full(n, rho * vp),
                   np.full(n, vp), temp, np.full(n, mat, np.uint8), materials, cfl * h / cp, steps,
                   info=dict(cells=cells, per_axis=per_axis, side=side, material=mat), courant=FIXED_DT_COURANT,
                   gradient=FIXED_DT_GRADIENT)
        out.append(sc)
    return BatchScenes(out, materials, steps, info=dict(seed=seed, counts=counts, grids=grids, cp=cp))
