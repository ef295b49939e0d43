"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline leg
import this module, as the checker or as the timed CPU baseline. The product
(``mpm_amd``) never imports it.

The oracle restates Summer-Particle-Lab's ``MPM3`` (src/three/mpm.js:7-136)
and the constitutive and seeding functions it calls; see oracle/mpm_oracle.cpp
for the per-function citations.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

D = C.c_double
PD = C.POINTER(C.c_double)
PI32 = C.POINTER(C.c_int32)
PU8 = C.POINTER(C.c_uint8)


class OrcThermal(C.Structure):
    _fields_ = [(k, D) for k in ("Tref", "Thot", "hotRatio", "A", "H", "rate", "Q", "n", "stress",
                                 "kelvin_offset")]


class OrcPlastic(C.Structure):
    _fields_ = [(k, D) for k in ("A", "B", "n", "C", "rate0", "Tref", "Tm", "m")]


class OrcPressure(C.Structure):
    _fields_ = [(k, D) for k in ("cohesion", "slope", "dilatancy", "tension", "alpha", "plateau", "gamma", "densify",
                                 "beta", "epsD")]


class OrcFracture(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("onset", D), ("failure", D)]


class OrcPart(C.Structure):
    _fields_ = [("E", D), ("G", D), ("L", D), ("th", OrcThermal), ("pl", OrcPlastic), ("pr", OrcPressure),
                ("fr", OrcFracture)]


class OrcScene(C.Structure):
    _fields_ = [
        ("count", C.c_int),
        ("x", PD), ("y", PD), ("z", PD), ("vx", PD), ("vy", PD), ("vz", PD),
        ("mass", PD), ("volume", PD), ("temperature", PD), ("damage", PD),
        ("yield_strain", PD), ("friction", PD),
        ("model", PI32), ("part", PI32), ("pinned", PU8),
        ("F", PD), ("C", PD), ("Jp", PD), ("dilation", PD), ("stress", PD), ("plastic", PD),
        ("tau", PD), ("cachedTau", PD), ("cacheValid", PU8),
        ("nparts", C.c_int), ("parts", C.POINTER(OrcPart)),
        ("stiffness", D),
        ("lower", D * 3), ("upper", D * 3), ("origin", D * 3), ("H", D),
        ("N", C.c_int),
        ("gmass", PD), ("gv", PD), ("gfixed", PU8),
        ("vol_scale", D), ("grid_gravity", D * 3), ("wall_mode", C.c_int32 * 6), ("mass_eps", D),
        ("press_enabled", C.c_int32), ("press_y", D), ("press_velocity", D), ("press_impulse", D),
    ]


# Reference defaults: src/three/thermal-plasticity.js:3, metal-plasticity.js:3
THERMAL_DEFAULTS = dict(Tref=20.0, Thot=180.0, hotRatio=.08, A=.04, H=.04, rate=.0005, Q=35000.0, n=3.0,
                        stress=.04, kelvin_offset=273.15)
PLASTICITY_DEFAULTS = dict(A=.025, B=.08, n=.35, C=.025, rate0=1.0, Tref=20.0, Tm=1000.0, m=1.0)
# src/three/pressure-plasticity.js:2, src/three/fracture.js:7
PRESSURE_DEFAULTS = dict(cohesion=.025, slope=.6, dilatancy=.15, tension=.01, alpha=1.5, plateau=.015, gamma=.01,
                         densify=.02, beta=2.0, epsD=.7)
FRACTURE_DEFAULTS = dict(enabled=0, onset=.08, failure=.16)


class OrcSnapshot(C.Structure):
    _fields_ = [("count", C.c_uint32), ("render_count", C.c_uint32), ("sample_stride", C.c_uint32),
                ("solids", C.c_int64), ("liquids", C.c_int64), ("models", C.c_int64 * 12),
                ("mass", D), ("min_temperature", D), ("max_temperature", D), ("min_J", D), ("max_stress", D),
                ("max_plastic", D), ("nonfinite", C.c_int64)]


class OrcFrame(C.Structure):
    _fields_ = [("time_step_scale", D), ("max_substeps", C.c_int32), ("velocity_limit", D),
                ("bounds_lo", D * 3), ("bounds_hi", D * 3), ("anti_clump", C.c_int32), ("gravity_mode", C.c_int32),
                ("gravity_strength", D), ("gravity_direction", D * 3), ("gravity_center", D * 3),
                ("gravity_softening", D)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(_LIB_PATH)
        L.orc_random.restype = D
        L.orc_random.argtypes = [C.c_uint32, C.c_int32, C.c_int32, C.c_int32]
        L.orc_jitter_point.argtypes = [PD, C.c_int, PD, PD, C.c_int, C.c_int, C.c_int, D, C.c_uint32, D, PD]
        L.orc_domain.restype = C.c_int
        L.orc_domain.argtypes = [PD, PD, D, C.c_int, C.c_int, PD, PI32, PI32]
        L.orc_geometry_points.restype = C.c_int64
        L.orc_geometry_points.argtypes = [PD, PD, C.c_int, D, C.c_int64, PD]
        L.orc_svd3.argtypes = [PD, PD, PD, PD]
        L.orc_compose.argtypes = [PD, PD, PD, PD]
        L.orc_rotation.argtypes = [PD, PD]
        L.orc_stress3.argtypes = [PD, D, D, PD]
        L.orc_thermal_strength.restype = D
        L.orc_thermal_strength.argtypes = [D, D, C.POINTER(OrcThermal)]
        L.orc_creep_log_rate.restype = D
        L.orc_creep_log_rate.argtypes = [D, D, C.POINTER(OrcThermal)]
        L.orc_thermal_return.argtypes = [PD, C.c_int, D, D, D, C.POINTER(OrcThermal), D, PD, PD]
        L.orc_flow_stress.restype = D
        L.orc_flow_stress.argtypes = [C.c_int, D, D, D, C.POINTER(OrcPlastic)]
        L.orc_metal_return.argtypes = [PD, C.c_int, D, D, D, C.POINTER(OrcPlastic), D, PD, PD]
        L.orc_constitutive_stress.argtypes = [PD, C.c_int, D, D, D, PD]
        L.orc_principal_stress.argtypes = [PD, PD, C.c_int, D, PD]
        L.orc_von_mises.restype = D
        L.orc_von_mises.argtypes = [PD, D]
        L.orc_preview_pick.restype = C.c_uint32
        L.orc_preview_pick.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_snapshot.argtypes = [C.POINTER(OrcScene), C.c_uint32, PI32, C.POINTER(C.c_float),
                                   C.POINTER(OrcSnapshot)]
        L.orc_foam_strength.restype = D
        L.orc_foam_strength.argtypes = [D, C.POINTER(OrcPressure)]
        L.orc_pressure_return.argtypes = [PD, C.c_int, D, D, C.POINTER(OrcPressure), D, D, PD, PD, PD]
        L.orc_update_damage.argtypes = [C.POINTER(OrcScene), C.c_int]
        L.orc_project_plastic.restype = C.c_int
        L.orc_project_plastic.argtypes = [PD, C.c_int, D, D, D, D, D, D, D, C.POINTER(OrcPart),
                                          PD, PD, PD, PD, PD, PD, PD]
        for name in ("orc_update_stress",):
            getattr(L, name).argtypes = [C.POINTER(OrcScene), C.c_int]
        L.orc_refresh_diagnostics.argtypes = [C.POINTER(OrcScene)]
        L.orc_stable_step.restype = D
        L.orc_stable_step.argtypes = [C.POINTER(OrcScene), D, D]
        L.orc_apply_particle_gravity.argtypes = [C.POINTER(OrcScene), PD, D]
        L.orc_step.restype = C.c_int
        L.orc_step.argtypes = [C.POINTER(OrcScene), D]
        L.orc_step_to_grid.restype = C.c_int
        L.orc_step_to_grid.argtypes = [C.POINTER(OrcScene), D]
        L.orc_step_from_grid.restype = C.c_int
        L.orc_step_from_grid.argtypes = [C.POINTER(OrcScene), D, C.POINTER(PD)]
        L.orc_p2g.argtypes = [C.POINTER(OrcScene), D]
        L.orc_min_j.restype = D
        L.orc_min_j.argtypes = [C.POINTER(OrcScene)]
        L.orc_mechanics.argtypes = [C.POINTER(OrcScene), PD]
        L.orc_advance_frame.restype = C.c_int
        L.orc_advance_frame.argtypes = [C.c_void_p, C.POINTER(OrcFrame), D, PD]
        L.orc_regularize.argtypes = [C.c_void_p, D, PD, PD]
        L.orc_mlp_param_count.restype = C.c_int
        L.orc_mlp_forward.argtypes = [PD, C.c_int, PD, PD]
        L.orc_mlp_loss_grad.restype = D
        L.orc_mlp_loss_grad.argtypes = [PD, C.c_int, PD, PD, PD]
        _lib = L
    return _lib


def _p(a, ct=PD):
    return a.ctypes.data_as(ct)


def _f64(a, n=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    return a


# ------------------------------------------------------------------ functions
def random(seed, index, attempt, axis):
    return lib().orc_random(seed, index, attempt, axis)


def jitter_point(p, index, lo, hi, shape, fixed, enabled, strength, seed, spacing):
    out = np.zeros(3)
    p, lo, hi = _f64(p), _f64(lo), _f64(hi)
    lib().orc_jitter_point(_p(p), index, _p(lo), _p(hi), shape, int(fixed), int(enabled), strength, seed, spacing,
                           _p(out))
    return out


def domain(lower, upper, h, n=2, strict=True):
    lower, upper = _f64(lower), _f64(upper)
    origin = np.zeros(3)
    nodes = np.zeros(3, np.int32)
    cells = np.zeros(3, np.int32)
    rc = lib().orc_domain(_p(lower), _p(upper), h, n, int(strict), _p(origin), _p(nodes, PI32), _p(cells, PI32))
    if rc != 0:
        raise ValueError("invalid domain")
    return dict(origin=origin, nodes=nodes, cells=cells, cellSize=h, spacing=h / n)


def geometry_points(center, size, shape, spacing, capacity=1 << 40):
    center, size = _f64(center), _f64(size)
    n = lib().orc_geometry_points(_p(center), _p(size), shape, spacing, capacity, None)
    if n < 0:
        raise ValueError("over capacity")
    out = np.zeros((n, 3))
    lib().orc_geometry_points(_p(center), _p(size), shape, spacing, capacity, _p(out))
    return out


def svd3(F):
    F = _f64(F).reshape(9)
    u, v, s = np.zeros(9), np.zeros(9), np.zeros(3)
    lib().orc_svd3(_p(F), _p(u), _p(v), _p(s))
    return u, v, s


def compose(u, v, sigma):
    u, v, sigma = _f64(u), _f64(v), _f64(sigma)
    F = np.zeros(9)
    lib().orc_compose(_p(u), _p(v), _p(sigma), _p(F))
    return F


def rotation(F):
    F = _f64(F).reshape(9)
    R = np.zeros(9)
    lib().orc_rotation(_p(F), _p(R))
    return R


def stress3(F, mu, lam):
    F = _f64(F).reshape(9)
    t = np.zeros(9)
    lib().orc_stress3(_p(F), mu, lam, _p(t))
    return t


def thermal_params(**kw):
    p = dict(THERMAL_DEFAULTS)
    p.update(kw)
    return OrcThermal(**p)


def plastic_params(**kw):
    p = dict(PLASTICITY_DEFAULTS)
    p.update(kw)
    return OrcPlastic(**p)


def pressure_params(**kw):
    p = dict(PRESSURE_DEFAULTS)
    p.update(kw)
    return OrcPressure(**p)


def fracture_params(**kw):
    p = dict(FRACTURE_DEFAULTS)
    p.update(kw)
    p["enabled"] = int(p["enabled"])
    return OrcFracture(**p)


def foam_strength(ep, p=None):
    return lib().orc_foam_strength(ep, C.byref(p or pressure_params()))


def pressure_return(sigma, model, jp, ep, p=None, G=.4, L=.4):
    """(stretches, jp, increment) of the pressure-dependent return."""
    out, jo, inc = np.zeros(3), D(0), D(0)
    lib().orc_pressure_return(_p(_f64(sigma)), model, jp, ep, C.byref(p or pressure_params()), G, L, _p(out),
                              C.byref(jo), C.byref(inc))
    return out, jo.value, inc.value


def thermal_strength(ep, T, p=None):
    return lib().orc_thermal_strength(ep, T, C.byref(p or thermal_params()))


def creep_log_rate(q, T, p=None):
    return lib().orc_creep_log_rate(q, T, C.byref(p or thermal_params()))


def thermal_return(sigma, model, ep, dt, T, p=None, G=.4):
    sigma = _f64(sigma)
    out, inc = np.zeros(3), D(0)
    lib().orc_thermal_return(_p(sigma), model, ep, dt, T, C.byref(p or thermal_params()), G, _p(out), C.byref(inc))
    return out, inc.value


def flow_stress(model, ep, rate, T, p=None):
    return lib().orc_flow_stress(model, ep, rate, T, C.byref(p or plastic_params()))


def metal_return(sigma, model, ep, dt, T, p=None, G=.4):
    sigma = _f64(sigma)
    out, inc = np.zeros(3), D(0)
    lib().orc_metal_return(_p(sigma), model, ep, dt, T, C.byref(p or plastic_params()), G, _p(out), C.byref(inc))
    return out, inc.value


def constitutive_stress(F, model, mu, lam, jp=1.0):
    F = _f64(F).reshape(9)
    t = np.zeros(9)
    lib().orc_constitutive_stress(_p(F), model, mu, lam, jp, _p(t))
    return t


def principal_stress(u, sigma, model, lr=1.0):
    u, sigma = _f64(u), _f64(sigma)
    t = np.zeros(9)
    lib().orc_principal_stress(_p(u), _p(sigma), model, lr, _p(t))
    return t


def von_mises(t, J=1.0):
    t = _f64(t).reshape(9)
    return lib().orc_von_mises(_p(t), J)


def make_part(E=0.0, G=.4, L=.4, thermal=None, plastic=None, pressure=None, fracture=None):
    return OrcPart(E=E, G=G, L=L, th=thermal or thermal_params(), pl=plastic or plastic_params(),
                   pr=pressure or pressure_params(), fr=fracture or fracture_params())


def project_plastic(F, model, jp=1.0, dilation=0.0, yield_strain=.025, friction=35.0, ep=0.0, dt=1 / 240,
                    T=20.0, part=None):
    F = _f64(F).reshape(9)
    part = part or make_part()
    outF, u, v, sg = np.zeros(9), np.zeros(9), np.zeros(9), np.zeros(3)
    jp_o, dil_o, inc_o = D(0), D(0), D(0)
    rc = lib().orc_project_plastic(_p(F), model, jp, dilation, yield_strain, friction, ep, dt, T, C.byref(part),
                                   _p(outF), C.byref(jp_o), C.byref(dil_o), C.byref(inc_o), _p(u), _p(v), _p(sg))
    if rc != 0:
        raise ValueError(f"model {model} not restated")
    return dict(F=outF, jp=jp_o.value, dilation=dil_o.value, plasticIncrement=inc_o.value,
                decomposition=dict(u=u, v=v, sigma=sg))


# -------------------------------------------------------------------- scenes
class Scene:
    """Caller-owned FP64 scene stepped by the oracle's MPM3 restatement.

    ``vol_scale`` = 64 and ``E`` = 0 in every part reproduce the reference's
    internal-unit conventions (src/three/mpm.js:34,68); SI scenes pass the
    physical particle volume (vol_scale 1) and a fixed E per part.
    """

    def __init__(self, x, v, mass, volume, *, lower, upper, h, model=0, part=0, parts=None, temperature=20.0,
                 pinned=None, stiffness=1.0, vol_scale=64.0, grid_gravity=(0, 0, 0), wall_mode=(0,) * 6,
                 mass_eps=1e-12, strict_domain=False, ppa=2):
        x = np.asarray(x, np.float64)
        n = len(x)
        self.n = n
        self.x, self.y, self.z = [np.ascontiguousarray(x[:, a]) for a in range(3)]
        v = np.broadcast_to(np.asarray(v, np.float64), (n, 3))
        self.vx, self.vy, self.vz = [np.ascontiguousarray(v[:, a]) for a in range(3)]
        self.mass = np.ascontiguousarray(np.broadcast_to(np.asarray(mass, np.float64), (n,)))
        self.volume = np.ascontiguousarray(np.broadcast_to(np.asarray(volume, np.float64), (n,)))
        self.temperature = np.ascontiguousarray(np.broadcast_to(np.asarray(temperature, np.float64), (n,)))
        self.model = np.ascontiguousarray(np.broadcast_to(np.asarray(model, np.int32), (n,)))
        self.part = np.ascontiguousarray(np.broadcast_to(np.asarray(part, np.int32), (n,)))
        self.pinned = np.zeros(n, np.uint8) if pinned is None else np.ascontiguousarray(pinned, np.uint8)
        self.F = np.tile(np.eye(3).reshape(9), n).astype(np.float64)
        self.C = np.zeros(9 * n)
        self.Jp = np.ones(n)
        self.dilation = np.zeros(n)
        self.damage = np.zeros(n)
        self.stress = np.zeros(n)
        self.plastic = np.zeros(n)
        self.tau = np.zeros(9 * n)
        self.cachedTau = np.zeros(9 * n)
        self.cacheValid = np.zeros(n, np.uint8)
        need = int(self.part.max()) + 1 if n else 1
        if self.part.min(initial=0) < 0:
            raise ValueError("negative part index")
        parts = parts or [make_part() for _ in range(need)]
        if len(parts) < need:
            raise ValueError(f"part index {need - 1} without a part record ({len(parts)} given)")
        self._parts = (OrcPart * len(parts))(*parts)
        d = domain(lower, upper, h, ppa, strict=strict_domain)
        self.origin = d["origin"]
        self.N = int(d["nodes"].max())
        self.H = h
        nn = self.N ** 3
        self.gmass = np.zeros(nn)
        self.gv = np.zeros(3 * nn)
        self.gfixed = np.zeros(nn, np.uint8)
        s = OrcScene()
        s.count = n
        for k in ("x", "y", "z", "vx", "vy", "vz", "mass", "volume", "temperature", "F", "C", "Jp", "dilation",
                  "stress", "plastic", "tau", "cachedTau", "gmass", "gv", "damage"):
            setattr(s, k, _p(getattr(self, k)))
        s.yield_strain = None
        s.friction = None
        s.model = _p(self.model, PI32)
        s.part = _p(self.part, PI32)
        s.pinned = _p(self.pinned, PU8)
        s.cacheValid = _p(self.cacheValid, PU8)
        s.gfixed = _p(self.gfixed, PU8)
        s.nparts = len(parts)
        s.parts = self._parts
        s.stiffness = stiffness
        s.lower[:] = list(map(float, lower))
        s.upper[:] = list(map(float, upper))
        s.origin[:] = list(self.origin)
        s.H = h
        s.N = self.N
        s.vol_scale = vol_scale
        s.grid_gravity[:] = list(map(float, grid_gravity))
        s.wall_mode[:] = list(map(int, wall_mode))
        s.mass_eps = mass_eps
        self.s = s
        self.lower, self.upper = np.asarray(lower, float), np.asarray(upper, float)

    def set_press(self, y, velocity, impulse=0.0):
        """The platen of the reference's grid loop (s.press, src/three/mpm.js:97-101)."""
        self.s.press_enabled, self.s.press_y, self.s.press_velocity = 1, float(y), float(velocity)
        self.s.press_impulse = float(impulse)

    @property
    def press_impulse(self):
        return self.s.press_impulse

    def step(self, dt, particle_gravity=None):
        if particle_gravity is not None:
            g = _f64(particle_gravity)
            lib().orc_apply_particle_gravity(C.byref(self.s), _p(g), dt)
        rc = lib().orc_step(C.byref(self.s), dt)
        if rc != 0:
            raise ValueError("unsupported model or part index in scene")

    def step_to_grid(self, dt):
        """Stress, P2G and grid update; self.gmass, self.gv, self.gfixed then
        hold the grid (the externalGridContact hook's view)."""
        if lib().orc_step_to_grid(C.byref(self.s), dt) != 0:
            raise ValueError("unsupported model or part index in scene")

    def step_from_grid(self, dt, deltas=None):
        """G2P from the grid plus per-part velocity deltas ({part: [3 N^3]}),
        then the plastic return."""
        arr = None
        if deltas:
            self._deltas = {k: np.ascontiguousarray(v, np.float64).reshape(-1) for k, v in deltas.items()}
            arr = (PD * self.s.nparts)(*[_p(self._deltas[k]) if k in self._deltas else None
                                         for k in range(self.s.nparts)])
        if lib().orc_step_from_grid(C.byref(self.s), dt, arr) != 0:
            raise ValueError("unsupported model or part index in scene")

    def advance_frame(self, dt, *, time_step_scale=1.0, max_substeps=48, velocity_limit=160.0, bounds=None,
                      anti_clump=True, gravity_mode=1, gravity_strength=1.0, gravity_direction=(0, -1, 0),
                      gravity_center=(0, 0, 0), gravity_softening=8.0):
        """One Simulation3.step frame of an all-solid scene; returns (substeps,
        advanced time)."""
        f = OrcFrame()
        f.time_step_scale, f.max_substeps, f.velocity_limit = time_step_scale, max_substeps, velocity_limit
        lo, hi = bounds if bounds is not None else (self.lower, self.upper)
        f.bounds_lo[:], f.bounds_hi[:] = list(map(float, lo)), list(map(float, hi))
        f.anti_clump, f.gravity_mode, f.gravity_strength = int(anti_clump), int(gravity_mode), gravity_strength
        f.gravity_direction[:] = list(map(float, gravity_direction))
        f.gravity_center[:] = list(map(float, gravity_center))
        f.gravity_softening = gravity_softening
        adv = D(0)
        n = lib().orc_advance_frame(C.byref(self.s), C.byref(f), dt, C.byref(adv))
        if n < 0:
            raise ValueError("unsupported model or part index in scene")
        return n, adv.value

    def snapshot(self, limit=100_000, codes=None):
        """(stats dict, data [rows, 8] float32) of the service's frame pack."""
        lib().orc_refresh_diagnostics(C.byref(self.s))
        st = OrcSnapshot()
        lib().orc_snapshot(C.byref(self.s), limit, None, None, C.byref(st))
        data = np.zeros((st.render_count, 8), np.float32)
        c = None if codes is None else np.ascontiguousarray(codes, np.int32)
        lib().orc_snapshot(C.byref(self.s), limit, None if c is None else _p(c, PI32),
                           data.ctypes.data_as(C.POINTER(C.c_float)), C.byref(st))
        stats = {k: getattr(st, k) for k, _ in OrcSnapshot._fields_ if k != "models"}
        stats["models"] = list(st.models)
        return stats, data

    def p2g(self, dt):
        lib().orc_p2g(C.byref(self.s), dt)

    def stable_step(self, cap=0.0, cfl=0.0):
        return lib().orc_stable_step(C.byref(self.s), cap, cfl)

    def refresh_diagnostics(self):
        lib().orc_refresh_diagnostics(C.byref(self.s))

    def min_j(self):
        return lib().orc_min_j(C.byref(self.s))

    def mechanics(self):
        out = np.zeros(9)
        lib().orc_mechanics(C.byref(self.s), _p(out))
        return dict(mass=out[0], momentum=out[1:4], angular=out[4:7], kinetic=out[7], affine=out[8])

    def positions(self):
        return np.stack([self.x, self.y, self.z], 1)

    def velocities(self):
        return np.stack([self.vx, self.vy, self.vz], 1)


# ---------------------------------------------------------------- surrogate
def mlp_param_count():
    return lib().orc_mlp_param_count()


def mlp_forward(params, x):
    params, x = _f64(params), _f64(x)
    n = len(x)
    y = np.zeros((n, 6))
    lib().orc_mlp_forward(_p(params), n, _p(x), _p(y))
    return y


def mlp_loss_grad(params, x, t):
    params, x, t = _f64(params), _f64(x), _f64(t)
    g = np.zeros(len(params))
    loss = lib().orc_mlp_loss_grad(_p(params), len(x), _p(x), _p(t), _p(g))
    return loss, g


# ------------------------------------------------------------------ 2D solid
PF = C.POINTER(C.c_float)


class OrcScene2(C.Structure):
    _fields_ = [("count", C.c_int), ("x", PF), ("y", PF), ("vx", PF), ("vy", PF), ("mass", PF), ("volume", PF),
                ("pinned", PU8), ("F", PD), ("C", PD), ("stiffness", D), ("walls", C.c_int), ("bound", D),
                ("floor", D), ("top", D), ("active_nodes", C.c_int)]


def _lib2():
    L = lib()
    if not getattr(L, "_orc2", False):
        L.orc2_solid_stress.argtypes = [D, D, D, D, D, D, PD]
        L.orc2_stable_step.restype = D
        L.orc2_stable_step.argtypes = [C.POINTER(OrcScene2)]
        L.orc2_step.argtypes = [C.POINTER(OrcScene2), D]
        L.orc2_stats.argtypes = [C.POINTER(OrcScene2), PD]
        L.orc2_advance.restype = C.c_int
        L.orc2_advance.argtypes = [C.POINTER(OrcScene2), D, C.c_int, D, PD]
        L._orc2 = True
    return L


def solid_stress2(a, b, c, d, mu, lam):
    out = np.zeros(4)
    _lib2().orc2_solid_stress(a, b, c, d, mu, lam, _p(out))
    return out


# Wall boxes of the 2D driver: softbody demo, basin, and every other scene
# (src/solid-mpm.js:69-70).
WALLS2 = dict(softbody=(170, -145, 145), basin=(245, -175, 245), default=(555, -555, 555))


class Scene2:
    """2D scene of MPM solid particles stepped by the oracle's SolidMPM
    restatement; x, y, vx, vy, mass and volume are float32 as in the
    reference's 2D driver, F and C float64."""

    def __init__(self, x, v=(0, 0), mass=1.1, volume=1.0, pinned=None, stiffness=1.0, walls=None):
        x = np.asarray(x, np.float64).reshape(-1, 2)
        n = len(x)
        self.n = n
        f32 = lambda a: np.array(np.broadcast_to(np.asarray(a, np.float32), (n,)))  # noqa: E731  (own, writable)
        self.x, self.y = f32(x[:, 0]), f32(x[:, 1])
        v = np.broadcast_to(np.asarray(v, np.float64), (n, 2))
        self.vx, self.vy = f32(v[:, 0]), f32(v[:, 1])
        self.mass, self.volume = f32(mass), f32(volume)
        self.pinned = np.zeros(n, np.uint8) if pinned is None else np.ascontiguousarray(pinned, np.uint8)
        self.F = np.tile([1.0, 0.0, 0.0, 1.0], n)
        self.C = np.zeros(4 * n)
        s = OrcScene2()
        s.count = n
        for k in ("x", "y", "vx", "vy", "mass", "volume"):
            setattr(s, k, _p(getattr(self, k), PF))
        s.pinned = _p(self.pinned, PU8)
        s.F, s.C = _p(self.F), _p(self.C)
        s.stiffness = stiffness
        s.walls = 0 if walls is None else 1
        if walls is not None:
            s.bound, s.floor, s.top = WALLS2[walls] if isinstance(walls, str) else walls
        self.s = s

    def step(self, dt):
        _lib2().orc2_step(C.byref(self.s), dt)

    def stable_step(self):
        return _lib2().orc2_stable_step(C.byref(self.s))

    def stats(self):
        out = np.zeros(4)
        _lib2().orc2_stats(C.byref(self.s), _p(out))
        return dict(count=int(out[0]), minJ=out[1], maxJ=out[2], strain=out[3], nodes=self.s.active_nodes)

    def advance(self, dt=1 / 60, gravity=0, strength=1.0):
        adv = D(0)
        n = _lib2().orc2_advance(C.byref(self.s), dt, gravity, strength, C.byref(adv))
        return n, adv.value

    def positions(self):
        return np.stack([self.x, self.y], 1).astype(np.float64)

    def velocities(self):
        return np.stack([self.vx, self.vy], 1).astype(np.float64)
