"""Python host wrapper over the engine's C ABI.

``Engine`` owns one device engine (one HIP stream). It mirrors the method
surface of the reference ``MPM3`` class (src/three/mpm.js:7-136):
``configure`` -> constructor config, ``step``, ``stable_step``,
``refresh_diagnostics``, ``min_j``; state moves through ``upload`` /
``download`` of structure-of-arrays particle data.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import lib as L


@dataclass
class Material:
    """One constitutive record; strengths normalised by E (reference convention,
    src/three/thermal-plasticity.js:1-4, src/three/metal-plasticity.js:1-4)."""
    model: int = 0
    E: float = 1.0
    mu: float = 0.4
    lam: float = 0.4
    thermal: dict = field(default_factory=dict)
    plastic: dict = field(default_factory=dict)
    yield_strain: float = 0.0  # snow; 0: the reference default .025
    friction: float = 0.0      # sand, degrees; 0: the reference default 35
    pressure: dict = field(default_factory=dict)  # models 8, 9 (src/three/pressure-plasticity.js:2)
    fracture: dict = field(default_factory=dict)  # enabled, onset, failure (src/three/fracture.js:7)

    @staticmethod
    def from_E_nu(model, E, nu, **kw):
        mu = E / (2 * (1 + nu))
        lam = E * nu / ((1 + nu) * (1 - 2 * nu))
        return Material(model=model, E=E, mu=mu, lam=lam, **kw)

    def to_c(self):
        th = dict(Tref=20.0, Thot=180.0, hotRatio=.08, A=.04, H=.04, rate=.0005, Q=35000.0, n=3.0, stress=.04,
                  kelvin_offset=273.15)
        th.update(self.thermal)
        pl = dict(A=.025, B=.08, n=.35, C=.025, rate0=1.0, Tref=20.0, Tm=1000.0, m=1.0)
        pl.update(self.plastic)
        m = L.MpmMaterial()
        m.model = self.model
        m.E, m.mu, m.lambda_ = self.E, self.mu, self.lam
        m.yield_strain, m.friction = self.yield_strain, self.friction
        for k, v in th.items():
            setattr(m, "th_" + k, v)
        for k, v in pl.items():
            setattr(m, "pl_" + k, v)
        pr = dict(cohesion=.025, slope=.6, dilatancy=.15, tension=.01, alpha=1.5, plateau=.015, gamma=.01,
                  densify=.02, beta=2.0, epsD=.7)
        pr.update(self.pressure)
        for k, v in pr.items():
            setattr(m, "pr_" + k, v)
        fr = dict(enabled=False, onset=.08, failure=.16)
        fr.update(self.fracture)
        m.fracture, m.fr_onset, m.fr_failure = int(bool(fr["enabled"])), fr["onset"], fr["failure"]
        return m


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_config(lower, upper, cell_size, capacity=0, *, wall_mode=(0,) * 6, gravity=(0.0, 0.0, 0.0),
                gravity_on_particles=False, mass_eps=1e-12, sort_every=1, strict_domain=False, guard=True,
                cfl_number=0.0, cfl_cap=0.0, grad_number=0.0):
    """An mpm_config. The state guard (codes 7 and 8 at synchronize) is on by
    default with the reference's Courant number 1.4/8, gradient bound .08 and
    no cap term."""
    cfg = L.MpmConfig()
    cfg.lower[:] = list(map(float, lower))
    cfg.upper[:] = list(map(float, upper))
    cfg.cell_size = float(cell_size)
    cfg.capacity = int(capacity)
    cfg.wall_mode[:] = list(map(int, wall_mode))
    cfg.gravity[:] = list(map(float, gravity))
    cfg.gravity_on_particles = int(bool(gravity_on_particles))
    cfg.mass_eps = float(mass_eps)
    cfg.sort_every = int(sort_every)
    cfg.strict_domain = int(bool(strict_domain))
    cfg.guard_off = 0 if guard else 1
    cfg.cfl_number = float(cfl_number)
    cfg.cfl_cap = float(cfl_cap)
    cfg.grad_number = float(grad_number)
    return cfg


def particle_view(x, v, mass, volume, *, F=None, Cmat=None, temperature=None, plastic=None, material=None,
                  pinned=None, cached_tau=None, cache_valid=None, jp=None, dilation=None, damage=None,
                  dtype=np.float64):
    """An mpm_particles view of host arrays; returns (view, arrays to keep alive)."""
    x = np.asarray(x, np.float64)
    n = len(x)
    f = np.float64 if dtype == np.float64 else np.float32
    cols = lambda a: [np.ascontiguousarray(a[:, k], f) for k in range(3)]  # noqa: E731
    xs = cols(x)
    vs = cols(np.broadcast_to(np.asarray(v, np.float64), (n, 3)))
    F = np.tile(np.eye(3).reshape(9), n) if F is None else np.asarray(F).reshape(n * 9)
    Cm = np.zeros(9 * n) if Cmat is None else np.asarray(Cmat).reshape(n * 9)
    arr = lambda a, fill: np.ascontiguousarray(np.broadcast_to(np.asarray(fill if a is None else a, f), (n,)))  # noqa
    keep = dict(
        x=xs[0], y=xs[1], z=xs[2], vx=vs[0], vy=vs[1], vz=vs[2],
        F=np.ascontiguousarray(F, f), C=np.ascontiguousarray(Cm, f),
        mass=arr(mass, 0), volume=arr(volume, 0), temperature=arr(temperature, 20.0),
        plastic=arr(plastic, 0.0),
        material=np.ascontiguousarray(np.broadcast_to(np.asarray(0 if material is None else material, np.uint8),
                                                      (n,))),
        pinned=np.ascontiguousarray(np.broadcast_to(np.asarray(0 if pinned is None else pinned, np.uint8), (n,))),
    )
    if cached_tau is not None:
        keep["cached_tau"] = np.ascontiguousarray(np.asarray(cached_tau).reshape(n * 9), f)
        keep["cache_valid"] = np.ascontiguousarray(np.broadcast_to(np.asarray(cache_valid, np.uint8), (n,)))
    if jp is not None:
        keep["jp"] = arr(jp, 1.0)
    if dilation is not None:
        keep["dilation"] = arr(dilation, 0.0)
    if damage is not None:
        keep["damage"] = arr(damage, 0.0)
    p = L.MpmParticles()
    p.count = n
    p.dtype = L.MPM_F64 if f == np.float64 else L.MPM_F32
    for k, a in keep.items():
        setattr(p, k, _ptr(a))
    return p, keep


def download_view(n, dtype=np.float64):
    """An mpm_particles view of fresh host arrays for n particles."""
    f = dtype
    out = dict(x=np.zeros(n, f), y=np.zeros(n, f), z=np.zeros(n, f), vx=np.zeros(n, f), vy=np.zeros(n, f),
               vz=np.zeros(n, f), F=np.zeros(9 * n, f), C=np.zeros(9 * n, f), mass=np.zeros(n, f),
               volume=np.zeros(n, f), temperature=np.zeros(n, f), plastic=np.zeros(n, f),
               cached_tau=np.zeros(9 * n, f), material=np.zeros(n, np.uint8), pinned=np.zeros(n, np.uint8),
               cache_valid=np.zeros(n, np.uint8), jp=np.zeros(n, f), dilation=np.zeros(n, f),
               damage=np.zeros(n, f))
    p = L.MpmParticles()
    p.count = n
    p.dtype = L.MPM_F64 if f == np.float64 else L.MPM_F32
    for k, a in out.items():
        setattr(p, k, _ptr(a))
    return p, out


def finish_download(out):
    out["positions"] = np.stack([out["x"], out["y"], out["z"]], 1)
    out["velocities"] = np.stack([out["vx"], out["vy"], out["vz"]], 1)
    return out


def conservation_dict(r):
    return dict(particle_mass=r.particle_mass, grid_mass=r.grid_mass, particle_mass_q=r.particle_mass_q,
                grid_mass_q=r.grid_mass_q, mass_scale=r.mass_scale, momentum=np.array(r.momentum[:]),
                grid_momentum=np.array(r.grid_momentum[:]), angular=np.array(r.angular[:]),
                kinetic=r.kinetic, affine=r.affine, min_J=r.min_J, max_speed=r.max_speed,
                nonfinite=r.nonfinite)


class Engine:
    """One MLS-MPM engine on the current HIP device."""

    def __init__(self, lower, upper, cell_size, capacity, *, wall_mode=(0,) * 6, gravity=(0.0, 0.0, 0.0),
                 gravity_on_particles=False, mass_eps=1e-12, sort_every=1, strict_domain=False, guard=True,
                 cfl_number=0.0, cfl_cap=0.0, grad_number=0.0):
        self._lib = L.load()
        cfg = make_config(lower, upper, cell_size, capacity, wall_mode=wall_mode, gravity=gravity,
                          gravity_on_particles=gravity_on_particles, mass_eps=mass_eps, sort_every=sort_every,
                          strict_domain=strict_domain, guard=guard, cfl_number=cfl_number, cfl_cap=cfl_cap,
                          grad_number=grad_number)
        h = C.c_void_p()
        L.check(self._lib.mpm_create(C.byref(cfg), C.byref(h)))
        self.h = h
        self.config = cfg
        self.capacity = int(capacity)
        self.count = 0

    def close(self):
        if self.h:
            self._lib.mpm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        L.check(rc, self.h)

    def set_materials(self, materials):
        arr = (L.MpmMaterial * len(materials))(*[m.to_c() for m in materials])
        self._check(self._lib.mpm_set_materials(self.h, arr, len(materials)))

    def upload(self, x, v, mass, volume, **kw):
        p, keep = particle_view(x, v, mass, volume, **kw)
        self._check(self._lib.mpm_upload(self.h, C.byref(p)))
        self.count = p.count

    def download(self, dtype=np.float64):
        p, out = download_view(self.count, dtype)
        self._check(self._lib.mpm_download(self.h, C.byref(p)))
        return finish_download(out)

    def step(self, dt, nsteps=1):
        self._check(self._lib.mpm_step(self.h, float(dt), int(nsteps)))

    def synchronize(self):
        self._check(self._lib.mpm_synchronize(self.h))

    def stable_step(self, cap=0.0, cfl=0.0):
        out = C.c_double()
        self._check(self._lib.mpm_stable_dt(self.h, cap, cfl, C.byref(out)))
        return out.value

    def conservation(self):
        r = L.MpmConservation()
        self._check(self._lib.mpm_conservation_report(self.h, C.byref(r)))
        return conservation_dict(r)

    def min_j(self):
        return self.conservation()["min_J"]

    def grid_nodes(self):
        out = C.c_int32()
        self._check(self._lib.mpm_debug_p2g(self.h, 0.0, None, None, C.byref(out)))
        return out.value

    def step_to_grid(self, dt):
        """First half of a split step: stress, P2G and grid update."""
        self._check(self._lib.mpm_step_to_grid(self.h, float(dt)))

    def grid(self):
        """(mass [N^3], v [N^3, 3], fixed [N^3]) of the pending split step,
        node index (z N + y) N + x."""
        N = self.grid_nodes()
        m, v, f = np.zeros(N ** 3), np.zeros(3 * N ** 3), np.zeros(N ** 3, np.uint8)
        self._check(self._lib.mpm_grid_download(self.h, m.ctypes.data_as(C.POINTER(C.c_double)),
                                                v.ctypes.data_as(C.POINTER(C.c_double)),
                                                f.ctypes.data_as(C.POINTER(C.c_uint8))))
        return m, v.reshape(-1, 3), f

    def set_grid_velocity(self, v):
        v = np.ascontiguousarray(v, np.float64).reshape(-1)
        self._check(self._lib.mpm_grid_upload_velocity(self.h, v.ctypes.data_as(C.POINTER(C.c_double))))

    def step_from_grid(self, dt, deltas=None):
        """Second half of a split step; deltas maps a material index to a
        [N^3, 3] velocity added where that material's particles gather."""
        if not deltas:
            self._check(self._lib.mpm_step_from_grid(self.h, float(dt), None, 0, None, 0))
            return
        mats = sorted(deltas)
        slot = np.full(64, -1, np.int32)
        for k, m in enumerate(mats):
            slot[m] = k
        d = np.ascontiguousarray(np.concatenate([np.asarray(deltas[m], np.float64).reshape(-1) for m in mats]))
        self._check(self._lib.mpm_step_from_grid(self.h, float(dt), slot.ctypes.data_as(C.POINTER(C.c_int32)), 64,
                                                 d.ctypes.data_as(C.POINTER(C.c_double)), len(mats)))

    def snapshot(self, preview_limit=100_000, codes=None):
        """(totals dict, preview rows [render_count, 8] float32) of the
        resident state (mpm_snapshot)."""
        stride, rows = C.c_uint32(), C.c_uint32()
        L.check(self._lib.mpm_preview_layout(self.count, preview_limit, C.byref(stride), C.byref(rows)))
        data = np.zeros((rows.value, 8), np.float32)
        c = None if codes is None else np.ascontiguousarray(codes, np.int32)
        st = L.MpmSnapshotStats()
        self._check(self._lib.mpm_snapshot(self.h, preview_limit,
                                           None if c is None else c.ctypes.data_as(C.POINTER(C.c_int32)),
                                           0 if c is None else len(c), data.ctypes.data_as(C.POINTER(C.c_float)),
                                           C.byref(st)))
        stats = {k: getattr(st, k) for k, _ in L.MpmSnapshotStats._fields_ if k != "models"}
        stats["models"] = list(st.models)
        return stats, data

    def refresh_diagnostics(self):
        out = np.zeros(self.count)
        self._check(self._lib.mpm_refresh_diagnostics(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out

    def stress_tensors(self):
        """The Kirchhoff stress the next step applies, (count, 3, 3) (MPM3.tau)."""
        out = np.zeros(9 * self.count)
        self._check(self._lib.mpm_stress_tensors(self.h, out.ctypes.data_as(C.POINTER(C.c_double))))
        return out.reshape(self.count, 3, 3)

    def profile(self, enable=True, reset=True):
        self._check(self._lib.mpm_profile(self.h, int(enable), int(reset)))

    def kernel_times(self):
        n = C.c_int32()
        self._check(self._lib.mpm_kernel_times(self.h, None, 0, C.byref(n)))
        arr = (L.MpmKernelTime * max(n.value, 1))()
        self._check(self._lib.mpm_kernel_times(self.h, arr, n.value, C.byref(n)))
        return {a.name.decode(): (a.launches, a.total_ms) for a in arr[:n.value]}

    def debug_bin(self):
        n = self.count
        kin, kout, perm = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        P = C.POINTER(C.c_uint32)
        self._check(self._lib.mpm_debug_bin(self.h, kin.ctypes.data_as(P), kout.ctypes.data_as(P),
                                            perm.ctypes.data_as(P)))
        return kin, kout, perm

    def advance_frame(self, dt, **settings):
        """One display frame of Simulation3.step for an all-solid scene
        (mpm_advance_frame). Settings override the reference defaults by
        field name of mpm_frame_config (bounds as bounds_lo / bounds_hi).
        Returns (substeps, advanced time)."""
        c = L.MpmFrameConfig()
        self._check(self._lib.mpm_frame_defaults(self.h, C.byref(c)))
        for k, v in settings.items():
            if k in ("bounds_lo", "bounds_hi", "gravity_direction", "gravity_center"):
                getattr(c, k)[:] = list(map(float, v))
            else:
                setattr(c, k, v)
        r = L.MpmFrameReport()
        self._check(self._lib.mpm_advance_frame(self.h, C.byref(c), float(dt), C.byref(r)))
        return r.substeps, r.advanced

    def set_surrogate(self, material_index, params=None, norm=None):
        """Use a trained surrogate network for particles of one material
        (params None switches it off)."""
        if params is None:
            self._check(self._lib.mpm_set_surrogate(self.h, -1, None, None))
            return
        P = np.ascontiguousarray(np.asarray(params, np.float32).reshape(L.MLP_NPARAM))
        self._check(self._lib.mpm_set_surrogate(self.h, material_index, P.ctypes.data_as(C.POINTER(C.c_float)),
                                                C.byref(norm)))

    def surrogate_counts(self, reset=False):
        """Particle-steps since the last reset by branch: network, elastic,
        plain return."""
        u, e, f = C.c_uint64(), C.c_uint64(), C.c_uint64()
        self._check(self._lib.mpm_surrogate_counts(self.h, C.byref(u), C.byref(e), C.byref(f), int(reset)))
        return dict(network=u.value, elastic=e.value, fallback=f.value)

    def debug_p2g(self, dt):
        N = C.c_int32()
        self._check(self._lib.mpm_debug_p2g(self.h, dt, None, None, C.byref(N)))
        nn = N.value ** 3
        m, p = np.zeros(nn), np.zeros(3 * nn)
        P = C.POINTER(C.c_double)
        self._check(self._lib.mpm_debug_p2g(self.h, dt, m.ctypes.data_as(P), p.ctypes.data_as(P), C.byref(N)))
        return m, p.reshape(nn, 3), N.value


def sort_pairs(keys, key_bits=32):
    """Stable device radix sort of uint32 keys; returns (sorted keys, permutation)."""
    lib = L.load()
    keys = np.ascontiguousarray(keys, np.uint32)
    n = len(keys)
    ko, perm = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    P = C.POINTER(C.c_uint32)
    L.check(lib.mpm_sort_pairs(keys.ctypes.data_as(P), n, key_bits, ko.ctypes.data_as(P), perm.ctypes.data_as(P)))
    return ko, perm


def _f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def return_map_batch(material, F, ep, T, dt):
    """Device plain return map on a batch (mpm_return_map_batch): returns
    (F_out [n,9], ep_out [n], tau_out [n,6]) in FP32."""
    lib = L.load()
    F = np.ascontiguousarray(np.asarray(F, np.float32).reshape(-1, 9))
    n = len(F)
    ep = np.ascontiguousarray(np.broadcast_to(np.asarray(ep, np.float32), (n,)))
    T = np.ascontiguousarray(np.broadcast_to(np.asarray(T, np.float32), (n,)))
    Fo, epo, to = np.zeros((n, 9), np.float32), np.zeros(n, np.float32), np.zeros((n, 6), np.float32)
    m = material.to_c()
    L.check(lib.mpm_return_map_batch(C.byref(m), n, float(dt), _f32p(F), _f32p(ep), _f32p(T), _f32p(Fo), _f32p(epo),
                                     _f32p(to)))
    return Fo, epo, to


def return_map_state_batch(material, F, ep, T, jp, dilation, dt):
    """Device plain return with the history variables (mpm_return_map_state_batch):
    (F_out [n,9], ep_out [n], jp_out [n], dilation_out [n], tau_out [n,6])."""
    lib = L.load()
    F = np.ascontiguousarray(np.asarray(F, np.float32).reshape(-1, 9))
    n = len(F)
    col = lambda a: np.ascontiguousarray(np.broadcast_to(np.asarray(a, np.float32), (n,)))  # noqa: E731
    ep, T, jp, dilation = col(ep), col(T), col(jp), col(dilation)
    Fo, epo, jpo, dilo = np.zeros((n, 9), np.float32), np.zeros(n, np.float32), np.zeros(n, np.float32), \
        np.zeros(n, np.float32)
    to = np.zeros((n, 6), np.float32)
    m = material.to_c()
    L.check(lib.mpm_return_map_state_batch(C.byref(m), n, float(dt), _f32p(F), _f32p(ep), _f32p(T), _f32p(jp),
                                           _f32p(dilation), _f32p(Fo), _f32p(epo), _f32p(jpo), _f32p(dilo),
                                           _f32p(to)))
    return Fo, epo, jpo, dilo, to


def stress_batch(material, F):
    """Device step stress of a batch of F (mpm_stress_batch), [n,6] FP32."""
    lib = L.load()
    F = np.ascontiguousarray(np.asarray(F, np.float32).reshape(-1, 9))
    n = len(F)
    to = np.zeros((n, 6), np.float32)
    m = material.to_c()
    L.check(lib.mpm_stress_batch(C.byref(m), n, _f32p(F), _f32p(to)))
    return to


def sym6_to_full(t):
    """(xx, yy, zz, xy, xz, yz) -> row-major 3x3."""
    t = np.asarray(t)
    return np.stack([t[..., 0], t[..., 3], t[..., 4], t[..., 3], t[..., 1], t[..., 5], t[..., 4], t[..., 5],
                     t[..., 2]], -1)


class Batch:
    """Many independent scenes on one GPU (mpm_batch_*): scene i has its own
    domain, particle count and step size; all scenes share one material table
    and step together."""

    def __init__(self, configs, counts, *, sort_every=16, use_graph=False, particle_blocks=0):
        self._lib = L.load()
        cfgs = (L.MpmConfig * len(configs))(*configs)
        cnt = np.ascontiguousarray(counts, np.uint32)
        opt = L.MpmBatchOptions(int(sort_every), int(bool(use_graph)), int(particle_blocks))
        h = C.c_void_p()
        L.check(self._lib.mpm_batch_create(cfgs, cnt.ctypes.data_as(C.POINTER(C.c_uint32)), len(configs),
                                           C.byref(opt), C.byref(h)))
        self.h = h
        self.counts = cnt
        self.scenes = len(configs)

    def debug_bin(self):
        """The batch's binning (mpm_batch_debug_bin) over the whole pool:
        keys by particle id, sorted keys, ids in sorted order."""
        n = int(self.counts.sum())
        kin, kout, perm = np.zeros(n, np.uint32), np.zeros(n, np.uint32), np.zeros(n, np.uint32)
        P = C.POINTER(C.c_uint32)
        self._check(self._lib.mpm_batch_debug_bin(self.h, kin.ctypes.data_as(P), kout.ctypes.data_as(P),
                                                  perm.ctypes.data_as(P)))
        return kin, kout, perm

    def close(self):
        if self.h:
            self._lib.mpm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        L.check(rc, batch=self.h)

    def set_materials(self, materials):
        arr = (L.MpmMaterial * len(materials))(*[m.to_c() for m in materials])
        self._check(self._lib.mpm_batch_set_materials(self.h, arr, len(materials)))

    def upload(self, scene, x, v, mass, volume, **kw):
        p, keep = particle_view(x, v, mass, volume, **kw)
        self._check(self._lib.mpm_batch_upload(self.h, int(scene), C.byref(p)))

    def download(self, scene, dtype=np.float64):
        p, out = download_view(int(self.counts[scene]), dtype)
        self._check(self._lib.mpm_batch_download(self.h, int(scene), C.byref(p)))
        return finish_download(out)

    def set_dt(self, dt):
        d = np.ascontiguousarray(np.broadcast_to(np.asarray(dt, np.float64), (self.scenes,)))
        self._check(self._lib.mpm_batch_set_dt(self.h, d.ctypes.data_as(C.POINTER(C.c_double))))

    def step(self, nsteps=1):
        self._check(self._lib.mpm_batch_step(self.h, int(nsteps)))

    def synchronize(self):
        self._check(self._lib.mpm_batch_synchronize(self.h))

    def conservation(self, scene):
        r = L.MpmConservation()
        self._check(self._lib.mpm_batch_conservation(self.h, int(scene), C.byref(r)))
        return conservation_dict(r)

    @staticmethod
    def _tuning_dict(t):
        return dict(buckets=t.buckets, bucket_edge=list(t.bucket_edge)[:t.buckets - 1],
                    bucket_scenes=list(t.bucket_scenes)[:t.buckets],
                    particles_per_group=list(t.particles_per_group)[:t.buckets], sort_every=t.sort_every,
                    particle_blocks=t.particle_blocks, ms_per_step=t.ms_per_step,
                    bucket_ms=list(t.bucket_ms)[:t.buckets])

    def tuning(self):
        t = L.MpmBatchTuning()
        self._check(self._lib.mpm_batch_get_tuning(self.h, C.byref(t)))
        return self._tuning_dict(t)

    def set_tuning(self, particles_per_group=None, sort_every=None, particle_blocks=None):
        t = L.MpmBatchTuning()
        self._check(self._lib.mpm_batch_get_tuning(self.h, C.byref(t)))
        if particles_per_group is not None:
            for k, v in enumerate(particles_per_group):
                t.particles_per_group[k] = int(v)
        if sort_every is not None:
            t.sort_every = int(sort_every)
        if particle_blocks is not None:
            t.particle_blocks = int(particle_blocks)
        self._check(self._lib.mpm_batch_set_tuning(self.h, C.byref(t)))

    def tune(self, trial_steps=8):
        t = L.MpmBatchTuning()
        self._check(self._lib.mpm_batch_tune(self.h, int(trial_steps), C.byref(t)))
        return self._tuning_dict(t)
