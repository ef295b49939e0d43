"""Benchmark: particle-steps/s of the MLS-MPM step (P2G + grid + G2P) on one MI355X.

Headline workload (the line's `value`): the large single scene, 256^3 cells,
16.5M particles (a seeded thermal-J2 steel block plus 32 falling
fixed-corotated spheres, T 300-1200 K, dt 1e-5, FP32 state, FP64
conservation sums), plain constitutive update, stepped K times after W
warm-up steps with the particle state resident in HBM. The re-bin period is
the largest divisor of K not above 64, so the K timed steps hold exactly K/P
re-bins. One JSON line on stdout.

Fields besides the contract's:
  roofline      the dominant kernel's algorithmic HBM bytes per launch over
                its average launch duration, measured here with HIP events
                stamped by the launches on the engine's stream over the timed
                steps: a second engine built from the same scene replays the
                warm-up and the timed steps with per-launch events (events on
                every launch slow the stream, so the timed pass itself runs
                without them). `traffic` is the rocprofv3 FETCH_SIZE +
                WRITE_SIZE per launch of that kernel on this workload, read
                from the committed profile (null when there is none).
  cpu_baseline  the FP64 CPU restatement of the reference step (oracle),
                single thread, on a bounded sample of the same workload, with
                the host's CPU model and core count.
  surrogate     the headline workload with the learned constitutive surrogate
                on the steel block (network trained on the device first,
                timed separately; every yielding trial state inside the
                training envelope takes the network), K steps timed as above,
                the branch mix, and the stress error on held-out trial states.
  workloads     the 128^3 thermal-boxes scene (2.1M particles, dt 2e-5) with
                its own kernel roofline, and its two yielding variants, each
                plain and with the surrogate: per-particle velocities N(0, 0.5)
                (the benchmark configuration's wording) and per-box squeezes
                (plastic flow on every step), the latter also with an
                all-elastic twin (yield strength out of reach) whose fused-kernel
                time gives the cost of the plastic return.
  batched       256 independent scenes (2k-200k particles, 32^3/64^3/96^3
                grids, mixed elastic / thermal-J2) in one batch: autotuned on
                a batch of its own, then timed with graph replay, without
                graphs, and the same scenes one after another.
  solid2d       the 2D SolidMPM drop-in at the reference's particle cap, with
                the FP64 2D oracle timed on the same scene.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "summer-particle-lab_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes, from the step's per-unit byte budget: a particle-step
# reads x 12, v 12, C 36, F 36, m 4, V 4 and writes x, v, C, F (96): 200 B for
# an elastic particle; a thermal-J2 particle adds ep read and write (8) and T
# (4): 212 B; a scene of several materials adds the material id (1 B). A
# touched grid node costs 56 B per step (P2G writes 16, the grid update reads
# 16 and writes 12, G2P reads 12). A launch of the fused kernel performs the
# particle part of one whole step (G2P of step s, P2G of step s+1) and the
# node part of P2G (16 B) and G2P (12 B); the grid kernel performs the other
# 28 B per node.
ELASTIC_BYTES, J2_BYTES, MATERIAL_ID_BYTES = 200, 212, 1
NODE_BYTES = {"g2p2g": 28, "grid": 28}
FUSED = ("g2p2g",)


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1000)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--sort-every", type=int, default=64, help="upper bound of the re-bin period (50 for 1000 steps: measured best on the large scene and the thermal boxes)")
    p.add_argument("--workload", default="large_scene",
                   choices=["large_scene", "thermal_boxes", "thermal_boxes_particle_v", "thermal_boxes_squeezed"],
                   help="headline workload")
    p.add_argument("--headline-only", action="store_true", help="only the headline leg (for profiling)")
    p.add_argument("--cpu-steps", type=int, default=2)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-surrogate", action="store_true")
    p.add_argument("--no-workloads", action="store_true")
    p.add_argument("--workload-steps", type=int, default=500)
    p.add_argument("--no-batched", action="store_true")
    p.add_argument("--batched-scenes", type=int, default=256)
    p.add_argument("--batched-only", action="store_true",
                   help="only the batched leg, without the one-by-one pass (for profiling)")
    p.add_argument("--no-2d", action="store_true", help="skip the 2D SolidMPM leg")
    p.add_argument("--frames-2d", type=int, default=120)
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                   help="first copy MB megabytes device to device (a kernel of known traffic for scaling "
                        "rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def period_for(steps, bound):
    """The largest divisor of `steps` not above `bound`."""
    for p in range(min(bound, steps), 0, -1):
        if steps % p == 0:
            return p
    return 1


def make_scene(name, steps):
    from mpm_amd import scenes
    if name == "large_scene":
        return scenes.large_scene(steps=steps)
    if name == "thermal_boxes":
        return scenes.thermal_boxes(steps=steps)
    if name == "thermal_boxes_particle_v":
        return scenes.thermal_boxes(steps=steps, per_particle=True)
    return scenes.thermal_boxes(steps=steps, squeeze=0.5)


WORKLOAD_TEXT = {
    "large_scene": "large_scene: 256^3 cells, thermal-J2 steel block + 32 falling fixed-corotated spheres, "
                   "T 300-1200 K, dt 1e-5",
    "thermal_boxes": "thermal_boxes: 128^3 cells, 6 seeded boxes, thermal J2, per-box velocities N(0, 0.5), "
                     "dt 2e-5",
    "thermal_boxes_particle_v": "thermal_boxes with per-particle velocities N(0, 0.5 m/s), dt 2e-5",
    "thermal_boxes_squeezed": "thermal_boxes with per-box vertical squeezes of 0.5 m/s (plastic flow), dt 2e-5",
}


def timed_steps(eng, sc, steps, warmup):
    import torch
    if warmup:
        eng.step(sc.dt, warmup)
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(sc.dt, steps)
    eng.synchronize()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def replay_profile(sc, steps, warmup, period, surrogate=None):
    """Per-kernel (launches, total ms) over the timed steps: a fresh engine
    from the same scene (and surrogate) steps the warm-up, then the timed
    steps with per-launch events on its stream."""
    rep = sc.engine(sort_every=period)
    if surrogate is not None:
        rep.set_surrogate(0, *surrogate)
    rep.step(sc.dt, warmup)
    rep.synchronize()
    rep.profile(True, True)
    rep.step(sc.dt, steps)
    rep.synchronize()
    times = rep.kernel_times()
    rep.close()
    return times


def touched_nodes(x, sc):
    """Grid nodes inside the 3x3x3 stencils of the particles at positions x."""
    import numpy as np
    from mpm_amd import scenes
    origin, nodes = scenes.domain(sc.lower, sc.upper, sc.h)
    N = int(nodes.max())
    b = np.clip(np.floor((x - origin) / sc.h - .5).astype(np.int64), 0, N - 3)
    base = np.unique((b[:, 2] * N + b[:, 1]) * N + b[:, 0])
    mark = np.zeros(N ** 3, bool)
    for k in range(3):
        for j in range(3):
            for i in range(3):
                mark[base + ((k * N + j) * N + i)] = True
    return int(mark.sum()), N


def particle_bytes(sc):
    """Algorithmic particle bytes of one step of the scene."""
    import numpy as np
    models = np.array([m.model for m in sc.materials])[sc.material]
    per = np.where(models >= 2, J2_BYTES, ELASTIC_BYTES).astype(np.int64)
    if len(sc.materials) > 1:
        per += MATERIAL_ID_BYTES
    return int(per.sum())


def kernel_table(times, sc, touched):
    pbytes = particle_bytes(sc)
    out = {}
    for name, (count, ms) in times.items():
        byts = None
        if name in FUSED or name in NODE_BYTES:
            byts = (pbytes if name in FUSED else 0) + touched * NODE_BYTES.get(name, 0)
        out[name] = dict(launches=count, avg_us=round(ms / count * 1e3, 2), bytes=byts)
    return out


def roofline_of(kernels, workload):
    cands = [k for k in kernels if kernels[k]["bytes"] is not None]
    dom = max(cands, key=lambda k: kernels[k]["avg_us"] * kernels[k]["launches"])
    d = kernels[dom]
    achieved = d["bytes"] / (d["avg_us"] * 1e-6) / 1e9
    traffic, tsrc = None, None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom}_{workload}.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")
        tsrc = os.path.relpath(tpath, ROOT)
    return dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, traffic_source=tsrc, kernel=dom,
                algorithmic_bytes_per_launch=d["bytes"], avg_launch_us=d["avg_us"], launches=d["launches"])


def run_plain(sc, steps, warmup, period, workload):
    """Timed plain leg plus its per-kernel replay and roofline."""
    import numpy as np
    eng = sc.engine(sort_every=period)
    elapsed = timed_steps(eng, sc, steps, warmup)
    cons = eng.conservation()
    assert cons["grid_mass_q"] == cons["particle_mass_q"] and cons["nonfinite"] == 0, cons
    x = eng.download(np.float32)["positions"]
    eng.close()
    times = replay_profile(sc, steps, warmup, period)
    touched, N = touched_nodes(x, sc)
    kernels = kernel_table(times, sc, touched)
    return dict(value=sc.n * steps / elapsed, elapsed=elapsed, ms_per_step=elapsed / steps * 1e3,
                kernels=kernels, roofline=roofline_of(kernels, workload), touched=touched, N=N,
                min_J=cons["min_J"], max_speed=cons["max_speed"])


def held_out_states(n, sigma, T_lo, T_hi, seed=2024):
    """Trial states F = exp(eps) R: eps components N(0, sigma), R a random
    rotation, T uniform."""
    import numpy as np
    rng = np.random.default_rng(seed)
    e = rng.normal(0, sigma, (n, 6))
    E = np.stack([e[:, 0], e[:, 3], e[:, 4], e[:, 3], e[:, 1], e[:, 5], e[:, 4], e[:, 5], e[:, 2]], 1)
    lam, V = np.linalg.eigh(E.reshape(n, 3, 3))
    S = np.einsum("nij,nj,nkj->nik", V, np.exp(lam), V)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y),
                  2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x),
                  2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], 1).reshape(n, 3, 3)
    return (S @ R).reshape(n, 9), rng.uniform(T_lo, T_hi, n)


_NETWORKS = {}


def train_surrogate(sc, T_hi):
    """Network for the scene's steel (material 0): training on the device
    (once per material and temperature range), min_yield_ratio 1 (every
    yielding trial state inside the envelope takes the network), and its
    stress error on held-out trial states against the plain FP32 return."""
    mat = sc.materials[0]
    key = (repr(mat), T_hi, sc.dt)
    if key not in _NETWORKS:
        _NETWORKS[key] = _train_surrogate(sc, T_hi)
    return _NETWORKS[key]


def _train_surrogate(sc, T_hi):
    import numpy as np
    from mpm_amd import surrogate as S
    from mpm_amd.engine import return_map_batch
    mat = sc.materials[0]
    # Strain scale of the training distribution: the material's yield strain
    # scale A (sigma_y / E), so the yield surface sits inside the sampled range.
    cfg = S.TrainConfig(strain_sigma=mat.thermal["A"], dt=sc.dt, T_lo=300.0, T_hi=T_hi)
    t = time.perf_counter()
    P, norm, hist = S.train(mat, cfg)
    train_s = time.perf_counter() - t
    norm.min_yield_ratio = 1.0
    F, T = held_out_states(20000, cfg.strain_sigma, cfg.T_lo, cfg.T_hi)
    _, _, tau, br = S.return_batch(mat, P, norm, F, 0.0, T, sc.dt)
    _, _, tau_p = return_map_batch(mat, F, 0.0, T, sc.dt)
    w = np.array([1, 1, 1, 2, 2, 2], np.float64)  # symmetric 6-vectors: off-diagonals count twice

    def rel(mask):
        if not mask.any():
            return None
        d = ((tau[mask].astype(np.float64) - tau_p[mask]) ** 2) @ w
        return round(float(np.sqrt(d.sum() / ((tau_p[mask].astype(np.float64) ** 2) @ w).sum())), 5)

    info = dict(stress_rel_l2_vs_plain_fp32=rel(br != 2), stress_rel_l2_network_states=rel(br == 0),
                held_out=dict(states=len(F), network=int((br == 0).sum()), elastic=int((br == 1).sum()),
                              fallback=int((br == 2).sum())),
                training=dict(samples=cfg.samples, epochs=cfg.epochs, batch=cfg.batch, strain_sigma=cfg.strain_sigma,
                              seconds=round(train_s, 2), loss_first=float(hist[0]), loss_last=float(hist[-1])),
                min_yield_ratio=1.0)
    return (P, norm), info


def run_surrogate(sc, steps, warmup, period, net, plain_value, profile=True):
    eng = sc.engine(sort_every=period)
    eng.set_surrogate(0, *net)
    if warmup:
        eng.step(sc.dt, warmup)
    eng.surrogate_counts(reset=True)
    elapsed = timed_steps(eng, sc, steps, 0)
    counts = eng.surrogate_counts(reset=True)
    cons = eng.conservation()
    assert cons["grid_mass_q"] == cons["particle_mass_q"] and cons["nonfinite"] == 0, cons
    eng.close()
    value = sc.n * steps / elapsed
    tot = max(sum(counts.values()), 1)
    out = dict(value=round(value, 1), ms_per_step=round(elapsed / steps * 1e3, 4),
               speedup_vs_plain=round(value / plain_value, 3),
               branches={k: round(v / tot, 4) for k, v in counts.items()},
               branch_particle_steps=counts)
    if profile:
        times = replay_profile(sc, steps, warmup, period, surrogate=net)
        out["kernels"] = {k: dict(launches=c, avg_us=round(ms / c * 1e3, 2)) for k, (c, ms) in times.items()}
    return out


def workload_leg(name, args):
    """A thermal-boxes workload: plain (with its kernel roofline), with the
    surrogate, and for the squeezed variant an all-elastic twin."""
    import copy

    import numpy as np
    steps = args.workload_steps
    period = period_for(steps, args.sort_every)
    sc = make_scene(name, steps)
    plain = run_plain(sc, steps, 20, period, name)
    out = dict(workload=WORKLOAD_TEXT[name], particles=sc.n, steps=steps, sort_every=period,
               value=round(plain["value"], 1), unit="particle-steps/s", ms_per_step=round(plain["ms_per_step"], 4),
               roofline=plain["roofline"], kernels=plain["kernels"], touched_nodes=plain["touched"])
    eng = sc.engine(sort_every=period)
    eng.step(sc.dt, 20 + steps)
    st = eng.download(np.float32)
    eng.close()
    out["plastic"] = dict(share_yielded=round(float((st["plastic"] > 0).mean()), 4),
                          max=float(st["plastic"].max()), mean=float(st["plastic"].mean()))
    if not args.no_surrogate:
        net, info = train_surrogate(sc, 900.0)
        out["surrogate"] = run_surrogate(sc, steps, 20, period, net, plain["value"])
        out["surrogate"].update(info)
    if name == "thermal_boxes_squeezed":
        # The same scene with the yield strength out of reach: every particle
        # takes the closed-form elastic branch, so the difference of the fused
        # kernel's time is the cost of the plastic return on this scene.
        twin = copy.copy(sc)
        mat = copy.deepcopy(sc.materials[0])
        mat.thermal = dict(mat.thermal, A=1e3)
        twin.materials = [mat]
        el = run_plain(twin, steps, 20, period, name + "_elastic")
        g_pl, g_el = plain["kernels"]["g2p2g"]["avg_us"], el["kernels"]["g2p2g"]["avg_us"]
        out["elastic_twin"] = dict(value=round(el["value"], 1), ms_per_step=round(el["ms_per_step"], 4),
                                   g2p2g_avg_us=g_el, plastic_return_share_of_g2p2g=round((g_pl - g_el) / g_pl, 4))
    return out


def batched_leg(args):
    """256 scenes in one batch vs the same scenes one by one."""
    import torch
    from mpm_amd import scenes
    bs = scenes.batched_scenes(nscenes=args.batched_scenes)
    steps = bs.steps
    total = bs.total
    out = dict(workload=f"{len(bs.scenes)} scenes, counts log-uniform 2k-200k, grids 32^3/64^3/96^3 cells, "
                        f"mixed fixed-corotated / thermal J2, {steps} steps", scenes=len(bs.scenes), particles=total)
    # The autotuner advances the state it times on, so it runs on a batch of
    # its own; the graph and eager passes start from the same fresh state.
    t = bs.batch(sort_every=16, use_graph=False)
    out["tuning"] = t.tune(trial_steps=8)
    t.close()
    # Every pass runs the scenes' 300 steps: two re-bin periods of warm-up
    # (graph capture), the rest timed. Past about 310 steps a soft thermal-J2
    # block crushed on the floor (scene 120) outgrows the fixed dt and the
    # guard reports it, as the reference's stable step would.
    warm = 2 * out["tuning"]["sort_every"]
    steps = bs.steps - warm
    b = bs.batch(sort_every=out["tuning"]["sort_every"], use_graph=True)
    b.set_tuning(particles_per_group=out["tuning"]["particles_per_group"])
    b.step(warm)
    b.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    b.synchronize()
    el = time.perf_counter() - t0
    for i in range(len(bs.scenes)):
        c = b.conservation(i)
        assert c["grid_mass_q"] == c["particle_mass_q"] and c["nonfinite"] == 0, (i, c)
    out.update(value=round(total * steps / el, 1), unit="particle-steps/s", ms_per_step=round(el / steps * 1e3, 4))
    b.close()
    e = bs.batch(sort_every=out["tuning"]["sort_every"], use_graph=False)
    e.set_tuning(particles_per_group=out["tuning"]["particles_per_group"])
    e.step(warm)
    e.synchronize()
    t0 = time.perf_counter()
    e.step(steps)
    e.synchronize()
    el_e = time.perf_counter() - t0
    out["eager"] = dict(value=round(total * steps / el_e, 1), ms_per_step=round(el_e / steps * 1e3, 4))
    e.close()
    if args.batched_only:
        return out
    # the same scenes, one single-scene engine each, stepped one after another
    engines = [sc.engine(sort_every=out["tuning"]["sort_every"]) for sc in bs.scenes]
    for eng, sc in zip(engines, bs.scenes):
        eng.step(sc.dt, warm)
    for eng in engines:
        eng.synchronize()
    t0 = time.perf_counter()
    for eng, sc in zip(engines, bs.scenes):
        eng.step(sc.dt, steps)
    for eng in engines:
        eng.synchronize()
    el_s = time.perf_counter() - t0
    for eng in engines:
        eng.close()
    out["one_by_one"] = dict(value=round(total * steps / el_s, 1), ms_per_step=round(el_s / steps * 1e3, 4))
    out["speedup_vs_one_by_one"] = round(el_s / el, 2)
    return out


def solid2d_leg(args):
    """The 2D SolidMPM drop-in (src/solid-mpm.js) on a scene at the reference's
    particle cap: four 216 x 216 panels at the reference's spacing 3.6
    (14,400 particles; the 2D driver holds at most 16,000,
    src/simulation.js:7), centre gravity, the default walls. Frames of 1/60 s
    through the device frame loop (stable substeps as the driver takes them),
    the FP64 2D oracle timed on the first frames beside it."""
    import numpy as np
    from mpm_amd.solid2d import Solid2D, panel, settings
    pts = [panel(x0, y0, x0 + 216, y0 + 216)[0] for x0 in (-420, 60) for y0 in (-420, 60)]
    x = np.concatenate(pts)
    n = len(x)
    s = settings()
    e = Solid2D(n)
    e.upload(x)
    e.advance(1 / 60, "center", s=s)
    e.stats()
    t0 = time.perf_counter()
    substeps = 0
    frames = args.frames_2d
    for _ in range(frames):
        k, _ = e.advance(1 / 60, "center", s=s)
        substeps += k
    st = e.stats()  # reads back: the frames are done
    el = time.perf_counter() - t0
    assert st["grid_mass_q"] == st["particle_mass_q"] and st["min_J"] > 0, st
    cpu = None
    if not args.no_cpu:
        from oracle import binding as o
        sc = o.Scene2(x, walls="default")
        sc.advance(1 / 60, gravity=2)
        cpu_frames, cpu_sub = 3, 0
        t1 = time.perf_counter()
        for _ in range(cpu_frames):
            k, _ = sc.advance(1 / 60, gravity=2)
            cpu_sub += k
        el_cpu = time.perf_counter() - t1
        cpu = dict(value=round(n * cpu_sub / el_cpu, 1), unit="particle-substeps/s", cores=1, kind="port",
                   sample=f"{cpu_frames} frames ({cpu_sub} substeps) of the same scene, FP64 2D oracle",
                   seconds=round(el_cpu, 3))
    return dict(workload=f"2D SolidMPM: 4 panels, {n} particles at spacing 3.6, 149^2 grid, centre gravity, "
                         f"default walls, {frames} frames of 1/60 s", particles=n, frames=frames,
                substeps=substeps, value=round(n * substeps / el, 1), unit="particle-substeps/s",
                ms_per_frame=round(el / frames * 1e3, 4), us_per_substep=round(el / max(substeps, 1) * 1e6, 2),
                cpu_baseline=cpu)


def host_cpu():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = int(subprocess.run(["nproc"], capture_output=True, text=True).stdout.strip())
    except (OSError, ValueError):
        nproc = os.cpu_count()
    return model, nproc


def cpu_baseline(sc, steps):
    """Time the FP64 oracle (one thread) on `steps` steps of the same scene."""
    import numpy as np
    from oracle import binding as o
    parts = []
    for m in sc.materials:
        th = o.thermal_params(**m.thermal) if m.thermal else None
        parts.append(o.make_part(E=m.E, G=m.mu / m.E, L=m.lam / m.E, thermal=th))
    models = np.array([m.model for m in sc.materials], np.int32)[sc.material]
    orc = o.Scene(sc.x, sc.v, sc.mass, sc.volume, lower=sc.lower, upper=sc.upper, h=sc.h, vol_scale=1.0,
                  parts=parts, model=models, part=sc.material.astype(np.int32), temperature=sc.temperature,
                  wall_mode=sc.wall_mode, grid_gravity=sc.gravity)
    t = time.perf_counter()
    for _ in range(steps):
        orc.step(sc.dt)
    el = time.perf_counter() - t
    model, nproc = host_cpu()
    N = orc.N
    return dict(value=sc.n * steps / el, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{steps} steps of the full workload ({sc.n} particles, {N}^3 dense grid), FP64 oracle "
                       "(CPU restatement of the reference's JS step), one thread",
                seconds=round(el, 3), host_cpu=model, host_nproc=nproc)


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine runs a scene on a single MI355X; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    import torch

    torch.cuda.set_device(0)
    if args.copy_calibration > 0:
        a = torch.ones(args.copy_calibration * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()
        del a
    if args.headline_only:
        args.no_cpu = args.no_surrogate = args.no_workloads = args.no_batched = args.no_2d = True
    if args.batched_only:
        batched = batched_leg(args)
        log(f"batched: {batched['ms_per_step']} ms/step")
        print(json.dumps({"batched": batched}))
        return
    K, W = args.steps, args.warmup
    period = period_for(K, args.sort_every)
    t = time.perf_counter()
    sc = make_scene(args.workload, K)
    log(f"headline {args.workload}: {sc.n} particles, generated in {time.perf_counter() - t:.1f} s")
    plain = run_plain(sc, K, W, period, args.workload)
    log(f"headline plain: {plain['ms_per_step']:.4f} ms/step, {plain['value']:.4e} particle-steps/s")
    sur = None
    if not args.no_surrogate:
        net, info = train_surrogate(sc, 1200.0 if args.workload == "large_scene" else 900.0)
        sur = run_surrogate(sc, K, W, period, net, plain["value"])
        sur.update(info)
        log(f"headline surrogate: {sur['ms_per_step']} ms/step, branches {sur['branches']}")
    workloads = None
    if not args.no_workloads:
        workloads = {}
        for name in ("thermal_boxes", "thermal_boxes_particle_v", "thermal_boxes_squeezed"):
            if name != args.workload:
                workloads[name] = workload_leg(name, args)
                log(f"{name}: {workloads[name]['ms_per_step']} ms/step")
    batched = None
    if not args.no_batched:
        batched = batched_leg(args)
        log(f"batched: {batched['ms_per_step']} ms/step")
    solid2d = None if args.no_2d else solid2d_leg(args)
    cpu = None if args.no_cpu else cpu_baseline(sc, args.cpu_steps)
    config = {"workload": WORKLOAD_TEXT[args.workload] + f", {K} steps", "particles": sc.n,
              "grid_nodes": int(plain["N"]) ** 3, "touched_nodes": plain["touched"], "sort_every": period,
              "constitutive": "plain", "materials": [m.model for m in sc.materials]}
    out = {
        "metric": "particle-steps/s (P2G+grid+G2P), plain vs MLP-constitutive at <=1% stress err",
        "value": round(plain["value"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(plain["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": config,
        "roofline": plain["roofline"],
        "cpu_baseline": cpu,
        "kernels": plain["kernels"],
        "min_J": plain["min_J"],
        "surrogate": sur,
        "workloads": workloads,
        "batched": batched,
        "solid2d": solid2d,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
