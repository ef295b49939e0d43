"""Benchmark: particle-steps/s of the MLS-MPM step (P2G + grid + G2P) on one MI355X.

Workload: the 128^3-cell thermal-J2 scene (~2.13M particles in 6 seeded
boxes, FP32 state, dt = 2e-5), stepped K times after W warm-up steps with
the particle state resident in HBM. One JSON line on stdout.

Extra fields:
  roofline      the dominant kernel's algorithmic HBM bytes per launch over
                its average launch duration, measured here with HIP events
                on the engine's stream over the timed steps: a second engine
                built from the same scene replays the warm-up and the timed
                steps with per-launch events (events on every launch slow the
                stream, so the timed pass itself runs without them).
  cpu_baseline  the FP64 CPU restatement of the reference step (oracle),
                single thread, on a bounded sample of the same workload.
  batched       256 independent scenes (2k-200k particles, 32^3/64^3/96^3
                grids, mixed elastic / thermal-J2) in one batch: autotuned on
                a batch of its own, then 300 steps timed with graph replay and
                the same without graphs (each from the fresh state), and the
                same scenes stepped one after another by single-scene engines.
  solid2d       the 2D SolidMPM drop-in at the reference's particle cap
                (14,400 particles, frames of 1/60 s through the device frame
                loop), with the FP64 2D oracle timed on the same scene.
  surrogate     the same workload with the learned constitutive surrogate:
                network trained on the device first (timed separately), then
                K steps timed as above; its stress error on held-out trial
                states against the engine's plain FP32 return map, and the
                share of particle-steps per branch (network, closed-form
                elastic, plain return outside the envelope).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "summer-particle-lab_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0

# Algorithmic HBM bytes per launch, from the step's per-unit byte budget:
# 212 B per thermal-J2 particle-step (read x 12, v 12, C 36, F 36, m 4, V 4,
# ep 4, T 4; write x, v, C, F 96, ep 4) and 56 B per touched grid node per
# step (P2G writes 16, the grid update reads 16 and writes 12, G2P reads 12).
# A launch of the fused kernel performs the particle part of one whole step
# (G2P of step s, P2G of step s+1) and the node part of P2G (16 B) and G2P
# (12 B); the grid kernel performs the other 28 B per node.
PARTICLE_STEP_BYTES = 212
PARTICLE_BYTES = {"g2p2g": PARTICLE_STEP_BYTES}
NODE_BYTES = {"g2p2g": 28, "grid": 28}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=500)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--sort-every", type=int, default=int(os.environ.get("MPM_SORT_EVERY", "32")))
    p.add_argument("--cpu-steps", type=int, default=8)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-surrogate", action="store_true")
    p.add_argument("--no-batched", action="store_true")
    p.add_argument("--no-2d", action="store_true", help="skip the 2D SolidMPM leg")
    p.add_argument("--frames-2d", type=int, default=120)
    p.add_argument("--large", action="store_true",
                   help="also run the large single scene (256^3 cells, ~16.5M particles), plain and surrogate")
    p.add_argument("--large-steps", type=int, default=1000)
    p.add_argument("--batched-scenes", type=int, default=256)
    p.add_argument("--copy-calibration", type=int, default=0, metavar="MB",
                   help="first copy MB megabytes device to device (a kernel of known traffic for scaling "
                        "rocprofv3 FETCH_SIZE / WRITE_SIZE counters)")
    return p.parse_args()


def timed_steps(eng, sc, steps, warmup):
    import torch
    eng.step(sc.dt, warmup)
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(sc.dt, steps)
    eng.synchronize()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


def kernel_profile(eng, sc, steps):
    """Per-kernel (launches, total ms) from HIP events on the engine stream."""
    eng.profile(True, True)
    eng.step(sc.dt, steps)
    eng.synchronize()
    times = eng.kernel_times()
    eng.profile(False, True)
    return times


def replay_profile(sc, args, surrogate=None):
    """Per-kernel times over the timed steps: a fresh engine from the same
    scene (and surrogate) steps the warm-up, then the timed steps with
    per-launch events."""
    rep = sc.engine(sort_every=args.sort_every)
    if surrogate is not None:
        rep.set_surrogate(0, *surrogate)
    rep.step(sc.dt, args.warmup)
    rep.synchronize()
    times = kernel_profile(rep, sc, args.steps)
    del rep
    return times


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
    del t
    b = bs.batch(sort_every=out["tuning"]["sort_every"], use_graph=True)
    b.set_tuning(particles_per_group=out["tuning"]["particles_per_group"])
    b.step(2 * out["tuning"]["sort_every"])
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
    del b
    e = bs.batch(sort_every=out["tuning"]["sort_every"], use_graph=False)
    e.set_tuning(particles_per_group=out["tuning"]["particles_per_group"])
    e.step(2 * out["tuning"]["sort_every"])
    e.synchronize()
    t0 = time.perf_counter()
    e.step(steps)
    e.synchronize()
    el_e = time.perf_counter() - t0
    out["eager"] = dict(value=round(total * steps / el_e, 1), ms_per_step=round(el_e / steps * 1e3, 4))
    del e
    # the same scenes, one single-scene engine each, stepped one after another
    engines = [sc.engine(sort_every=out["tuning"]["sort_every"]) for sc in bs.scenes]
    for eng, sc in zip(engines, bs.scenes):
        eng.step(sc.dt, 2)
    for eng in engines:
        eng.synchronize()
    t0 = time.perf_counter()
    for eng, sc in zip(engines, bs.scenes):
        eng.step(sc.dt, steps)
    for eng in engines:
        eng.synchronize()
    el_s = time.perf_counter() - t0
    out["one_by_one"] = dict(value=round(total * steps / el_s, 1), ms_per_step=round(el_s / steps * 1e3, 4))
    out["speedup_vs_one_by_one"] = round(el_s / el, 2)
    return out


def large_leg(args):
    """The large single scene: plain step, then the surrogate on the J2 block."""
    import numpy as np
    import torch
    from mpm_amd import scenes
    from mpm_amd import surrogate as S
    t = time.perf_counter()
    sc = scenes.large_scene(steps=args.large_steps)
    gen_s = time.perf_counter() - t
    out = dict(workload="large_scene: 256^3 cells, J2 steel block + 32 fixed-corotated spheres, T 300-1200 K, "
                        f"dt 1e-5, {sc.steps} steps", particles=sc.n, generation_s=round(gen_s, 1))
    eng = sc.engine(sort_every=args.sort_every)
    el = timed_steps(eng, sc, sc.steps, 10)
    cons = eng.conservation()
    assert cons["grid_mass_q"] == cons["particle_mass_q"] and cons["nonfinite"] == 0, cons
    plain = sc.n * sc.steps / el
    out["plain"] = dict(value=round(plain, 1), ms_per_step=round(el / sc.steps * 1e3, 4), min_J=cons["min_J"],
                        max_speed=cons["max_speed"])
    del eng
    mat = sc.materials[0]
    cfg = S.TrainConfig(strain_sigma=mat.thermal["A"], dt=sc.dt, T_lo=300.0, T_hi=1200.0)
    P, norm, hist = S.train(mat, cfg)
    eng = sc.engine(sort_every=args.sort_every)
    eng.set_surrogate(0, P, norm)
    eng.step(sc.dt, 10)
    eng.surrogate_counts(reset=True)
    el = timed_steps(eng, sc, sc.steps, 0)
    counts = eng.surrogate_counts(reset=True)
    cons = eng.conservation()
    assert cons["grid_mass_q"] == cons["particle_mass_q"] and cons["nonfinite"] == 0, cons
    tot = max(sum(counts.values()), 1)
    out["surrogate"] = dict(value=round(sc.n * sc.steps / el, 1), ms_per_step=round(el / sc.steps * 1e3, 4),
                            speedup_vs_plain=round(sc.n * sc.steps / el / plain, 3),
                            branches={k: round(v / tot, 4) for k, v in counts.items()},
                            training_loss_last=float(hist[-1]))
    return out


def surrogate_leg(sc, args, plain_value):
    """Train the surrogate for the workload's material, check its stress error
    against the plain FP32 return, and time the same steps with it."""
    import numpy as np
    from mpm_amd import surrogate as S
    from mpm_amd.engine import return_map_batch
    mat = sc.materials[0]
    # Strain scale of the training distribution: the material's yield strain
    # scale A (sigma_y / E), so that the yield surface sits inside the
    # sampled range as it does for A = 0.02 with N(0, 0.02).
    cfg = S.TrainConfig(strain_sigma=mat.thermal["A"], dt=sc.dt, T_lo=300.0, T_hi=900.0)
    t = time.perf_counter()
    P, norm, hist = S.train(mat, cfg)
    train_s = time.perf_counter() - t
    F, T = held_out_states(20000, cfg.strain_sigma, cfg.T_lo, cfg.T_hi)
    _, _, tau, br = S.return_batch(mat, P, norm, F, 0.0, T, sc.dt)
    _, _, tau_p = return_map_batch(mat, F, 0.0, T, sc.dt)
    inside = br != 2
    w = np.array([1, 1, 1, 2, 2, 2], np.float64)  # symmetric 6-vectors: off-diagonals count twice
    d = (tau[inside].astype(np.float64) - tau_p[inside]) ** 2 @ w
    err = float(np.sqrt(d.sum() / ((tau_p[inside].astype(np.float64) ** 2) @ w).sum()))
    net = br == 0
    err_net = float(np.sqrt((((tau[net].astype(np.float64) - tau_p[net]) ** 2) @ w).sum()
                            / (((tau_p[net].astype(np.float64)) ** 2) @ w).sum())) if net.any() else None

    eng = sc.engine(sort_every=args.sort_every)
    eng.set_surrogate(0, P, norm)
    eng.step(sc.dt, args.warmup)
    eng.surrogate_counts(reset=True)
    elapsed = timed_steps(eng, sc, args.steps, 0)
    counts = eng.surrogate_counts(reset=True)
    cons = eng.conservation()
    assert cons["grid_mass_q"] == cons["particle_mass_q"] and cons["nonfinite"] == 0, cons
    del eng
    times = replay_profile(sc, args, surrogate=(P, norm))
    value = sc.n * args.steps / elapsed
    tot = max(sum(counts.values()), 1)
    return dict(value=round(value, 1), ms_per_step=round(elapsed / args.steps * 1e3, 4),
                speedup_vs_plain=round(value / plain_value, 3),
                branches={k: round(v / tot, 4) for k, v in counts.items()},
                stress_rel_l2_vs_plain_fp32=round(err, 5), stress_rel_l2_network_states=(
                    None if err_net is None else round(err_net, 5)),
                held_out=dict(states=len(F), network=int(net.sum()), elastic=int((br == 1).sum()),
                              fallback=int((br == 2).sum())),
                training=dict(samples=cfg.samples, epochs=cfg.epochs, batch=cfg.batch, strain_sigma=cfg.strain_sigma,
                              seconds=round(train_s, 2), loss_first=float(hist[0]), loss_last=float(hist[-1])),
                kernels={k: dict(launches=c, avg_us=round(ms / c * 1e3, 2)) for k, (c, ms) in times.items()})


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


def cpu_baseline(sc, steps):
    """Time the FP64 oracle (one thread) on `steps` steps of the same scene."""
    import numpy as np
    from oracle import binding as o
    mat = sc.materials[0]
    part = o.make_part(E=mat.E, G=mat.mu / mat.E, L=mat.lam / mat.E, thermal=o.thermal_params(**mat.thermal))
    orc = o.Scene(sc.x, sc.v, sc.mass, sc.volume, lower=sc.lower, upper=sc.upper, h=sc.h, vol_scale=1.0,
                  parts=[part], model=mat.model, temperature=sc.temperature, wall_mode=sc.wall_mode,
                  grid_gravity=sc.gravity)
    t = time.perf_counter()
    for _ in range(steps):
        orc.step(sc.dt)
    el = time.perf_counter() - t
    return dict(value=sc.n * steps / el, unit="particle-steps/s", cores=1, kind="port",
                sample=f"{steps} steps of the full workload ({sc.n} particles, 133^3 dense grid), FP64 oracle",
                seconds=round(el, 3))


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the engine runs a scene on a single MI355X; "
              "multi-GPU runs are not part of this build", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from mpm_amd import scenes

    torch.cuda.set_device(0)
    if args.copy_calibration > 0:
        a = torch.ones(args.copy_calibration * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        torch.empty_like(a).copy_(a)
        torch.cuda.synchronize()
        del a
    sc = scenes.thermal_boxes()
    eng = sc.engine(sort_every=args.sort_every)
    elapsed = timed_steps(eng, sc, args.steps, args.warmup)
    value = sc.n * args.steps / elapsed
    cons = eng.conservation()
    assert cons["grid_mass_q"] == cons["particle_mass_q"] and cons["nonfinite"] == 0, cons

    # Per-kernel average durations over the timed steps (event replay).
    times = replay_profile(sc, args)
    # Touched nodes: nodes within the 3x3x3 stencils of the particles.
    x = eng.download()["positions"]
    origin, nodes = scenes.domain(sc.lower, sc.upper, sc.h)
    N = int(nodes.max())
    b = np.clip(np.floor((x - origin) / sc.h - .5).astype(np.int64), 0, N - 3)
    touched = set()
    base = np.unique((b[:, 2] * N + b[:, 1]) * N + b[:, 0])
    off = np.array([(k * N + j) * N + i for k in range(3) for j in range(3) for i in range(3)])
    touched = np.unique((base[:, None] + off[None, :]).ravel()).size
    per_kernel = {}
    for name, (count, ms) in times.items():
        avg_us = ms / count * 1e3
        byts = None
        if name in PARTICLE_BYTES or name in NODE_BYTES:
            byts = sc.n * PARTICLE_BYTES.get(name, 0) + touched * NODE_BYTES.get(name, 0)
        per_kernel[name] = dict(launches=count, avg_us=round(avg_us, 2), bytes=byts)
    step_kernels = [k for k in per_kernel if k in PARTICLE_BYTES or k in NODE_BYTES]
    dom = max(step_kernels, key=lambda k: per_kernel[k]["avg_us"])
    d = per_kernel[dom]
    achieved = d["bytes"] / (d["avg_us"] * 1e-6) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{dom}_thermal_boxes.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get("bytes_per_launch")
    roofline = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                    frac=round(achieved / HBM_PEAK_GBS, 4), traffic=traffic, kernel=dom,
                    algorithmic_bytes_per_launch=d["bytes"], avg_launch_us=d["avg_us"])
    del eng
    sur = None if args.no_surrogate else surrogate_leg(sc, args, value)
    batched = None if args.no_batched else batched_leg(args)
    large = large_leg(args) if args.large else None
    solid2d = None if args.no_2d else solid2d_leg(args)
    cpu = None if args.no_cpu else cpu_baseline(sc, args.cpu_steps)
    out = {
        "metric": "particle-steps/s (P2G+grid+G2P), plain vs MLP-constitutive at <=1% stress err",
        "value": round(value, 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "thermal_boxes: 128^3 cells, 6 seeded boxes, thermal J2 plain return map, "
                               "dt 2e-5", "particles": sc.n, "grid_nodes": int(N) ** 3, "touched_nodes": int(touched),
                   "sort_every": args.sort_every, "constitutive": "plain"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": per_kernel,
        "surrogate": sur,
        "batched": batched,
        "large": large,
        "solid2d": solid2d,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
