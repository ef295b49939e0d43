"""Benchmark: particle-steps/s of the MLS-MPM step (P2G + grid + G2P) on one MI355X.

Workload: the 128^3-cell thermal-J2 scene (~2.13M particles in 6 seeded
boxes, FP32 state, dt = 2e-5), stepped K times after W warm-up steps with
the particle state resident in HBM. One JSON line on stdout.

Extra fields:
  roofline      the dominant kernel's algorithmic HBM bytes per launch over
                its average launch duration, measured here with HIP events
                on the engine's stream (profiled pass after the timed one).
  cpu_baseline  the FP64 CPU restatement of the reference step (oracle),
                single thread, on a bounded sample of the same workload.
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
    p.add_argument("--sort-every", type=int, default=int(os.environ.get("MPM_SORT_EVERY", "8")))
    p.add_argument("--cpu-steps", type=int, default=8)
    p.add_argument("--no-cpu", action="store_true")
    return p.parse_args()


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
    sc = scenes.thermal_boxes()
    eng = sc.engine(sort_every=args.sort_every)
    eng.step(sc.dt, args.warmup)
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    eng.step(sc.dt, args.steps)
    eng.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    value = sc.n * args.steps / elapsed
    cons = eng.conservation()
    assert cons["grid_mass_q"] == cons["particle_mass_q"] and cons["nonfinite"] == 0, cons

    # Profiled pass: per-kernel average duration from HIP events on the
    # engine stream.
    prof_steps = min(args.steps, 100)
    eng.profile(True, True)
    eng.step(sc.dt, prof_steps)
    eng.synchronize()
    times = eng.kernel_times()
    eng.profile(False, True)
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
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
