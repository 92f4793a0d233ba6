"""Benchmark of the ODE 0.12 step path on one MI355X.

Workload: 4096 independent worlds of uneven size (body count per world
clamp(Poisson(64), 8, 256) boxes/spheres plus 0-16 ball/hinge chains of 4-12
links), h = 0.01, 20 SOR iterations, synthetic seeded scenes.  One step =
broadphase + narrowphase + quickstep (rows, SOR-LCP, integration) of every
world.

metric: body-steps/s of the batched persistent path (all worlds in one launch
per step, HIP-graph replay).  Also reported: the plain one-world-per-launch
path on the same worlds, the LCP complementarity residual, device counters,
the HBM roofline of the step kernel and the reference CPU build on the host.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--worlds N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "ode-0.12_amd"))

HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(nb, ng, pairs, contacts, rows):
    """Compulsory HBM bytes of one step (FP32): per body 180 B (state in and
    out, mass data, external force), per geom 40 B, per candidate pair 16 B
    plus 128 B of narrowphase reads, per contact 36 B, per row 4 B."""
    return 180 * nb + 40 * ng + 16 * pairs + 128 * pairs + 36 * contacts + 4 * rows


def cpu_baseline(worlds, steps):
    """Reference ODE 0.12 single-precision build (oracle/_ref) on one host core,
    on the first worlds of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    if not po.have_ref("single"):
        return None
    lib = po.ref_lib("single")
    t0 = time.perf_counter()
    bodies = 0
    for sc in worlds:
        po.run(lib, sc, steps, 0)
        bodies += sc.nb
    dt = time.perf_counter() - t0
    return bodies * steps / dt, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--cpu-sample", type=int, default=256, help="worlds in the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--launch", default=None,
                    help="fixed launch config 'wg,cell_scale,lds,split,grid' instead of autotuning")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import odemi
    from odemi import scenes

    torch.cuda.set_device(0)
    t0 = time.perf_counter()
    worlds = scenes.uneven_worlds(args.worlds, seed=7)
    nb = sum(w.nb for w in worlds)
    ng = sum(w.ng for w in worlds)
    print(f"# workload: {len(worlds)} worlds, {nb} bodies, {ng} geoms (built in {time.perf_counter() - t0:.1f}s)",
          file=sys.stderr, flush=True)

    b = odemi.Batch(worlds, odemi.ORDER_COLORED)
    if args.launch:
        wg, cs, lds, split, grid = args.launch.split(",")
        b.set_launch_config(int(wg), float(cs), int(lds), int(split), int(grid))
    elif args.autotune:
        b.autotune(2)
    cfg = b.launch_config()
    print(f"# launch config {cfg}", file=sys.stderr, flush=True)
    if args.warmup:
        b.step(args.warmup)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    dev_ms = b.step_timed(args.steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    cnt = b.counters()
    # residual of one more step
    b.track_residual(True)
    b.step(1)
    res = b.residual()
    b.track_residual(False)
    # per-phase workgroup time (in-kernel clocks) over a few more steps
    b.profile(True)
    b.step(5)
    ph = {k: v / 5 for k, v in b.phase_times().items()}
    b.profile(False)
    ms_step = wall / args.steps * 1e3
    value = nb * args.steps / wall
    kernel_ms = dev_ms / args.steps
    B = algorithmic_bytes(nb, ng, cnt["pairs"], cnt["contacts"], cnt["rows"])
    achieved = B / (kernel_ms * 1e-3) / 1e9
    print(f"# batched: {ms_step:.3f} ms/step wall, {kernel_ms:.3f} ms/step device, counters {cnt}", file=sys.stderr,
          flush=True)

    # plain path: one world per launch
    plain = None
    if args.plain_steps > 0:
        bp = odemi.Batch(worlds, odemi.ORDER_COLORED)
        bp.set_launch_config(cfg["wg"], cfg["cell_scale"], cfg["lds"], cfg["split"])
        bp.step_plain(1)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        bp.step_plain(args.plain_steps)
        torch.cuda.synchronize()
        pw = time.perf_counter() - t2
        plain = nb * args.plain_steps / pw
        print(f"# plain: {pw / args.plain_steps * 1e3:.1f} ms/step", file=sys.stderr, flush=True)
        bp.close()

    cpu = None
    if not args.no_cpu:
        sample = worlds[: args.cpu_sample]
        csteps = args.warmup + args.steps
        r = cpu_baseline(sample, csteps)
        if r:
            cpu = dict(value=r[0], unit="body-steps/s", cores=1, kind="reference",
                       sample=f"reference ODE 0.12 dSINGLE build, dSpaceCollide+dWorldQuickStep+dJointGroupEmpty, "
                              f"first {len(sample)} worlds ({sum(w.nb for w in sample)} bodies) x steps 0-{csteps}, "
                              f"{r[1]:.1f}s on 1 host core")

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "traffic_uneven_worlds.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("bytes_per_launch")
        except Exception:
            traffic = None

    line = {
        "metric": "body-steps/sec batched vs plain GPU quickstep at LCP residual <=1e-4",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes)",
        "config": {"workload": f"batched uneven worlds: {len(worlds)} worlds, body count clamp(Poisson(64),8,256) "
                               f"box/sphere piles + 0-16 ball/hinge chains, h=0.01, 20 iterations, seed 7",
                   "worlds": len(worlds), "bodies": nb, "geoms": ng, "solver_order": "colored",
                   "launch": cfg},
        "plain_value": plain,
        "batched_over_plain": (value / plain) if plain else None,
        "lcp_residual": {"max": float(np.max(res)), "median": float(np.median(res))},
        "counters_per_step": cnt,
        "phase_wg_ms_per_step": ph,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": B, "kernel_ms": kernel_ms},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
