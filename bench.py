"""Benchmark of the ODE 0.12 step path on one MI355X.

Workload: 4096 independent worlds of uneven size (body count per world
clamp(Poisson(64), 8, 256) boxes/spheres plus 0-16 ball/hinge chains of 4-12
links, spawned without interpenetration), h = 0.005, mu = 0.6, synthetic
seeded scenes, a 200-step rollout.  One step = broadphase + narrowphase +
quickstep (rows, SOR-LCP, integration) of every world.  The timed window is
the last --steps steps of the rollout, whatever --warmup says: every step
before it runs untimed in the same mode.

metric: body-steps/s of the batched persistent path (all worlds in one launch
per step, HIP-graph replay) with every world's SOR iterated until its LCP
residual is <= 1e-4 (warm-started from the previous step's impulses), over
every step of the rollout; the fraction of the window's solves that reached
the tolerance and the largest final residual are reported with it.  Also
reported: the plain one-world-per-launch path continuing the same batch and
state, the same two paths at the quickstep's fixed 20 iterations, device
counters, per-phase clocks, the HBM roofline of the step and the reference
CPU build on the host, timed on the same window.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--worlds N]
       python bench.py --workload adjoint   (the differentiable piles only)
       python bench.py --workload pile4096|vehicles   (the single 4096-body
       pile and the 8192 vehicle worlds at the quickstep's 20 iterations, in
       the reference and the coloured order, with the reference CPU build on a
       bounded sample)
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


def _cpu_worker(job):
    """One host process: the reference single-precision build over a slice of
    worlds; returns (body-steps, seconds) of steps skip..skip+steps (the
    driver times dSpaceCollide + dWorldQuickStep + dJointGroupEmpty of those
    steps only)."""
    scenes_, skip, steps = job
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    lib = po.ref_lib("single")
    secs = 0.0
    bodies = 0
    for sc in scenes_:
        r = po.run(lib, sc, skip + steps, 0, time_from=skip)
        secs += r["sec_collide"] + r["sec_step"]
        bodies += sc.nb
    return bodies * steps, secs


def cpu_baseline(worlds, skip, steps, procs):
    """Reference ODE 0.12 single-precision build (oracle/_ref) on `procs` host
    processes (ODE keeps its RNG and collider table in process globals, so
    worlds are split over processes, not threads), each stepping a slice of
    the sample; the rate of steps skip..skip+steps (the GPU's timed window).
    Runs before the GPU is touched."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import multiprocessing as mp
    import pyoracle as po
    if not po.have_ref("single"):
        return None
    jobs = [(worlds[k::procs], skip, steps) for k in range(procs)]
    t0 = time.perf_counter()
    # the workers are closed and joined (they exit on their own) rather than
    # terminated by the context manager
    pool = mp.get_context("spawn").Pool(procs)
    try:
        res = pool.map(_cpu_worker, jobs)
    finally:
        pool.close()
        pool.join()
    work = sum(r[0] for r in res)
    wall = max(r[1] for r in res)
    return work / wall, wall, time.perf_counter() - t0


def host_cores() -> int:
    """Host cores this process may use, at most 16 (one GPU's share of the
    host on the GPU machines)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def host_description() -> str:
    model = "unknown CPU"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    comp = ""
    bi = os.path.join(ROOT, "oracle", "_ref", "buildinfo.txt")
    if os.path.exists(bi):
        comp = open(bi).read().strip()
    return f"{model}; {os.cpu_count()} logical CPUs on the host; reference build: {comp or 'g++ -O2'}"


def _timed(b, steps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dev_ms = b.step_timed(steps)
    torch.cuda.synchronize()
    return time.perf_counter() - t0, dev_ms


def _plain_rate(odemi, worlds, cfg, state, steps, tol, max_iters, check_every):
    import torch
    bp = odemi.Batch(worlds, odemi.ORDER_COLORED)
    bp.set_launch_config(cfg["wg"], cfg["cell_scale"], cfg["lds"], cfg["split"])
    bp.set_state(state)
    if tol > 0:
        bp.set_tolerance(tol, max_iters, check_every, accel=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bp.step_plain(steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    bp.close()
    return dt / steps


def adjoint_run(odemi, scenes, np, steps, interval, keep=False):
    """Differentiable batched piles: 1024 worlds of U{32..128} boxes/spheres
    with restitution, h=0.01; forward with the tape (checkpoint every
    `interval` steps, SOR iterated to residual 1e-5) and dBatchStepBackward
    of loss = sum_w |COM_w|^2 with respect to the initial velocities, mu and
    the restitution coefficient."""
    import torch
    worlds = scenes.diff_piles(1024, seed=5)
    nb = sum(w.nb for w in worlds)
    b = odemi.Batch(worlds, odemi.ORDER_COLORED)
    b.set_tolerance(1e-5, 1000, 10, accel=True)
    b.tape_options(interval)
    b.tape_keep_steps(keep)
    b.tape_begin(steps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.step(steps)
    torch.cuda.synchronize()
    fwd = time.perf_counter() - t0
    st = b.state().astype(np.float64)
    gout = np.zeros((nb, 13), np.float32)
    off = 0
    for sc in worlds:
        m = sc.body_mass
        com = (m[:, None] * st[off:off + sc.nb, :3]).sum(0) / m.sum()
        gout[off:off + sc.nb, :3] = 2 * com[None, :] * (m / m.sum())[:, None]
        off += sc.nb
    t1 = time.perf_counter()
    gi, gmu, gb = b.step_backward(steps, gout)
    bwd = time.perf_counter() - t1
    stats = b.tape_stats()
    wc = b.world_counters()
    b.close()
    offs = np.concatenate([[0], np.cumsum([w.nb for w in worlds])])
    fin_w = np.array([np.isfinite(gi[offs[k]:offs[k + 1]]).all() and np.isfinite(gmu[k]) for k in range(len(worlds))])
    return {"workload": f"differentiable piles: {len(worlds)} worlds, U{{32..128}} boxes/spheres ({nb} bodies), "
                        f"bounce 0.2, h=0.01, {steps} steps, checkpoint every {interval}",
            "bodies": nb, "forward_ms_per_step": fwd / steps * 1e3, "backward_ms_per_step": bwd / steps * 1e3,
            "backward_body_steps_per_s": nb * steps / bwd, "adjoint_solve_iterations": stats["solve_iterations"],
            "collider_mismatches": stats["collider_mismatches"],
            "worlds_with_finite_gradients": int(fin_w.sum()), "frozen_worlds": int(((wc[:, 4] & 16) != 0).sum())}


def _oracle_traj(job):
    """FP64 (or FP32) CPU restatement of the reference order over one world:
    positions every `every` steps."""
    sc, nsteps, every, prec = job
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    return po.run(po.oracle_lib(prec), sc, nsteps, every)["traj"][:, :, :3]


def divergence_run(odemi, np, worlds, nsteps, tol, max_iters, procs, label):
    """Position divergence over `nsteps` steps on a sample of worlds, sampled
    every 10 steps (max over bodies and samples), between: the GPU reference
    order and the FP64 build of the reference's arithmetic (the CPU
    restatement, bit-exact to the reference double build); the reference's
    own FP32 build and FP64 build (its FP32 envelope); the GPU coloured order
    and the GPU reference order (both at the quickstep's iteration count);
    the headline path (coloured, iterated to the tolerance) and the FP64
    build."""
    import multiprocessing as mp
    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    every = 10
    has_ctrl = worlds[0].step_ctrl is not None
    ctrl = np.concatenate([w.step_ctrl[:nsteps] for w in worlds], axis=1) if has_ctrl else None

    def gpu(order, tolerance):
        b = odemi.Batch(worlds, order)
        if tolerance:
            b.set_tolerance(tol, max_iters, 10, accel=True)
            b.set_warm_start(1.0)
        out = [b.state()[:, :3].astype(np.float64)]
        for k in range(nsteps):
            if has_ctrl:
                b.set_controls(ctrl[k])
            b.step(1)
            if (k + 1) % every == 0:
                out.append(b.state()[:, :3].astype(np.float64))
            if (k + 1) % 50 == 0:
                log(f"divergence ({label}): GPU {'reference' if order == odemi.ORDER_REFERENCE else 'coloured'} "
                    f"order{' at the tolerance' if tolerance else ''}, step {k + 1} of {nsteps}")
        b.close()
        return np.stack(out)

    t0 = time.perf_counter()
    ref = gpu(odemi.ORDER_REFERENCE, False)
    col = gpu(odemi.ORDER_COLORED, False)
    head = gpu(odemi.ORDER_COLORED, True)
    log(f"divergence ({label}): GPU runs done in {time.perf_counter() - t0:.1f}s")

    # both restatements in one pool of host processes, in order
    jobs = [(w, nsteps, every, prec) for prec in ("f64", "f32") for w in worlds]
    out = []
    pool = mp.get_context("spawn").Pool(procs)
    try:
        log(f"divergence ({label}): {len(jobs)} restatement runs on {procs} host processes")
        for k, tr in enumerate(pool.imap(_oracle_traj, jobs)):
            out.append(tr)
            if (k + 1) % 8 == 0:
                log(f"divergence ({label}): {k + 1} of {len(jobs)} restatement runs")
    finally:
        pool.close()
        pool.join()
    f64 = np.concatenate(out[:len(worlds)], axis=1)
    f32 = np.concatenate(out[len(worlds):], axis=1)
    d = lambda a, b: float(np.abs(a - b).max())
    return {"sample": f"{label}: {len(worlds)} worlds ({sum(w.nb for w in worlds)} bodies), {nsteps} steps, "
                      f"positions every {every} steps",
            "reference_order_vs_fp64_m": d(ref, f64), "reference_fp32_vs_fp64_m": d(f32, f64),
            "coloured_vs_reference_order_m": d(col, ref), "headline_vs_fp64_m": d(head, f64)}


def config_run(args):
    """The single 4096-body pile (500-step config: seed 12345, h=0.005) or the
    8192 vehicle worlds (per-step steer stops and wheel torques applied before
    every step, as the reference frame loop does) at the quickstep's fixed 20
    iterations, in the reference order and the coloured order; one JSON line."""
    import numpy as np
    import odemi
    from odemi import scenes
    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    steps, warmup = args.steps, args.warmup
    if args.workload == "pile4096":
        worlds = [scenes.random_pile(4096, seed=12345)]
        desc = "single-world random pile: 4096 bodies (boxes/spheres), h=0.005, mu=0.6, 20 iterations, seed 12345"
        sample, procs = worlds, 1
    else:
        worlds = scenes.vehicle_worlds(8192, warmup + steps + 1, seed=11)
        desc = ("batched vehicle worlds: 8192 worlds, chassis + 4 sphere wheels on hinge2 joints, 32 static ramps, "
                "per-step steer/torque controls, h=0.01, mu=1.0, 20 iterations, seed 11")
        sample, procs = worlds[:2048], (args.cpu_procs or host_cores())
    nb = sum(w.nb for w in worlds)
    log(f"workload: {desc}; {len(worlds)} worlds, {nb} bodies")
    cpu = None
    if not args.no_cpu:
        r = cpu_baseline(sample, warmup, steps, procs)
        if r:
            cpu = dict(value=r[0], unit="body-steps/s", cores=procs, kind="reference",
                       sample=f"reference ODE 0.12 dSINGLE build, dSpaceCollide + dWorldQuickStep (20 "
                              f"iterations) + dJointGroupEmpty, {len(sample)} world(s) ({sum(w.nb for w in sample)} "
                              f"bodies) x steps {warmup}-{warmup + steps} (the GPU's timed window; only those "
                              f"steps timed) on {procs} host process(es), {r[1]:.1f}s timed, {r[2]:.1f}s in all",
                       host=host_description())
            log(f"cpu baseline: {r[0]:.3g} body-steps/s on {procs} process(es) ({r[1]:.1f}s)")
    import torch
    torch.cuda.set_device(0)
    has_ctrl = worlds[0].step_ctrl is not None
    ctrl = None
    if has_ctrl:
        ctrl = np.concatenate([w.step_ctrl[:warmup + steps] for w in worlds], axis=1)   # (steps, nj, 4)
    out = {}
    orders = [(odemi.ORDER_REFERENCE, "reference_order"), (odemi.ORDER_COLORED, "coloured_order")]
    if args.order != "both":
        orders = [o for o in orders if o[1].startswith(args.order)]
    for order, name in orders:
        b = odemi.Batch(worlds, order)
        if args.launch:
            wg, cs, lds, split, grid = args.launch.split(",")
            b.set_launch_config(int(wg), float(cs), int(lds), int(split), int(grid))
        else:
            b.autotune(2)
        cfg = b.launch_config()

        def run(k0, k1):
            if has_ctrl:
                for k in range(k0, k1):
                    b.set_controls(ctrl[k])
                    b.step(1)
            else:
                b.step(k1 - k0)
            torch.cuda.synchronize()

        run(0, warmup)
        t0 = time.perf_counter()
        run(warmup, warmup + steps)
        wall = time.perf_counter() - t0
        b.track_residual(True)
        b.step(1)
        res = b.residual()
        b.track_residual(False)
        b.profile(True)
        b.step(3)
        ph = {k: v / 3 for k, v in b.phase_times().items()}
        b.profile(False)
        cnt = b.counters()
        b.close()
        out[name] = {"value": nb * steps / wall, "ms_per_step": wall / steps * 1e3, "launch": cfg,
                     "lcp_residual_median": float(np.median(res)), "counters_per_step": cnt,
                     "phase_wg_ms_per_step": ph}
        log(f"{name}: {wall / steps * 1e3:.3f} ms/step, {nb * steps / wall:.3g} body-steps/s, launch {cfg}")
    head = out.get("coloured_order") or out["reference_order"]
    line = {"metric": "body-steps/s at the quickstep's 20 iterations", "value": head["value"],
            "unit": "body-steps/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded scenes)",
            "config": {"workload": desc, "worlds": len(worlds), "bodies": nb},
            "reference_order": out.get("reference_order"), "coloured_order": out.get("coloured_order"),
            "cpu_baseline": cpu}
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None,
                    help="timed steps: the last --steps steps of the rollout (uneven worlds, default 20); the "
                         "timed steps after --warmup untimed steps (pile4096 / vehicles, default 100)")
    ap.add_argument("--warmup", type=int, default=None,
                    help="uneven worlds: recorded only (every step before the timed window runs untimed); "
                         "pile4096 / vehicles: untimed steps before the timed ones (default 100)")
    ap.add_argument("--rollout", type=int, default=200, help="uneven worlds: steps of the rollout")
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--autotune", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-4, help="LCP residual the solver iterates down to")
    ap.add_argument("--max-iters", type=int, default=2000)
    ap.add_argument("--check-every", type=int, default=10)
    ap.add_argument("--accel", type=int, default=1,
                    help="nonsmooth nonlinear CG steps between the sweeps of the iterate-to-tolerance solver (0: off)")
    ap.add_argument("--queue-order", choices=("cost", "size"), default="size",
                    help="persistent queue: worlds by size (bodies + geoms) or by estimated solver work (cost)")
    ap.add_argument("--warm", type=float, default=1.0,
                    help="warm start of the iterate-to-tolerance solver (factor on the previous impulses; 0: off)")
    ap.add_argument("--adjoint-steps", type=int, default=100, help="steps of the adjoint run (0: skip)")
    ap.add_argument("--cpu-sample", type=int, default=4096, help="worlds in the CPU baseline sample")
    ap.add_argument("--cpu-procs", type=int, default=0,
                    help="host processes of the CPU baseline (0: the host cores available, at most 16)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--divergence-worlds", type=int, default=64,
                    help="uneven worlds: worlds of the trajectory-divergence sample (C3 and vehicles; 0: skip)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="run one device copy of this many bytes first (calibrates HBM byte counters)")
    ap.add_argument("--launch", default=None,
                    help="fixed launch config 'wg,cell_scale,lds,split,grid' instead of autotuning")
    ap.add_argument("--launch-tol", default=None, help="the same for the iterate-to-tolerance phase")
    ap.add_argument("--workload", default="uneven", choices=("uneven", "pile4096", "vehicles", "adjoint"),
                    help="uneven: the metric's batched uneven worlds; pile4096 / vehicles: the other configurations; "
                         "adjoint: the differentiable piles (forward with the tape, then dBatchStepBackward)")
    ap.add_argument("--order", default="both", choices=("both", "reference", "coloured"),
                    help="pile4096 / vehicles: the solver order(s) to run")
    args = ap.parse_args()
    if args.steps is None:
        args.steps = 20 if args.workload == "uneven" else 100
    if args.warmup is None:
        args.warmup = 5 if args.workload == "uneven" else 100
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.workload == "adjoint":
        import numpy as np
        import torch
        import odemi
        from odemi import scenes
        torch.cuda.set_device(0)
        adj = adjoint_run(odemi, scenes, np, args.adjoint_steps, 10)
        print("#", adj, file=sys.stderr, flush=True)
        kept = adjoint_run(odemi, scenes, np, args.adjoint_steps, 10, keep=True)
        print("# every step kept (no replay):", kept, file=sys.stderr, flush=True)
        adj["kept_steps"] = {k: kept[k] for k in ("forward_ms_per_step", "backward_ms_per_step",
                                                  "backward_body_steps_per_s", "worlds_with_finite_gradients",
                                                  "collider_mismatches")}
        steps = args.adjoint_steps
        nb = adj["bodies"]
        print(json.dumps({"metric": "body-steps/s of the step adjoint (dBatchStepBackward, replay + reverse)",
                          "value": adj["backward_body_steps_per_s"], "unit": "body-steps/s", "n_gpus": 1,
                          "steps": steps, "warmup": 0, "ms_per_step": adj["backward_ms_per_step"],
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
                          "data": "synthetic (seeded scenes)", "config": {"workload": adj["workload"], "bodies": nb},
                          "adjoint": adj}), flush=True)
        return
    if args.workload != "uneven":
        config_run(args)
        return

    import numpy as np
    import odemi
    from odemi import scenes

    ROLL = args.rollout
    K = args.steps
    if not 0 < K <= ROLL - 20:
        print(f"bench.py: --steps must be in 1..{ROLL - 20} (the timed window is the last --steps steps of the "
              f"{ROLL}-step rollout)", file=sys.stderr)
        sys.exit(2)
    t_from = ROLL - K
    t0 = time.perf_counter()
    worlds = scenes.uneven_worlds(args.worlds, seed=7)
    nb = sum(w.nb for w in worlds)
    ng = sum(w.ng for w in worlds)
    log = lambda *a: print("#", *a, file=sys.stderr, flush=True)
    log(f"workload: {len(worlds)} worlds, {nb} bodies, {ng} geoms (built in {time.perf_counter() - t0:.1f}s); "
        f"timed window: steps {t_from}-{ROLL} of the {ROLL}-step rollout")

    # reference CPU baseline first, before this process touches the GPU
    cpu = None
    if not args.no_cpu:
        procs = args.cpu_procs or host_cores()
        sample = worlds[: args.cpu_sample]
        r = cpu_baseline(sample, t_from, K, procs)
        if r:
            cpu = dict(value=r[0], unit="body-steps/s", cores=procs, kind="reference",
                       sample=f"reference ODE 0.12 dSINGLE build, dSpaceCollide + dWorldQuickStep (20 iterations) "
                              f"+ dJointGroupEmpty, {len(sample)} worlds ({sum(w.nb for w in sample)} bodies) x "
                              f"steps {t_from}-{ROLL} (the GPU's timed window; only those steps timed), worlds "
                              f"split over {procs} host processes, {r[1]:.1f}s timed, {r[2]:.1f}s in all",
                       host=host_description())
            log(f"cpu baseline: {r[0]:.3g} body-steps/s on {procs} processes ({r[1]:.1f}s timed)")

    import torch
    torch.cuda.set_device(0)
    if args.copy_bytes > 0:
        src = torch.ones(args.copy_bytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # (1) the quickstep's fixed iteration count (20, as dWorldQuickStep):
    # steps 0..t_from untimed, then the window
    b = odemi.Batch(worlds, odemi.ORDER_COLORED)
    b.set_queue_order(args.queue_order == "cost")
    if args.launch:
        wg, cs, lds, split, grid = args.launch.split(",")
        b.set_launch_config(int(wg), float(cs), int(lds), int(split), int(grid))
    elif args.autotune:
        b.autotune(2)
    cfg = b.launch_config()
    log(f"launch config {cfg}")
    b.step(t_from)
    s_from20 = b.state()
    wall20, dev20 = _timed(b, K)
    cnt20 = b.counters()
    b.track_residual(True)
    b.step(1)
    res20 = b.residual()
    b.track_residual(False)
    b.close()
    log(f"fixed 20 iterations: {wall20 / K * 1e3:.3f} ms/step, residual at step {ROLL + 1}: median "
        f"{np.median(res20):.3g} max {res20.max():.3g}, counters {cnt20}")

    # (2) the metric's operating point: the whole rollout at the tolerance,
    # warm-started; the launch shape autotuned on warm-started steps just
    # before the window (state restored), the last K steps timed
    b = odemi.Batch(worlds, odemi.ORDER_COLORED)
    b.set_launch_config(cfg["wg"], cfg["cell_scale"], cfg["lds"], cfg["split"], cfg["grid"])
    b.set_tolerance(args.tol, args.max_iters, args.check_every, accel=bool(args.accel))
    b.set_warm_start(args.warm)
    b.set_queue_order(args.queue_order == "cost")
    if args.launch_tol:
        wg, cs, lds, split, grid = args.launch_tol.split(",")
        b.set_launch_config(int(wg), float(cs), int(lds), int(split), int(grid))
    tw = time.perf_counter()
    b.step(t_from)
    log(f"steps 0-{t_from} at tolerance (untimed): {time.perf_counter() - tw:.1f}s")
    if not args.launch_tol and args.autotune:
        b.autotune(4, keep_cell=True, fused_only=True)
    cfg_tol = b.launch_config()
    log(f"launch config at tolerance {cfg_tol}")
    b.reset_solve_stats()
    wall, dev_ms = _timed(b, K)
    cnt = b.counters()
    wc = b.world_counters()
    unconv, resmax = b.solve_stats()
    dead = (wc[:, 4] & 16) != 0
    live = ~dead
    live_bodies = int(sum(w.nb for w, d in zip(worlds, dead) if not d))
    iters = wc[:, 5]
    solves = int(live.sum()) * K
    conv_solves = 1.0 - float(unconv[live].sum()) / max(1, solves)
    conv_worlds = float(np.mean(unconv[live] == 0)) if live.any() else 0.0
    log(f"tolerance {args.tol:g}: {wall / K * 1e3:.3f} ms/step, iterations at step {ROLL}: mean "
        f"{iters.mean():.1f} max {iters.max()}; over the window: {conv_solves:.4f} of the solves and "
        f"{conv_worlds:.4f} of the worlds converged, largest final residual {resmax[live].max():.3g}, "
        f"frozen worlds {int(dead.sum())}")
    # per-phase workgroup time (in-kernel clocks) over a few more steps
    b.profile(True)
    b.step(3)
    ph = {k: v / 3 for k, v in b.phase_times().items()}
    b.profile(False)
    # the plain path at the same tolerance, continuing the same state (and
    # warm-start records): one world per launch
    t_tol = None
    if args.plain_steps > 0:
        torch.cuda.synchronize()
        tp = time.perf_counter()
        b.step_plain(args.plain_steps)
        torch.cuda.synchronize()
        t_tol = (time.perf_counter() - tp) / args.plain_steps
    b.close()

    ms_step = wall / K * 1e3
    value = live_bodies * K / wall   # frozen (non-finite) worlds do no work and are not counted
    kernel_ms = dev_ms / K
    B = algorithmic_bytes(nb, ng, cnt["pairs"], cnt["contacts"], cnt["rows"])
    achieved = B / (kernel_ms * 1e-3) / 1e9

    plain = plain20 = None
    if args.plain_steps > 0:
        t_20 = _plain_rate(odemi, worlds, cfg, s_from20, args.plain_steps, 0.0, 0, 0)
        plain, plain20 = nb / t_tol, nb / t_20
        log(f"plain: {t_tol * 1e3:.1f} ms/step at tolerance, {t_20 * 1e3:.1f} ms/step at 20 iterations")

    adjoint = None
    if args.adjoint_steps > 0:
        adjoint = adjoint_run(odemi, scenes, np, args.adjoint_steps, 10)
        log(f"adjoint: {adjoint}")

    divergence = None
    if args.divergence_worlds > 0:
        nd = args.divergence_worlds
        procs = args.cpu_procs or host_cores()
        divergence = {
            "uneven_worlds": divergence_run(odemi, np, worlds[:nd], ROLL, args.tol, args.max_iters, procs,
                                            "batched uneven worlds (the first of the workload)"),
            "vehicles": divergence_run(odemi, np, scenes.vehicle_worlds(nd, ROLL + 1, seed=11), ROLL, args.tol,
                                       args.max_iters, procs, "vehicle worlds (seed 11)")}
        log(f"divergence: {divergence}")

    line = {
        "metric": "body-steps/sec batched vs plain GPU quickstep at LCP residual <=1e-4",
        "value": value,
        "unit": "body-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scenes)",
        "config": {"workload": f"batched uneven worlds: {len(worlds)} worlds, body count clamp(Poisson(64),8,256) "
                               f"box/sphere piles + 0-16 ball/hinge chains of 4-12 links (spawned without "
                               f"interpenetration), h=0.005, mu=0.6, {ROLL}-step rollout, timed steps "
                               f"{t_from}-{ROLL}, seed 7",
                   "worlds": len(worlds), "bodies": nb, "geoms": ng, "solver_order": "colored",
                   "solver": f"coloured Gauss-Seidel {'with nonsmooth nonlinear CG steps ' if args.accel else ''}"
                             f"until LCP residual "
                             f"<= {args.tol:g} (checked every {args.check_every}, cap {args.max_iters}), w=1.0, "
                             f"warm start {args.warm:g} x the previous step's impulses, every step of the rollout",
                   "launch": cfg_tol, "launch_fixed_iterations": cfg,
                   "queue_order": "by estimated solver work (rows x sweeps)" if args.queue_order == "cost"
                                  else "by size (bodies + geoms), largest first"},
        "plain_value": plain,
        "batched_over_plain": (value / plain) if plain else None,
        "lcp_residual": {"tol": args.tol, "converged_fraction": conv_solves,
                         "worlds_converged_every_step": conv_worlds,
                         "unconverged_solves": int(unconv[live].sum()), "solves": solves,
                         "frozen_worlds": int(dead.sum()),
                         "max_final_residual": float(resmax[live].max()) if live.any() else None,
                         "iterations_mean_last_step": float(iters.mean()),
                         "iterations_max_last_step": int(iters.max())},
        "contacts_per_step": cnt["contacts"],
        "fixed_20_iterations": {"value": nb * K / wall20, "ms_per_step": wall20 / K * 1e3,
                                "plain_value": plain20,
                                "batched_over_plain": (nb * K / wall20 / plain20) if plain20 else None,
                                "lcp_residual_median": float(np.median(res20)),
                                "lcp_residual_max": float(np.max(res20)), "counters_last_step": cnt20},
        "counters_last_step": cnt,
        "phase_wg_ms_per_step": ph,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": B, "kernel_ms": kernel_ms},
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "trajectory_divergence": divergence,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
