"""dBatchStepBackward on the GPU (FP32, through the C-ABI): gradients of the
squared centre of mass after N steps with respect to initial velocities, mu
and restitution, against FP64 central differences through the CPU
restatement (converged solver) and against the host double adjoint, for
contact-only worlds and for worlds with ball joints and hinges; the
checkpoint interval must not change the result; the torch.autograd.Function
wrapper returns the same gradients."""
import copy
import ctypes as C
import os

import numpy as np
import pytest

from adjoint.harness import (Rollout, com_loss, contact_scene, converged_oracle, fd_grad, forward_f64, host_adjoint,
                             jointed_scene, vehicle_scene)
from conftest import say
from odemi import scenes

pytestmark = pytest.mark.gpu

N = 40
TOL = 1e-3   # relative to the largest velocity gradient of the world


def _scenes():
    return [contact_scene(4, 4, (0,), bounce=0.5), contact_scene(5, 4, (1,)), contact_scene(6, 6, (0, 1))]


def _gpu_grad(odemi_mod, scs, interval=10, warm=0.0, keep=False):
    b = odemi_mod.Batch(scs, odemi_mod.ORDER_COLORED)
    b.set_tolerance(1e-6, 4000, 10)
    b.set_warm_start(warm)
    b.tape_options(interval)
    b.tape_keep_steps(keep)
    b.tape_begin(N)
    b.step(N)
    st = b.state().astype(np.float64)
    finals, off = [], 0
    for sc in scs:
        finals.append(st[off:off + sc.nb])
        off += sc.nb
    L, gl = com_loss(scs, finals)
    gi, gmu, gb = b.step_backward(N, np.concatenate(gl))
    stats = b.tape_stats()
    b.close()
    return L, gi.astype(np.float64), gmu, gb, stats


def test_gpu_adjoint_matches_central_differences_and_host(odemi_mod):
    scs = _scenes()
    say("adjoint: GPU backward over", N, "steps,", len(scs), "worlds")
    L, gi, gmu, gb, stats = _gpu_grad(odemi_mod, scs)
    say("adjoint stats", stats)
    assert stats["collider_mismatches"] == 0
    off = 0
    for w, sc in enumerate(scs):
        g = gi[off:off + sc.nb]
        _, hg, hmu, hb, _ = host_adjoint(sc, N)
        gmax = float(np.abs(hg[:, 7:13]).max())
        err_host = float(np.abs(g[:, 7:13] - hg[:, 7:13]).max()) / gmax
        worst = 0.0
        for b in range(sc.nb):
            for field, col in (("body_lvel", 7), ("body_avel", 10)):
                k = (b + col) % 3
                fd = fd_grad(sc, N, field, (b, k))
                worst = max(worst, abs(fd - g[b, col + k]) / gmax)
        fd_mu = fd_grad(sc, N, "mu")
        say(f"world {w}: max rel err vs FD {worst:.2e}, vs host adjoint {err_host:.2e}, "
            f"dL/dmu gpu {gmu[w]:.6g} fd {fd_mu:.6g}")
        assert worst <= TOL and err_host <= TOL
        assert abs(gmu[w] - fd_mu) <= TOL * max(gmax, abs(fd_mu))
        if sc.bounce > 0:
            fd_b = fd_grad(sc, N, "bounce")
            assert abs(gb[w] - fd_b) <= TOL * max(gmax, abs(fd_b))
        off += sc.nb


@pytest.mark.parametrize("warm", [0.0, 1.0], ids=["cold", "warm"])
def test_checkpoint_interval_does_not_change_gradient(odemi_mod, warm):
    """Segments are replayed from checkpoints; a checkpoint holds the
    warm-start records too, so replays of warm-started steps reproduce the
    forward and the gradient does not depend on the interval."""
    scs = _scenes()
    say("adjoint: checkpoint interval 3 vs 40, warm start", warm)
    a = _gpu_grad(odemi_mod, scs, interval=3, warm=warm)
    b = _gpu_grad(odemi_mod, scs, interval=40, warm=warm)
    assert a[4]["collider_mismatches"] == 0 and b[4]["collider_mismatches"] == 0
    scale = float(np.abs(a[1]).max())
    assert np.abs(a[1] - b[1]).max() <= 1e-5 * scale
    assert np.allclose(a[2], b[2], rtol=1e-5, atol=1e-6 * scale)


def test_kept_steps_give_the_replayed_gradient(odemi_mod):
    """With every step kept by the forward pass the backward pass runs no
    replay; the records are those a replay writes, so the gradient is the
    checkpointed one."""
    scs = _scenes()
    say("adjoint: kept steps vs checkpoint interval 10")
    a = _gpu_grad(odemi_mod, scs, interval=10)
    b = _gpu_grad(odemi_mod, scs, interval=10, keep=True)
    assert b[4]["collider_mismatches"] == 0
    scale = float(np.abs(a[1]).max())
    assert np.abs(a[1] - b[1]).max() <= 1e-5 * scale
    assert np.allclose(a[2], b[2], rtol=1e-5, atol=1e-6 * scale)
    say("kept vs replayed: max |d grad| / max |grad|", float(np.abs(a[1] - b[1]).max()) / scale)


def test_torch_autograd_function(odemi_mod):
    import torch
    scs = _scenes()
    say("adjoint: torch.autograd.Function")
    Fn = Rollout.make(torch, odemi_mod, scs, N)
    lv = torch.tensor(np.concatenate([s.body_lvel for s in scs]), dtype=torch.float64, requires_grad=True)
    av = torch.tensor(np.concatenate([s.body_avel for s in scs]), dtype=torch.float64, requires_grad=True)
    mu = torch.tensor([s.mu for s in scs], dtype=torch.float64, requires_grad=True)
    bo = torch.tensor([s.bounce for s in scs], dtype=torch.float64, requires_grad=True)
    pos = Fn.apply(lv, av, mu, bo)
    loss = 0
    off = 0
    for sc in scs:
        m = torch.tensor(sc.body_mass, dtype=torch.float64)
        com = (m[:, None] * pos[off:off + sc.nb]).sum(0) / m.sum()
        loss = loss + (com ** 2).sum()
        off += sc.nb
    loss.backward()
    _, gi, gmu, gb, _ = _gpu_grad(odemi_mod, scs)
    assert np.allclose(lv.grad.numpy(), gi[:, 7:10], rtol=1e-4, atol=1e-6)
    assert np.allclose(av.grad.numpy(), gi[:, 10:13], rtol=1e-4, atol=1e-6)
    assert np.allclose(mu.grad.numpy(), gmu, rtol=1e-4, atol=1e-6)


def _perturbed(sc, field, idx, delta):
    s2 = copy.deepcopy(sc)
    if field in ("mu", "bounce"):
        setattr(s2, field, getattr(sc, field) + delta)
    else:
        a = getattr(s2, field).copy()
        a[idx] += delta
        setattr(s2, field, a)
    return s2


def _contact_structure(sc, n):
    import pyoracle as po
    r = po.run(po.oracle_lib("f64"), sc, n, n, diag_steps=n, max_pairs=10 ** 6, max_contacts=10 ** 6)
    return r["contact_step_count"], r["pairs"]


def _c5_gpu(odemi_mod, scs, n):
    b = odemi_mod.Batch(scs, odemi_mod.ORDER_COLORED)
    b.set_tolerance(1e-6, 4000, 10, accel=True)
    b.tape_options(10)
    b.tape_begin(n)
    b.step(n)
    st = b.state().astype(np.float64)
    finals, off = [], 0
    for sc in scs:
        finals.append(st[off:off + sc.nb])
        off += sc.nb
    _, gl = com_loss(scs, finals)
    gi, gmu, gb = b.step_backward(n, np.concatenate(gl))
    stats = b.tape_stats()
    b.close()
    assert stats["collider_mismatches"] == 0
    assert np.isfinite(gi).all()
    return finals, gi.astype(np.float64), gmu, gb


@pytest.mark.parametrize("n", [30, 100])
def test_differentiable_pile_config(odemi_mod, n):
    """The differentiable-pile configuration: worlds of its U{32..128}
    box/sphere distribution with restitution, h = 0.01, the tape checkpointed
    every 10 steps, loss = sum of squared centres of mass; gradients with
    respect to sampled initial velocities, mu and restitution.

    * The host double adjoint (the same adjoint source in double, on the
      FP64 run of the restatement) against FP64 central differences (h =
      1e-6) through the restatement: within 1e-4 of the largest gradient, at
      30 and at the configuration's 100 steps.
    * The GPU adjoint (FP32, forward iterated to 1e-6) against the host
      adjoint and the central differences: within 1e-3 at 30 steps.  By 100
      steps the FP32 forward has left the FP64 trajectory (a pile is
      chaotic: the final positions of these worlds differ by up to 0.17 m),
      so the GPU gradient there is that of another trajectory: it must be
      finite, and its distance from the FP64 gradient is reported next to
      the trajectory divergence, not bounded.
    A component whose central difference crosses a change of the contact
    structure (the contacts or pairs of some step differ between +h and -h)
    is ill-posed for finite differences: it is reported and not compared."""
    ws = sorted(scenes.diff_piles(64, seed=5), key=lambda s: s.nb)[:2]
    scs = []
    for sc in ws:
        sc = copy.deepcopy(sc)
        sc.iterations = 20000   # the FP64 restatement converges too
        scs.append(sc)
    say("differentiable piles:", [s.nb for s in scs], "bodies,", n, "steps, checkpoint 10")
    finals, gi, gmu, gb = _c5_gpu(odemi_mod, scs, n)
    strict = n <= 30
    off = 0
    for w, sc in enumerate(scs):
        g = gi[off:off + sc.nb]
        _, hg, hmu, hb, _ = host_adjoint(sc, n)
        f64 = forward_f64(sc, n)
        div = float(np.abs(finals[w][:, :3] - f64[:, :3]).max())
        gmax = float(np.abs(hg[:, 7:13]).max())
        err_host = float(np.abs(g[:, 7:13] - hg[:, 7:13]).max()) / gmax
        say(f"world {w} ({sc.nb} bodies): FP32 vs FP64 final positions {div:.2e} m; GPU vs host adjoint "
            f"{err_host:.2e} of max |grad| {gmax:.3g}")
        assert not strict or err_host <= TOL
        rng = np.random.default_rng(w)
        comps = [("body_lvel", (int(rng.integers(sc.nb)), int(rng.integers(3)))),
                 ("body_avel", (int(rng.integers(sc.nb)), int(rng.integers(3)))), ("mu", None), ("bounce", None)]
        compared = 0
        for field, idx in comps:
            a, c = (_contact_structure(_perturbed(sc, field, idx, s * 1e-6), n) for s in (1, -1))
            if not (np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])):
                say(f"  {field} {idx}: ill-posed for central differences (contacts switch within +-1e-6)")
                continue
            fd = fd_grad(sc, n, field, idx)
            if field == "mu":
                got, host = float(gmu[w]), hmu
            elif field == "bounce":
                got, host = float(gb[w]), hb
            else:
                col = 7 if field == "body_lvel" else 10
                got, host = float(g[idx[0], col + idx[1]]), float(hg[idx[0], col + idx[1]])
            scale = max(gmax, abs(fd))
            say(f"  {field} {idx}: fd {fd:.6g} host {host:.6g} gpu {got:.6g}")
            assert abs(host - fd) <= 1e-4 * scale
            assert not strict or abs(got - fd) <= TOL * scale
            compared += 1
        assert compared >= 1
        off += sc.nb


def test_world_quickstep_backward_matches_central_differences(odemi_mod):
    """dWorldQuickStepBackward through the ODE API (the scene driver, an
    unchanged ODE frame loop, records its last dWorldQuickStep with
    dWorldSetTape and differentiates it): dL/d(state before the last step)
    for L = squared centre of mass after it, against FP64 central
    differences (h = 1e-6) of that one step through the restatement from the
    same pre-step state, for every position and velocity component, on two
    contact scenes, a ball/hinge chain with joints to the static world and a
    car on hinge2 wheels with its steering stops; the reference-order solve
    runs 20000 iterations, so it converges."""
    import pyoracle as po
    from conftest import ROOT
    lib = po.load(os.path.join(ROOT, "tests", "_build", "libscene_odemi.so"))
    orc = po.oracle_lib("f64")
    n = 8
    car = vehicle_scene(5, n)
    car.step_ctrl = None   # steering locked by its stops at 0; no wheel torques (body torques are not taped)
    for sc in (contact_scene(4, 4, (0,), bounce=0.5), contact_scene(5, 4, (1,)), jointed_scene(3, 4), car):
        say("dWorldQuickStepBackward:", sc.nb, "bodies, step", n)
        pre = po.run(lib, sc, n - 1, n - 1)["traj"][-1]
        fin = po.run(lib, sc, n, n)["traj"][-1]
        _, (gout,) = com_loss([sc], [fin])
        r = po.run(lib, sc, n, n, grad_out=gout)
        assert r["status"] == 0, "dWorldQuickStepBackward failed"
        gin = r["grad_in"]
        assert np.isfinite(gin).all()

        def loss_from(state):
            orc.orc_set_state_override(state.ctypes.data_as(C.POINTER(C.c_double)), sc.nb)
            try:
                end = po.run(orc, sc, 1, 1)["traj"][-1]
            finally:
                orc.orc_set_state_override(None, 0)
            return com_loss([sc], [end])[0]

        orc.orc_set_state_override.argtypes = [C.POINTER(C.c_double), C.c_int]
        base = np.ascontiguousarray(pre, np.float64)
        fd = np.zeros_like(gin)
        for b in range(sc.nb):
            for col in list(range(0, 3)) + list(range(7, 13)):
                vals = []
                for s in (1, -1):
                    st = base.copy()
                    st[b, col] += s * 1e-6
                    vals.append(loss_from(st))
                fd[b, col] = (vals[0] - vals[1]) / 2e-6
        cols = list(range(0, 3)) + list(range(7, 13))
        scale = float(np.abs(fd[:, cols]).max())
        err = float(np.abs(gin[:, cols] - fd[:, cols]).max()) / scale
        say(f"  max |grad| {scale:.4g}, max rel err vs FD {err:.2e}")
        assert err <= TOL


def test_gpu_adjoint_with_joints(odemi_mod):
    """Worlds with a ball/hinge chain on a plane and a hinge and a ball joint
    to the static world: the GPU gradients (FP32, forward iterated to 1e-6)
    against the host double adjoint and FP64 central differences of the
    restatement iterated to 1e-12, every velocity component of every body
    and mu, within 1e-3 of the largest gradient; the kept-steps backward
    pass gives the checkpointed gradient.  The seeds are ones whose FP32
    trajectory stays within 1e-6 m of the FP64 one over the 25 steps (the
    FP32 restatement, checked on the host): in others a contact event is
    decided by the last bits (seed 8: 1.5e-4 m apart after step 22), the
    FP32 run then differentiates another branch, and the comparison with
    FP64 is ill-posed."""
    n = 25
    scs = [jointed_scene(3, 4), jointed_scene(9, 5), jointed_scene(12, 5)]
    say("adjoint with joints:", [(s.nb, s.nj) for s in scs], "bodies/joints,", n, "steps")
    b = odemi_mod.Batch(scs, odemi_mod.ORDER_COLORED)
    b.set_tolerance(1e-6, 20000, 10, accel=True)
    b.tape_options(10)
    b.tape_begin(n)
    b.step(n)
    st = b.state().astype(np.float64)
    finals, off = [], 0
    for sc in scs:
        finals.append(st[off:off + sc.nb])
        off += sc.nb
    _, gl = com_loss(scs, finals)
    gi, gmu, gb = b.step_backward(n, np.concatenate(gl))
    stats = b.tape_stats()
    b.close()
    say("adjoint stats", stats)
    assert stats["collider_mismatches"] == 0
    gi = gi.astype(np.float64)
    assert np.isfinite(gi).all()
    with converged_oracle(1e-12):
        off = 0
        for w, sc in enumerate(scs):
            g = gi[off:off + sc.nb]
            div = float(np.abs(finals[w][:, :3] - forward_f64(sc, n)[:, :3]).max())
            say(f"world {w}: GPU vs FP64 final positions {div:.2e} m")
            assert div <= 1e-4
            _, hg, hmu, _, hst = host_adjoint(sc, n)
            assert hst[1] == 0
            gmax = float(np.abs(hg[:, 7:13]).max())
            err_host = float(np.abs(g[:, 7:13] - hg[:, 7:13]).max()) / gmax
            worst = 0.0
            for bb in range(sc.nb):
                for field, col in (("body_lvel", 7), ("body_avel", 10)):
                    for k in range(3):
                        fd = fd_grad(sc, n, field, (bb, k))
                        worst = max(worst, abs(fd - g[bb, col + k]) / gmax)
            fd_mu = fd_grad(sc, n, "mu")
            say(f"world {w}: max rel err vs FD {worst:.2e}, vs host adjoint {err_host:.2e}, "
                f"dL/dmu gpu {gmu[w]:.6g} host {hmu:.6g} fd {fd_mu:.6g}")
            assert worst <= TOL and err_host <= TOL
            assert abs(gmu[w] - fd_mu) <= TOL * max(gmax, abs(fd_mu))
            off += sc.nb
    # the kept-steps backward pass records the joints as the replay does
    b = odemi_mod.Batch(scs, odemi_mod.ORDER_COLORED)
    b.set_tolerance(1e-6, 20000, 10, accel=True)
    b.tape_options(10)
    b.tape_keep_steps(True)
    b.tape_begin(n)
    b.step(n)
    gk, gmuk, _ = b.step_backward(n, np.concatenate(gl))
    b.close()
    scale = float(np.abs(gi).max())
    assert np.abs(gk.astype(np.float64) - gi).max() <= 1e-5 * scale
    assert np.allclose(gmuk, gmu, rtol=1e-5, atol=1e-6 * scale)


def test_gpu_adjoint_on_uneven_worlds_with_chains(odemi_mod):
    """Worlds of the batched uneven-worlds workload that hold ball/hinge
    chains are accepted by the tape; 20 steps at its operating point
    (coloured order to 1e-4, warm start) are differentiated to finite
    gradients with no collider mismatch."""
    ws = scenes.uneven_worlds(48, seed=7)
    picked = [w for w in ws if w.nj > 0 and w.nb <= 96][:6]
    assert len(picked) >= 2
    say("uneven worlds with chains:", [(w.nb, w.nj) for w in picked])
    b = odemi_mod.Batch(picked, odemi_mod.ORDER_COLORED)
    b.set_tolerance(1e-4, 2000, 10, accel=True)
    b.set_warm_start(1.0)
    b.tape_options(10)
    b.tape_begin(20)
    b.step(20)
    st = b.state().astype(np.float64)
    finals, off = [], 0
    for sc in picked:
        finals.append(st[off:off + sc.nb])
        off += sc.nb
    _, gl = com_loss(picked, finals)
    gi, gmu, _ = b.step_backward(20, np.concatenate(gl))
    stats = b.tape_stats()
    b.close()
    say("stats", stats, "max |grad|", float(np.abs(gi).max()))
    assert stats["collider_mismatches"] == 0
    assert np.isfinite(gi).all() and np.isfinite(gmu).all()


def _vehicle_gpu(odemi_mod, scs, n, interval, keep=False, settled=False):
    # the FP32 solve of a landed car stays above a 1e-6 residual, so every
    # solve runs to the cap.  settled: 100 000 sweeps let the impulses
    # settle, as the adjoint of the converged LCP assumes; at 20 000 the
    # gradients depend on where the sweep order leaves them (up to 3e-2 off
    # FD), which is enough to compare replays with the kept steps
    b = odemi_mod.Batch(scs, odemi_mod.ORDER_COLORED)
    if settled:
        b.set_tolerance(3e-7, 100000, 10, accel=True)
    else:
        b.set_tolerance(1e-6, 20000, 10, accel=True)
    b.tape_options(interval)
    b.tape_keep_steps(keep)
    b.tape_begin(n)
    ctrl = np.concatenate([sc.step_ctrl[:n] for sc in scs], axis=1)
    for k in range(n):
        b.set_controls(ctrl[k])
        b.step(1)
    st = b.state().astype(np.float64)
    finals, off = [], 0
    for sc in scs:
        finals.append(st[off:off + sc.nb])
        off += sc.nb
    _, gl = com_loss(scs, finals)
    gi, gmu, _ = b.step_backward(n, np.concatenate(gl))
    stats = b.tape_stats()
    b.close()
    return finals, gi.astype(np.float64), gmu, stats


def test_gpu_adjoint_of_vehicles(odemi_mod):
    """Cars of the vehicle workload (hinge2 suspension, steering stops and
    wheel torques set per step by dBatchSetControls) over 50 steps in which
    they land: the GPU gradients against the host double adjoint and FP64
    central differences, every velocity component of every body, within
    1e-3 of the largest gradient (forward solves of 100 000 sweeps, steps
    kept); the replays re-apply each step's controls, so checkpoint
    intervals 7 and 50 and the kept-steps pass agree (20 000 sweeps)."""
    n = 50
    scs = [vehicle_scene(5, n), vehicle_scene(6, n)]
    say("adjoint of vehicles:", len(scs), "cars,", n, "steps")
    finals, gi, gmu, stats = _vehicle_gpu(odemi_mod, scs, n, 10, True, settled=True)
    say("adjoint stats", stats)
    assert stats["collider_mismatches"] == 0 and np.isfinite(gi).all()
    off = 0
    for w, sc in enumerate(scs):
        g = gi[off:off + sc.nb]
        div = float(np.abs(finals[w][:, :3] - forward_f64(sc, n)[:, :3]).max())
        _, hg, _, _, hst = host_adjoint(sc, n)
        assert hst[1] == 0
        gmax = float(np.abs(hg[:, 7:13]).max())
        err_host = float(np.abs(g[:, 7:13] - hg[:, 7:13]).max()) / gmax
        worst = 0.0
        for bb in range(sc.nb):
            for field, col in (("body_lvel", 7), ("body_avel", 10)):
                for k in range(3):
                    fd = fd_grad(sc, n, field, (bb, k))
                    worst = max(worst, abs(fd - g[bb, col + k]) / gmax)
        say(f"car {w}: GPU vs FP64 final positions {div:.2e} m, max rel err vs FD {worst:.2e}, "
            f"vs host adjoint {err_host:.2e}")
        assert div <= 1e-4 and worst <= TOL and err_host <= TOL
        off += sc.nb
    _, gk, _, _ = _vehicle_gpu(odemi_mod, scs, n, 10, True)
    scale = float(np.abs(gk).max())
    for interval in (7, 50):
        _, g2, _, _ = _vehicle_gpu(odemi_mod, scs, n, interval)
        say(f"interval {interval}: max |d grad| / max |grad| vs kept steps {float(np.abs(g2 - gk).max()) / scale:.2e}")
        assert np.abs(g2 - gk).max() <= 1e-5 * scale


def test_tape_refuses_joints_it_does_not_cover(odemi_mod):
    """Hinges with limits (the joint zoo) are refused by dBatchTapeBegin with
    an error, not differentiated."""
    say("tape refusal: joint zoo")
    b = odemi_mod.Batch([scenes.joint_zoo()], odemi_mod.ORDER_COLORED)
    try:
        with pytest.raises(RuntimeError):
            b.tape_begin(5)
    finally:
        b.close()
