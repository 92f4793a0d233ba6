"""Test harness of the step adjoint.

* ``contact_scene``: small seeded contact-only worlds (spheres/boxes on a
  plane, optional restitution) of the differentiable-pile kind.
* ``jointed_scene``: a ball/hinge chain on a plane with a hinge and a ball
  joint to the static world.
* ``vehicle_scene``: one car on hinge2 wheels with per-step controls.
* ``converged_oracle``: the FP64 restatement solving to a tight residual.
* ``com_loss``: sum over worlds of the squared centre of mass of the final
  positions, and its gradient with respect to the final state.
* ``host_adjoint``: the host double build of the adjoint
  (tests/_build/libadj_host.so, built from tests/adjoint/adj_host.cpp and
  ode-0.12_amd/csrc/adj_step.hpp) on a recorded run of the CPU restatement.
* ``fd_grad``: FP64 central differences through the CPU restatement with a
  converged solver.
* ``Rollout``: a torch.autograd.Function whose forward steps a batch on the
  GPU with the tape on and whose backward calls dBatchStepBackward.
"""
from __future__ import annotations

import copy
import ctypes as C
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "ode-0.12_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import pyoracle as po  # noqa: E402
from odemi import scenes  # noqa: E402

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def contact_scene(seed: int, n: int, kinds=(0, 1), bounce: float = 0.0, spacing: float = 0.35,
                  iterations: int = 20000, dt: float = 0.01, mu: float = 0.6) -> scenes.Scene:
    """n bodies (kinds cycle: 0 sphere, 1 box) stacked loosely over a plane,
    with seeded horizontal and angular velocities."""
    rng = np.random.default_rng(seed)
    b = scenes._Builder()
    b.geom(scenes.PLANE, (0.0, 0.0, 1.0, 0.0))
    for i in range(n):
        kind = kinds[i % len(kinds)]
        pos = (rng.uniform(-0.4, 0.4), rng.uniform(-0.4, 0.4), 0.2 + spacing * i)
        v = (rng.uniform(-1, 1), rng.uniform(-1, 1), 0.0)
        w = tuple(rng.uniform(-1, 1, 3))
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        if kind == 0:
            b.sphere_body(pos, rng.uniform(0.15, 0.25), rng.uniform(0.5, 2), lvel=v, avel=w)
        else:
            b.box_body(pos, tuple(rng.uniform(0.25, 0.45, 3)), rng.uniform(0.5, 2), lvel=v, avel=w, q=tuple(q))
    mode = scenes.APPROX1 | (scenes.BOUNCE if bounce > 0 else 0)
    return b.scene(dt=dt, mu=mu, seed=seed, iterations=iterations, surf_mode=mode, bounce=bounce, bounce_vel=0.1)


def jointed_scene(seed: int, n_links: int = 4, iterations: int = 20000, dt: float = 0.01,
                  mu: float = 0.6) -> scenes.Scene:
    """A chain of n_links boxes joined alternately by ball joints and hinges
    (scenes._chain) falling onto a plane, a box hanging from a hinge to the
    static world and a sphere on a ball joint to the static world, all with
    seeded linear and angular velocities."""
    rng = np.random.default_rng(seed)
    b = scenes._Builder()
    b.geom(scenes.PLANE, (0.0, 0.0, 1.0, 0.0))
    scenes._chain(b, rng, n_links, (-0.6, 0.0, 0.25))
    p = b.box_body((0.9, 0.0, 0.75), (0.1, 0.1, 0.3), 0.8)
    b.joint(scenes.HINGE, p, -1, (0.9, 0.0, 0.9), axis1=(0.0, 1.0, 0.0))
    s = b.sphere_body((0.9, 0.6, 0.5), 0.12, 0.6)
    b.joint(scenes.BALL, s, -1, (0.9, 0.6, 0.75))
    sc = b.scene(dt=dt, mu=mu, seed=seed, iterations=iterations, surf_mode=scenes.APPROX1)
    sc.body_lvel = rng.uniform(-0.8, 0.8, (sc.nb, 3))
    sc.body_avel = rng.uniform(-1.0, 1.0, (sc.nb, 3))
    return sc


def vehicle_scene(seed: int, nsteps: int) -> scenes.Scene:
    """One car of the vehicle workload (scenes.vehicle_world: chassis on four
    hinge2 wheels with suspension, per-step steering stops on the front
    wheels and wheel torques) without ramps, with seeded small initial
    velocities; it lands on the plane within the first 40 steps."""
    rng = np.random.default_rng(seed)
    sc = scenes.vehicle_world(rng, nsteps, n_ramps=0, seed=seed)
    sc.body_lvel = rng.uniform(-0.5, 0.5, (sc.nb, 3))
    sc.body_avel = rng.uniform(-0.5, 0.5, (sc.nb, 3))
    sc.iterations = 20000
    return sc


def com_loss(scs, finals):
    """L = sum_w |COM_w|^2 and dL/d(final state) per world (nb x 13)."""
    L = 0.0
    grads = []
    for sc, f in zip(scs, finals):
        m = np.asarray(sc.body_mass, np.float64)
        com = (m[:, None] * f[:, :3]).sum(0) / m.sum()
        L += float((com ** 2).sum())
        g = np.zeros((sc.nb, 13))
        g[:, :3] = 2 * com[None, :] * (m / m.sum())[:, None]
        grads.append(g)
    return L, grads


class converged_oracle:
    """Within the block the FP64 restatement solves every step's LCP in the
    coloured order to the residual `tol` (conjugate-gradient steps, cap
    400000 sweeps), so that finite differences and the recorded impulses are
    those of the converged LCP the adjoint differentiates."""

    def __init__(self, tol: float = 1e-12):
        self.tol = tol

    def __enter__(self):
        lib = po.oracle_lib("f64")
        lib.orc_set_colored(1)
        lib.orc_set_colored_solver(C.c_double(self.tol), 400000, 10, 1)
        return self

    def __exit__(self, *exc):
        lib = po.oracle_lib("f64")
        lib.orc_set_colored(0)
        lib.orc_set_colored_solver(C.c_double(0.0), 0, 10, 0)
        return False


def forward_f64(sc, nsteps):
    return po.run(po.oracle_lib("f64"), sc, nsteps, nsteps)["traj"][-1]


def fd_grad(sc, nsteps, field, idx=None, eps=1e-6):
    """Central difference of com_loss of one world through the FP64
    restatement; field is body_lvel / body_avel (idx = (body, axis)), mu or
    bounce."""
    vals = []
    for s in (+1, -1):
        s2 = copy.deepcopy(sc)
        if field in ("mu", "bounce"):
            setattr(s2, field, getattr(sc, field) + s * eps)
        else:
            a = getattr(s2, field).copy()
            a[idx] += s * eps
            setattr(s2, field, a)
        vals.append(com_loss([s2], [forward_f64(s2, nsteps)])[0])
    return (vals[0] - vals[1]) / (2 * eps)


def _geom_rot(sc):
    out = np.zeros((sc.ng, 9))
    for g in range(sc.ng):
        w, x, y, z = sc.geom_q[g]
        out[g] = [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y), 2 * (x * y + w * z),
                  1 - 2 * (x * x + z * z), 2 * (y * z - w * x), 2 * (x * z - w * y), 2 * (y * z + w * x),
                  1 - 2 * (x * x + y * y)]
    return out


def host_adjoint(sc, nsteps, lib_path=None):
    """(loss, dL/d(initial state) nb x 13, dL/dmu, dL/dbounce, stats) of one
    world from the host double adjoint on an FP64 recorded run."""
    lib = C.CDLL(lib_path or os.path.join(ROOT, "tests", "_build", "libadj_host.so"))
    t = po.tape_run(po.oracle_lib("f64"), sc, nsteps)
    L, (gout,) = com_loss([sc], [t["states"][-1]])
    params = np.array([sc.dt, sc.erp, sc.cfm, sc.min_depth, 1e30 if math.isinf(sc.max_vel) else sc.max_vel,
                       *sc.gravity, sc.mu, sc.bounce, sc.bounce_vel,
                       1.0 if sc.surf_mode & scenes.BOUNCE else 0.0, sc.max_contacts, 1e-9, 1e-13, 4000],
                      np.float64)
    arr = lambda a, t=np.float64: np.ascontiguousarray(a, t)
    keep = [arr(sc.body_mass), arr(sc.body_I), arr(sc.geom_type, np.int32), arr(sc.geom_body, np.int32),
            arr(sc.geom_param), arr(sc.geom_pos), arr(_geom_rot(sc)), params, arr(t["states"]),
            arr(t["nk"], np.int32), arr(t["contacts"]), arr(t["joints"]), arr(gout)]
    ptr = lambda a: a.ctypes.data_as(_ip if a.dtype == np.int32 else _dp)
    gin = np.zeros((sc.nb, 13))
    gmu, gb = C.c_double(), C.c_double()
    stats = np.zeros(2, np.int32)
    ok = lib.adj_host_backward(sc.nb, sc.ng, *[ptr(a) for a in keep[:8]], nsteps, ptr(keep[8]), ptr(keep[9]),
                               ptr(keep[10]), sc.nj, ptr(keep[11]), ptr(keep[12]), ptr(gin), C.byref(gmu),
                               C.byref(gb), ptr(stats))
    if not ok:
        raise ValueError("a recorded joint has rows the adjoint does not cover")
    return L, gin, gmu.value, gb.value, stats


class Rollout:
    """torch.autograd.Function over a GPU batch of contact-only worlds:
    forward(lvel0, avel0, mu, bounce) -> final positions (nb x 3) after
    `nsteps` steps in the coloured order iterated to `tol`; backward through
    dBatchStepBackward.  lvel0/avel0: (nb, 3) float64 tensors; mu/bounce: one
    value per world."""

    @staticmethod
    def make(torch, odemi, scs, nsteps, tol=1e-6, max_iters=4000, interval=10):
        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, lvel0, avel0, mu, bounce):
                worlds = []
                off = 0
                for w, sc in enumerate(scs):
                    s2 = copy.deepcopy(sc)
                    s2.body_lvel = lvel0[off:off + sc.nb].detach().cpu().numpy().astype(np.float64)
                    s2.body_avel = avel0[off:off + sc.nb].detach().cpu().numpy().astype(np.float64)
                    s2.mu = float(mu[w])
                    s2.bounce = float(bounce[w])
                    worlds.append(s2)
                    off += sc.nb
                b = odemi.Batch(worlds, odemi.ORDER_COLORED)
                b.set_tolerance(tol, max_iters, 10)
                b.tape_options(interval)
                b.tape_begin(nsteps)
                b.step(nsteps)
                st = b.state()
                ctx.batch = b
                ctx.nsteps = nsteps
                return torch.as_tensor(st[:, :3].astype(np.float64))

            @staticmethod
            def backward(ctx, gpos):
                b = ctx.batch
                gout = np.zeros((b.nb, 13), np.float32)
                gout[:, :3] = gpos.detach().cpu().numpy()
                gi, gmu, gb = b.step_backward(ctx.nsteps, gout)
                ctx.stats = b.tape_stats()
                b.close()
                t = lambda a: torch.as_tensor(np.asarray(a, np.float64))
                return t(gi[:, 7:10]), t(gi[:, 10:13]), t(gmu), t(gb)

        return Fn
