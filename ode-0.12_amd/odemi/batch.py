"""Batch of independent worlds on the GPU (dBatch* entry points)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from .capi import BatchDesc, BatchWorldDesc, load_library

ORDER_REFERENCE = 0
ORDER_COLORED = 1


def _cat(scenes, name, width, dtype=np.float64):
    parts = [np.asarray(getattr(s, name), dtype).reshape(-1, width) if width else
             np.asarray(getattr(s, name), dtype).reshape(-1) for s in scenes]
    a = np.concatenate(parts) if parts else np.zeros((0, width) if width else (0,), dtype)
    return np.ascontiguousarray(a)


class Batch:
    """Device state of a list of `odemi.scenes.Scene` worlds.

    `step(n)` runs the batched path (one persistent launch per step from a
    HIP graph); `step_plain(n)` runs one world per launch.
    """

    def __init__(self, scenes, order: int = ORDER_COLORED):
        self.L = load_library()
        self.scenes = list(scenes)
        self.nworlds = len(self.scenes)
        self.nb = sum(s.nb for s in self.scenes)
        self.nj = sum(s.nj for s in self.scenes)
        self.body_offsets = np.cumsum([0] + [s.nb for s in self.scenes])
        self.joint_offsets = np.cumsum([0] + [s.nj for s in self.scenes])
        W = (BatchWorldDesc * self.nworlds)()
        for w, s in zip(W, self.scenes):
            w.gravity[:] = list(s.gravity)
            for k in ("erp", "cfm", "iterations", "sor_w", "max_vel", "min_depth", "dt", "seed", "hash_minlevel",
                      "hash_maxlevel", "surf_mode", "mu", "mu2", "bounce", "bounce_vel", "soft_erp", "soft_cfm",
                      "max_contacts", "skip_connected"):
                setattr(w, k, getattr(s, k))
            w.nb, w.ng, w.nj = s.nb, s.ng, s.nj
        keep = [W]
        D = BatchDesc()
        D.nworlds = self.nworlds
        D.worlds = C.cast(W, C.POINTER(BatchWorldDesc))
        for name, width in (("body_pos", 3), ("body_q", 4), ("body_lvel", 3), ("body_avel", 3), ("body_mass", 0),
                            ("body_I", 9), ("geom_param", 4), ("geom_pos", 3), ("geom_q", 4),
                            ("joint_anchor", 3), ("joint_axis1", 3), ("joint_axis2", 3), ("joint_param", 14)):
            a = _cat(self.scenes, name, width)
            keep.append(a)
            setattr(D, name, a.ctypes.data_as(C.POINTER(C.c_double)))
        for name in ("geom_type", "geom_body", "joint_type", "joint_b1", "joint_b2"):
            a = _cat(self.scenes, name, 0, np.int32)
            keep.append(a)
            setattr(D, name, a.ctypes.data_as(C.POINTER(C.c_int)))
        self.h = self.L.dBatchCreate(C.byref(D), order)
        del keep
        if not self.h:
            raise RuntimeError("dBatchCreate failed")

    def close(self):
        if getattr(self, "h", None):
            self.L.dBatchDestroy(self.h)
            self.h = None

    __del__ = close

    def step(self, n: int = 1):
        if not self.L.dBatchStep(self.h, n):
            raise RuntimeError("dBatchStep failed")

    def step_timed(self, n: int = 1) -> float:
        """n batched steps; returns their device time in ms (HIP events)."""
        ms = C.c_float()
        if not self.L.dBatchStepTimed(self.h, n, C.byref(ms)):
            raise RuntimeError("dBatchStepTimed failed")
        return ms.value

    def step_plain(self, n: int = 1):
        if not self.L.dBatchStepPlain(self.h, n):
            raise RuntimeError("dBatchStepPlain failed")

    def set_controls(self, ctrl):
        if ctrl is None:
            self.L.dBatchSetControls(self.h, None)
            return
        a = np.ascontiguousarray(ctrl, np.float64).reshape(-1)
        assert a.size == 4 * self.nj
        if not self.L.dBatchSetControls(self.h, a.ctypes.data_as(C.POINTER(C.c_double))):
            raise RuntimeError("dBatchSetControls failed")

    def state(self) -> np.ndarray:
        out = np.zeros((self.nb, 13), np.float32)
        if not self.L.dBatchGetState(self.h, out.ctypes.data_as(C.POINTER(C.c_float))):
            raise RuntimeError("dBatchGetState failed")
        return out

    def set_state(self, st):
        a = np.ascontiguousarray(st, np.float32).reshape(self.nb, 13)
        if not self.L.dBatchSetState(self.h, a.ctypes.data_as(C.POINTER(C.c_float))):
            raise RuntimeError("dBatchSetState failed")

    def counters(self) -> dict:
        c = (C.c_longlong * 5)()
        self.L.dBatchGetCounters(self.h, c)
        return dict(pairs=c[0], contacts=c[1], rows=c[2], groups=c[3], overflow=c[4])

    def world_counters(self) -> np.ndarray:
        """(nworlds, 8) int32: of the last step pairs, contacts, rows,
        islands/colours, overflow bits, SOR iterations; since the last
        reset_solve_stats() the tolerance-mode solves that ended at the cap
        above the tolerance and the largest final residual (float bits; see
        solve_stats())."""
        out = np.zeros((self.nworlds, 8), np.int32)
        if not self.L.dBatchGetWorldCounters(self.h, out.ctypes.data_as(C.POINTER(C.c_int))):
            raise RuntimeError("dBatchGetWorldCounters failed")
        return out

    def reset_solve_stats(self):
        if not self.L.dBatchResetSolveStats(self.h):
            raise RuntimeError("dBatchResetSolveStats failed")

    def solve_stats(self):
        """Per world since the last reset: (unconverged solves, largest final
        residual of an iterate-to-tolerance solve)."""
        wc = self.world_counters()
        return wc[:, 6].copy(), wc[:, 7].copy().view(np.float32)

    def set_tolerance(self, tol: float, max_iterations: int = 400, check_every: int = 10, accel: bool = False):
        """Coloured order: SOR runs until the LCP residual <= tol (0: fixed
        count); accel adds the conjugate-gradient step between sweeps."""
        if not self.L.dBatchSetSolverToleranceAccel(self.h, tol, max_iterations, check_every, 1 if accel else 0):
            raise ValueError("bad solver tolerance settings")

    def set_warm_start(self, warm: float = 1.0):
        """Coloured order: start each solve from warm x the previous step's
        impulses (contacts matched by geom pair and index within the pair);
        0 restores the zero start."""
        if not self.L.dBatchSetWarmStart(self.h, warm):
            raise ValueError("bad warm-start factor")

    def set_exact_lcp(self, max_island_rows: int):
        """dWorldStep for every world (reference order): each island's LCP
        solved exactly, with workspaces for islands of up to
        max_island_rows rows; 0 switches back to the SOR iterations."""
        if not self.L.dBatchSetExactLCP(self.h, int(max_island_rows)):
            raise ValueError("exact LCP needs the reference order and a workspace allocation")

    # ---- adjoint
    def tape_options(self, checkpoint_interval: int = 10, act_tol: float = 1e-6, lin_tol: float = 1e-6,
                     lin_maxit: int = 500):
        if not self.L.dBatchTapeSetOptions(self.h, checkpoint_interval, act_tol, lin_tol, lin_maxit):
            raise ValueError("bad tape options (or the tape is recording)")

    def tape_keep_steps(self, on: bool = True):
        """Keep every recorded step on the device so that step_backward
        needs no replay (set before tape_begin)."""
        if not self.L.dBatchTapeKeepSteps(self.h, 1 if on else 0):
            raise RuntimeError("dBatchTapeKeepSteps failed (the tape is recording)")

    def tape_begin(self, max_steps: int):
        """Record the following steps for step_backward (checkpointed)."""
        if not self.L.dBatchTapeBegin(self.h, max_steps):
            raise RuntimeError("dBatchTapeBegin failed (see the error handler output)")

    def tape_end(self):
        self.L.dBatchTapeEnd(self.h)

    def step_backward(self, nsteps: int, grad_out: np.ndarray):
        """dL/d(state nsteps ago) and dL/d(mu), dL/d(bounce) per world, from
        dL/d(current state) (nb x 13: pos, q, lvel, avel)."""
        go = np.ascontiguousarray(grad_out, np.float32).reshape(self.nb, 13)
        gi = np.zeros((self.nb, 13), np.float32)
        gmu = np.zeros(self.nworlds, np.float32)
        gb = np.zeros(self.nworlds, np.float32)
        fp = C.POINTER(C.c_float)
        if not self.L.dBatchStepBackward(self.h, nsteps, go.ctypes.data_as(fp), gi.ctypes.data_as(fp),
                                         gmu.ctypes.data_as(fp), gb.ctypes.data_as(fp)):
            raise RuntimeError("dBatchStepBackward failed (see the error handler output)")
        return gi, gmu, gb

    def tape_stats(self) -> dict:
        out = (C.c_longlong * 2)()
        self.L.dBatchTapeStats(self.h, out)
        return dict(solve_iterations=out[0], collider_mismatches=out[1])

    def track_residual(self, on: bool = True):
        self.L.dBatchSetResidualTracking(self.h, 1 if on else 0)

    def residual(self) -> np.ndarray:
        out = np.zeros(self.nworlds, np.float32)
        self.L.dBatchGetResidual(self.h, out.ctypes.data_as(C.POINTER(C.c_float)))
        return out

    def autotune(self, trial_steps: int = 3, keep_cell: bool = False, fused_only: bool = False) -> int:
        """Times launch shapes on the live batch (state restored) in its
        current solver mode and keeps the fastest; returns the workgroup size."""
        flags = (1 if keep_cell else 0) | (2 if fused_only else 0)
        return self.L.dBatchAutotuneEx(self.h, trial_steps, flags)

    def launch_config(self):
        wg = C.c_int()
        sc = C.c_float()
        grid = self.L.dBatchGetLaunchConfig(self.h, C.byref(wg), C.byref(sc))
        on, swg, slds, sgrid = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        self.L.dBatchGetSorLaunch(self.h, C.byref(on), C.byref(swg), C.byref(slds), C.byref(sgrid))
        return dict(wg=wg.value, cell_scale=sc.value, grid=grid, lds=self.L.dBatchGetLds(self.h),
                    split=self.L.dBatchGetSplit(self.h),
                    sor=dict(on=on.value, wg=swg.value, lds=slds.value, grid=sgrid.value))

    def set_sor_launch(self, on: bool, wg: int = 128, lds: int = 0, grid: int = 0):
        """Coloured order at a tolerance: the sweeps in a kernel of their own
        with this workgroup size, LDS pool and grid (0: one workgroup per world)."""
        if not self.L.dBatchSetSorLaunch(self.h, 1 if on else 0, wg, lds, grid):
            raise RuntimeError("bad sweep-kernel launch config")

    def set_launch_config(self, wg: int, cell_scale: float = 1.0, lds: int | None = None,
                          split: int | None = None, grid: int | None = None):
        if not self.L.dBatchSetLaunchConfig(self.h, wg, cell_scale):
            raise RuntimeError("bad launch config")
        if lds is not None and not self.L.dBatchSetLds(self.h, lds):
            raise RuntimeError(f"bad LDS pool size {lds}")
        if split is not None and not self.L.dBatchSetSplit(self.h, int(split)):
            raise RuntimeError("the two-kernel step (split=1) is switched off")
        if grid is not None and not self.L.dBatchSetPersistentGrid(self.h, int(grid)):
            raise RuntimeError(f"bad grid {grid}")

    def set_queue_order(self, by_cost: bool):
        """Persistent launches pull worlds by estimated solver work (True) or
        by size (False, the default)."""
        if not self.L.dBatchSetQueueOrder(self.h, 1 if by_cost else 0):
            raise RuntimeError("dBatchSetQueueOrder failed")

    def queue_order(self) -> bool:
        return self.L.dBatchGetQueueOrder(self.h) == 1

    PHASES = ("aabb", "broadphase", "narrowphase", "bodies_info1", "info2_rhs", "order", "sor", "integrate")

    def profile(self, on: bool = True):
        """Turns the in-kernel phase clocks on or off (resets them)."""
        if not self.L.dBatchSetProfiling(self.h, 1 if on else 0):
            raise RuntimeError("dBatchSetProfiling failed")

    def phase_times(self) -> dict:
        """Workgroup-ms per phase summed over worlds since the last read."""
        out = (C.c_double * 8)()
        if not self.L.dBatchGetPhaseTimes(self.h, out):
            raise RuntimeError("dBatchGetPhaseTimes failed")
        return dict(zip(self.PHASES, list(out)))

    def world_phase_times(self) -> np.ndarray:
        """Workgroup-ms per phase and world since the last read,
        (nworlds, 8) in PHASES order."""
        out = np.zeros((self.nworlds, 8), np.float64)
        if not self.L.dBatchGetWorldPhaseTimes(self.h, out.ctypes.data_as(C.POINTER(C.c_double))):
            raise RuntimeError("dBatchGetWorldPhaseTimes failed")
        return out

    def pairs(self, w: int, maxn: int = 1 << 20) -> np.ndarray:
        out = np.zeros((maxn, 2), np.int32)
        n = self.L.dBatchGetPairs(self.h, w, maxn, out.ctypes.data_as(C.POINTER(C.c_int)))
        if n < 0:
            raise IndexError(w)
        return out[:min(n, maxn)].copy()

    def contacts(self, w: int, maxn: int = 1 << 18) -> np.ndarray:
        """Last step's contacts of world w: pos 3, normal 3, depth, g1, g2, lambda 3."""
        out = np.zeros((maxn, 12), np.float32)
        n = self.L.dBatchGetContacts(self.h, w, maxn, out.ctypes.data_as(C.POINTER(C.c_float)))
        if n < 0:
            raise IndexError(w)
        return out[:min(n, maxn)].copy()

    def world_state(self, st, w):
        return st[self.body_offsets[w]:self.body_offsets[w + 1]]
