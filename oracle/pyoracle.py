"""ctypes harness for the scene_run() contract of oracle/scene.h.

TEST INFRASTRUCTURE.  Only tests/, the smoke check and the CPU baseline leg of
bench.py import this module.  It loads

* the CPU restatement          oracle/liboracle_{f64,f32}.so
* the reference ODE 0.12 build oracle/_ref/libode_ref_{double,single}.so
  (built from the reference sources by oracle/ref.mk when they are present)

and runs Scene objects (see odemi.scenes) through any library that exports
scene_run(), including the scene driver linked against the MI355X library.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class SceneDesc(C.Structure):
    _fields_ = [
        ("gravity", C.c_double * 3), ("erp", C.c_double), ("cfm", C.c_double), ("iterations", C.c_int),
        ("sor_w", C.c_double), ("max_vel", C.c_double), ("min_depth", C.c_double), ("dt", C.c_double),
        ("seed", C.c_ulong), ("hash_minlevel", C.c_int), ("hash_maxlevel", C.c_int),
        ("surf_mode", C.c_int), ("mu", C.c_double), ("mu2", C.c_double), ("bounce", C.c_double),
        ("bounce_vel", C.c_double), ("soft_erp", C.c_double), ("soft_cfm", C.c_double),
        ("max_contacts", C.c_int), ("skip_connected", C.c_int),
        ("nb", C.c_int), ("body_pos", _dp), ("body_q", _dp), ("body_lvel", _dp), ("body_avel", _dp),
        ("body_mass", _dp), ("body_I", _dp),
        ("ng", C.c_int), ("geom_type", _ip), ("geom_body", _ip), ("geom_param", _dp), ("geom_pos", _dp),
        ("geom_q", _dp),
        ("nj", C.c_int), ("joint_type", _ip), ("joint_b1", _ip), ("joint_b2", _ip), ("joint_anchor", _dp),
        ("joint_axis1", _dp), ("joint_axis2", _dp), ("joint_param", _dp),
        ("step_ctrl", _dp), ("geom_space", _ip), ("world_step", C.c_int),
    ]


class SceneOut(C.Structure):
    _fields_ = [
        ("nsteps", C.c_int), ("record_every", C.c_int), ("traj", _dp),
        ("diag_steps", C.c_int), ("max_pairs", C.c_int), ("pairs", _ip), ("pair_step_count", _ip),
        ("max_contacts", C.c_int), ("contacts", _dp), ("contact_step_count", _ip), ("feedback", _dp),
        ("total_pairs", C.c_longlong), ("total_contacts", C.c_longlong),
        ("sec_collide", C.c_double), ("sec_step", C.c_double), ("status", C.c_int), ("time_from", C.c_int),
        ("residual", _dp), ("grad_out", _dp), ("grad_in", _dp),
    ]


def _d(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, np.int32)
    return a, a.ctypes.data_as(_ip)


def make_desc(sc):
    """SceneDesc for a Scene; returns (desc, keepalive)."""
    keep = []
    d = SceneDesc()
    d.gravity[:] = list(sc.gravity)
    for k in ("erp", "cfm", "iterations", "sor_w", "max_vel", "min_depth", "dt", "seed", "hash_minlevel",
              "hash_maxlevel", "surf_mode", "mu", "mu2", "bounce", "bounce_vel", "soft_erp", "soft_cfm",
              "max_contacts", "skip_connected"):
        setattr(d, k, getattr(sc, k))
    d.nb, d.ng, d.nj = sc.nb, sc.ng, sc.nj
    for k in ("body_pos", "body_q", "body_lvel", "body_avel", "body_mass", "body_I", "geom_param", "geom_pos",
              "geom_q", "joint_anchor", "joint_axis1", "joint_axis2", "joint_param"):
        a, p = _d(getattr(sc, k))
        keep.append(a)
        setattr(d, k, p)
    for k in ("geom_type", "geom_body", "joint_type", "joint_b1", "joint_b2"):
        a, p = _i(getattr(sc, k))
        keep.append(a)
        setattr(d, k, p)
    if sc.step_ctrl is not None:
        a, p = _d(sc.step_ctrl)
        keep.append(a)
        d.step_ctrl = p
    d.world_step = int(getattr(sc, "world_step", 0))
    if getattr(sc, "geom_space", None) is not None:
        a, p = _i(sc.geom_space)
        keep.append(a)
        d.geom_space = p
    return d, keep


_libs: dict = {}


def load(path: str):
    path = os.path.abspath(path)
    if path not in _libs:
        if not os.path.exists(path):
            raise FileNotFoundError(path)
        lib = C.CDLL(path, mode=os.RTLD_LOCAL)
        lib.scene_run.argtypes = [C.POINTER(SceneDesc), C.POINTER(SceneOut)]
        lib.scene_run.restype = C.c_int
        lib.scene_dreal_size.restype = C.c_int
        _libs[path] = lib
    return _libs[path]


def oracle_lib(precision: str = "f64"):
    return load(os.path.join(HERE, f"liboracle_{precision}.so"))


def ref_path(precision: str = "double") -> str:
    return os.path.join(HERE, "_ref", f"libode_ref_{precision}.so")


def have_ref(precision: str = "double") -> bool:
    return os.path.exists(ref_path(precision))


def ref_lib(precision: str = "double"):
    return load(ref_path(precision))


def run(lib, sc, nsteps: int, record_every: int = 1, diag_steps: int = 0, max_pairs: int = 0,
        max_contacts: int = 0, feedback: bool = False, time_from: int = 0, residual: bool = False,
        grad_out=None) -> dict:
    """Run `nsteps` of a Scene through `lib.scene_run`; returns numpy arrays.
    sec_collide / sec_step time the steps from `time_from` on; residual=True
    (CPU restatement only) records every step's LCP residual; grad_out
    (nb x 13, MI355X library's scene driver only) differentiates the last
    step through dWorldQuickStepBackward into the result's grad_in."""
    desc, keep = make_desc(sc)
    out = SceneOut()
    out.nsteps = nsteps
    out.time_from = time_from
    out.record_every = record_every
    nrec = (nsteps // record_every + 1) if record_every > 0 else 1
    traj = np.zeros((nrec, sc.nb, 13))
    out.traj = traj.ctypes.data_as(_dp)
    out.diag_steps = diag_steps
    pairs = np.full((max(max_pairs, 1), 2), -1, np.int32)
    psc = np.zeros(max(diag_steps, 1), np.int32)
    out.max_pairs = max_pairs
    out.pairs = pairs.ctypes.data_as(_ip)
    out.pair_step_count = psc.ctypes.data_as(_ip)
    contacts = np.zeros((max(max_contacts, 1), 9))
    csc = np.zeros(max(diag_steps, 1), np.int32)
    fb = np.zeros((max(max_contacts, 1), 12))
    out.max_contacts = max_contacts
    out.contacts = contacts.ctypes.data_as(_dp)
    out.contact_step_count = csc.ctypes.data_as(_ip)
    out.feedback = fb.ctypes.data_as(_dp) if feedback else None
    res = np.zeros(max(nsteps, 1))
    out.residual = res.ctypes.data_as(_dp) if residual else None
    gin = np.zeros((sc.nb, 13))
    if grad_out is not None:
        gout = np.ascontiguousarray(grad_out, np.float64).reshape(sc.nb, 13)
        out.grad_out = gout.ctypes.data_as(_dp)
        out.grad_in = gin.ctypes.data_as(_dp)
    status = lib.scene_run(C.byref(desc), C.byref(out))
    del keep
    npairs = int(psc[:diag_steps].sum()) if diag_steps else 0
    ncont = int(csc[:diag_steps].sum()) if diag_steps else 0
    return dict(status=status, traj=traj, pairs=pairs[:min(npairs, max_pairs)],
                pair_step_count=psc[:diag_steps].copy(), contacts=contacts[:min(ncont, max_contacts)],
                contact_step_count=csc[:diag_steps].copy(), feedback=fb[:min(ncont, max_contacts)],
                total_pairs=out.total_pairs, total_contacts=out.total_contacts,
                residual=res[:nsteps].copy() if residual else None,
                grad_in=gin if grad_out is not None else None,
                sec_collide=out.sec_collide, sec_step=out.sec_step)


def tape_run(lib, sc, nsteps: int, max_total: int = 200000) -> dict:
    """Step-by-step record of a run (orc_tape_run): states before every step
    (and the final one), every step's contacts with their row impulses, and
    every permanent joint's record after each step (64 values: impulses,
    row bounds and cfm, frames, limit/motor rows, torques; orc_tape_run)."""
    fn = lib.orc_tape_run
    fn.restype = C.c_long
    fn.argtypes = [C.POINTER(SceneDesc), C.c_int, _dp, _ip, _dp, C.c_long, _dp]
    desc, keep = make_desc(sc)
    states = np.zeros((nsteps + 1, sc.nb, 13))
    nk = np.zeros(nsteps, np.int32)
    con = np.zeros((max_total, 16))
    jrec = np.zeros((nsteps, max(sc.nj, 1), 64))
    total = fn(C.byref(desc), nsteps, states.ctypes.data_as(_dp), nk.ctypes.data_as(_ip),
               con.ctypes.data_as(_dp), max_total, jrec.ctypes.data_as(_dp))
    del keep
    if total < 0:
        raise RuntimeError("tape_run: contact buffer too small")
    return dict(states=states, nk=nk, contacts=con[:total].copy(), joints=jrec[:, :sc.nj].copy())
