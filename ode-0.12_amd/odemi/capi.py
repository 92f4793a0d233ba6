"""ctypes declarations of include/odemi.h (batch and adjoint entry points)."""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libodemi.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_fp = C.POINTER(C.c_float)


class BatchWorldDesc(C.Structure):
    _fields_ = [
        ("gravity", C.c_double * 3), ("erp", C.c_double), ("cfm", C.c_double), ("iterations", C.c_int),
        ("sor_w", C.c_double), ("max_vel", C.c_double), ("min_depth", C.c_double), ("dt", C.c_double),
        ("seed", C.c_ulong), ("hash_minlevel", C.c_int), ("hash_maxlevel", C.c_int),
        ("surf_mode", C.c_int), ("mu", C.c_double), ("mu2", C.c_double), ("bounce", C.c_double),
        ("bounce_vel", C.c_double), ("soft_erp", C.c_double), ("soft_cfm", C.c_double),
        ("max_contacts", C.c_int), ("skip_connected", C.c_int), ("nb", C.c_int), ("ng", C.c_int),
        ("nj", C.c_int),
    ]


class BatchDesc(C.Structure):
    _fields_ = [
        ("nworlds", C.c_int), ("worlds", C.POINTER(BatchWorldDesc)),
        ("body_pos", _dp), ("body_q", _dp), ("body_lvel", _dp), ("body_avel", _dp), ("body_mass", _dp),
        ("body_I", _dp), ("geom_type", _ip), ("geom_body", _ip), ("geom_param", _dp), ("geom_pos", _dp),
        ("geom_q", _dp), ("joint_type", _ip), ("joint_b1", _ip), ("joint_b2", _ip), ("joint_anchor", _dp),
        ("joint_axis1", _dp), ("joint_axis2", _dp), ("joint_param", _dp),
    ]


_lib = None


def load_library(path: str = LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it first (python -m odemi.build, or __graft_entry__.build())")
    L = C.CDLL(path, mode=os.RTLD_LOCAL)
    V = C.c_void_p
    sig = {
        "dBatchCreate": (V, [C.POINTER(BatchDesc), C.c_int]),
        "dBatchDestroy": (None, [V]),
        "dBatchStep": (C.c_int, [V, C.c_int]),
        "dBatchStepPlain": (C.c_int, [V, C.c_int]),
        "dBatchStepTimed": (C.c_int, [V, C.c_int, _fp]),
        "dBatchSetControls": (C.c_int, [V, _dp]),
        "dBatchGetState": (C.c_int, [V, _fp]),
        "dBatchSetState": (C.c_int, [V, _fp]),
        "dBatchGetCounters": (C.c_int, [V, C.POINTER(C.c_longlong)]),
        "dBatchSetResidualTracking": (C.c_int, [V, C.c_int]),
        "dBatchGetResidual": (C.c_int, [V, _fp]),
        "dBatchAutotune": (C.c_int, [V, C.c_int]),
        "dBatchGetLaunchConfig": (C.c_int, [V, _ip, _fp]),
        "dBatchSetLaunchConfig": (C.c_int, [V, C.c_int, C.c_float]),
        "dBatchGetLds": (C.c_int, [V]),
        "dBatchGetWorldCounters": (C.c_int, [V, _ip]),
        "dBatchResetSolveStats": (C.c_int, [V]),
        "dBatchSetSolverTolerance": (C.c_int, [V, C.c_float, C.c_int, C.c_int]),
        "dBatchSetSolverToleranceAccel": (C.c_int, [V, C.c_float, C.c_int, C.c_int, C.c_int]),
        "dBatchSetWarmStart": (C.c_int, [V, C.c_float]),
        "dBatchSetExactLCP": (C.c_int, [V, C.c_int]),
        "dBatchSetSorLaunch": (C.c_int, [V, C.c_int, C.c_int, C.c_int, C.c_int]),
        "dBatchGetSorLaunch": (C.c_int, [V, _ip, _ip, _ip, _ip]),
        "dBatchGetSplit": (C.c_int, [V]),
        "dBatchAutotuneEx": (C.c_int, [V, C.c_int, C.c_int]),
        "dBatchSetPersistentGrid": (C.c_int, [V, C.c_int]),
        "dBatchSetQueueOrder": (C.c_int, [V, C.c_int]),
        "dBatchGetQueueOrder": (C.c_int, [V]),
        "dBatchSetSplit": (C.c_int, [V, C.c_int]),
        "dBatchSetLds": (C.c_int, [V, C.c_int]),
        "dBatchSetProfiling": (C.c_int, [V, C.c_int]),
        "dBatchGetPhaseTimes": (C.c_int, [V, C.POINTER(C.c_double)]),
        "dBatchGetWorldPhaseTimes": (C.c_int, [V, C.POINTER(C.c_double)]),
        "dBatchGetPairs": (C.c_int, [V, C.c_int, C.c_int, _ip]),
        "dBatchGetContacts": (C.c_int, [V, C.c_int, C.c_int, _fp]),
        "dBatchDeviceState": (V, [V]),
        "dBatchStream": (V, [V]),
        "dBatchNumBodies": (C.c_int, [V]),
        "dBatchTapeBegin": (C.c_int, [V, C.c_int]),
        "dBatchTapeEnd": (C.c_int, [V]),
        "dBatchTapeSetOptions": (C.c_int, [V, C.c_int, C.c_float, C.c_float, C.c_int]),
        "dBatchTapeStats": (C.c_int, [V, C.POINTER(C.c_longlong)]),
        "dBatchTapeKeepSteps": (C.c_int, [V, C.c_int]),
        "dBatchStepBackward": (C.c_int, [V, C.c_int, _fp, _fp, _fp, _fp]),
        "odemiVersion": (C.c_char_p, []),
        "odemiDeviceCount": (C.c_int, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def lib():
    return load_library()
