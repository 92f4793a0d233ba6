"""The C-ABI library: it loads, and exports every entry point include/*.h
declares (no GPU needed)."""
import os
import re
import subprocess

import pytest


def _declared(root):
    names = set()
    for f in os.listdir(os.path.join(root, "include")):
        if f.endswith(".h"):
            names |= set(re.findall(r"ODE_API\s+[\w\s\*]+?\b(\w+)\s*\(", open(os.path.join(root, "include", f)).read()))
    return names


def test_library_exports_every_declared_symbol(root):
    lib = os.path.join(root, "ode-0.12_amd", "odemi", "libodemi.so")
    assert os.path.exists(lib), "build first"
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    syms = {l.split()[-1] for l in out.splitlines() if l.strip()}
    decl = _declared(root)
    assert len(decl) > 100
    assert sorted(decl - syms) == []


def test_library_loads_and_reports(root):
    import odemi
    L = odemi.lib()
    assert b"gfx950" in L.odemiVersion()


def test_struct_layouts_match_the_reference_abi(root):
    # dContactGeom 64 B, dContact 128 B, dSurfaceParameters 48 B in the
    # single-precision build (include/ode/contact.h:52-97)
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "odemi.h"
int main(void){ printf("%zu %zu %zu %zu %zu %zu\n", sizeof(dContactGeom), sizeof(dContact), sizeof(dSurfaceParameters),
  offsetof(dContactGeom, depth), offsetof(dContactGeom, g1), sizeof(dMass)); return 0; }'''
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        c = os.path.join(td, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(td, "t")
        subprocess.run(["gcc", "-I", os.path.join(root, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["64", "128", "48", "32", "40", "68"]


def test_scene_driver_links_against_the_library(root):
    p = os.path.join(root, "tests", "_build", "libscene_odemi.so")
    assert os.path.exists(p)
    out = subprocess.run(["nm", "-D", "--undefined-only", p], capture_output=True, text=True, check=True).stdout
    assert "dSpaceCollide" in out and "dWorldQuickStep" in out


def _bad_scenes():
    """Descriptors the ODE setters would refuse by assertion, one fault each."""
    import numpy as np
    from odemi import scenes
    out = []
    base = lambda: scenes.box_stack(3)
    sc = base(); sc.geom_body = sc.geom_body.copy(); sc.geom_body[-1] = sc.nb
    out.append(("geom body out of range", sc, "out of range"))
    sc = base(); sc.geom_body = sc.geom_body.copy(); sc.geom_body[0] = 0   # geom 0 is the plane
    out.append(("plane on a body", sc, "planes take none"))
    sc = base(); sc.geom_type = sc.geom_type.copy(); sc.geom_type[-1] = 7
    out.append(("unknown geom type", sc, "unknown type"))
    sc = base(); sc.geom_param = sc.geom_param.copy(); sc.geom_param[-1, 1] = -1.0
    out.append(("negative box side", sc, "size must be"))
    sc = base(); sc.body_mass = sc.body_mass.copy(); sc.body_mass[1] = 0.0
    out.append(("zero mass", sc, "mass must be"))
    sc = base(); sc.dt = 0.0
    out.append(("zero step size", sc, "step size"))
    ch = scenes.pile_world(np.random.default_rng(1), 8, 1, seed=3)
    assert ch.nj > 0
    sc = ch; sc.joint_b2 = sc.joint_b2.copy(); sc.joint_b1 = sc.joint_b1.copy(); sc.joint_b2[0] = sc.joint_b1[0]
    out.append(("joint on one body twice", sc, "out of range or equal"))
    sc = base(); sc.body_I = sc.body_I.copy(); sc.body_I[2] = np.diag([1.0, -1.0, 1.0]).reshape(9)
    out.append(("inertia not positive definite", sc, "positive definite"))
    sc = base(); sc.body_q = sc.body_q.copy(); sc.body_q[0] = 0.0
    out.append(("zero quaternion", sc, "zero quaternion"))
    sc = base(); sc.body_pos = sc.body_pos.copy(); sc.body_pos[1, 2] = float("nan")
    out.append(("NaN position", sc, "not finite"))
    return out


@pytest.mark.parametrize("case", range(10))
def test_batch_descriptor_faults_are_refused(root, capfd, case):
    """dBatchCreate refuses a descriptor the reference's setters would assert
    on, before any device work (no GPU needed), and says why."""
    import odemi
    name, sc, msg = _bad_scenes()[case]
    with pytest.raises(RuntimeError):
        odemi.Batch([sc], odemi.ORDER_REFERENCE)
    err = capfd.readouterr().err
    assert "dBatchCreate" in err and msg in err, (name, err)


def test_api_assertions_follow_the_reference(root, capfd):
    """The reference's argument assertions on the API path go to the debug
    handler and leave the object unchanged (no GPU work): dGeomSetBody on a
    plane ("geom must be placeable", collision_kernel.cpp:510) and
    dJointAttach of one body to itself ("can't have body1==body2",
    ode.cpp:1352)."""
    import ctypes as C
    import odemi
    L = odemi.lib()
    V = C.c_void_p
    for name, res, args in [("dWorldCreate", V, []), ("dWorldDestroy", None, [V]), ("dBodyCreate", V, [V]),
                            ("dCreatePlane", V, [V, C.c_float, C.c_float, C.c_float, C.c_float]),
                            ("dGeomSetBody", None, [V, V]), ("dGeomGetBody", V, [V]), ("dGeomDestroy", None, [V]),
                            ("dJointCreateBall", V, [V, V]), ("dJointAttach", None, [V, V, V]),
                            ("dJointGetBody", V, [V, C.c_int])]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    w = L.dWorldCreate()
    b = L.dBodyCreate(w)
    p = L.dCreatePlane(None, 0.0, 0.0, 1.0, 0.0)
    L.dGeomSetBody(p, b)
    assert not L.dGeomGetBody(p)
    j = L.dJointCreateBall(w, None)
    L.dJointAttach(j, b, b)
    assert not L.dJointGetBody(j, 0) and not L.dJointGetBody(j, 1)
    err = capfd.readouterr().err
    assert "geom must be placeable" in err and "body1==body2" in err, err
    L.dGeomDestroy(p)
    L.dWorldDestroy(w)


def test_api_teardown_under_address_sanitizer(root, tmp_path):
    """The API object model under AddressSanitizer (host-side build of the
    library, no device work): dBodyDestroy leaves the body's geoms at its last
    pose (ode.cpp:296-309), and dWorldDestroy before dJointGroupDestroy and
    dSpaceDestroy (demos/raycar/ode.cpp:26-29) touches no freed memory
    (tests/capi/teardown.c)."""
    import glob
    asan_dir = os.path.join(root, "tests", "_build", "asan")
    lib = os.path.join(asan_dir, "libodemi_asan.so")
    assert os.path.exists(lib), "build first (__graft_entry__.build())"
    rt = glob.glob("/opt/rocm*/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so")
    assert rt, "clang AddressSanitizer runtime not found"
    exe = str(tmp_path / "teardown")
    subprocess.run(["/opt/rocm/llvm/bin/clang", "-g", "-fsanitize=address", "-shared-libsan",
                    "-I", os.path.join(root, "include"), os.path.join(root, "tests", "capi", "teardown.c"),
                    "-o", exe, lib, f"-Wl,-rpath,{asan_dir}", f"-Wl,-rpath,{os.path.dirname(rt[0])}"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr[-3000:]
