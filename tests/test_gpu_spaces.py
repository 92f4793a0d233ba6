"""dSpaceCollide2 and spaces inside spaces through the C ABI, against the
reference's own single-precision build driven by the same API calls:
the near-callback sequence (pairs, with spaces among them, in order) and the
contact counts dCollide returns for each geom pair must be identical.
Covers space x space (either side smaller), space x geom, geom x space,
geom x geom, a space nested in another (the callback descends with
dSpaceCollide2) and sublevels."""
import ctypes as C
import os

import numpy as np
import pytest

import pyoracle as po
from conftest import ROOT, golden, say

pytestmark = pytest.mark.gpu

V, F, I = C.c_void_p, C.c_float, C.c_int
CB = C.CFUNCTYPE(None, V, V, V)


def _declare(L):
    for name, res, args in [
            ("dInitODE2", I, [C.c_uint]), ("dWorldCreate", V, []), ("dWorldDestroy", None, [V]),
            ("dBodyCreate", V, [V]), ("dBodySetPosition", None, [V, F, F, F]),
            ("dHashSpaceCreate", V, [V]), ("dSimpleSpaceCreate", V, [V]), ("dSpaceDestroy", None, [V]),
            ("dSpaceSetSublevel", None, [V, I]), ("dGeomIsSpace", I, [V]),
            ("dCreateBox", V, [V, F, F, F]), ("dCreateSphere", V, [V, F]), ("dCreatePlane", V, [V, F, F, F, F]),
            ("dGeomSetBody", None, [V, V]), ("dGeomSetPosition", None, [V, F, F, F]),
            ("dGeomSetData", None, [V, V]), ("dGeomGetData", V, [V]),
            ("dCollide", I, [V, V, I, V, I]), ("dSpaceCollide2", None, [V, V, V, CB])]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _run(L):
    """The same sequence of API calls on library L; returns the callback log."""
    L.dInitODE2(0)
    w = L.dWorldCreate()
    dyn = L.dHashSpaceCreate(None)
    stat = L.dHashSpaceCreate(None)
    sub = L.dSimpleSpaceCreate(stat)        # a space inside the static space
    names = {}

    def tag(g, name):
        names[g] = name
        L.dGeomSetData(g, C.c_void_p(len(names)))
        return g

    names[dyn], names[stat], names[sub] = "dyn", "stat", "sub"
    rng = np.random.default_rng(9)
    bodies = []
    for i in range(5):
        b = L.dBodyCreate(w)
        L.dBodySetPosition(b, float(rng.uniform(-2, 2)), float(rng.uniform(-2, 2)), float(rng.uniform(0.2, 1.2)))
        g = L.dCreateBox(dyn, 0.8, 0.6, 0.7) if i % 2 == 0 else L.dCreateSphere(dyn, 0.45)
        L.dGeomSetBody(g, b)
        tag(g, f"d{i}")
        bodies.append(b)
    tag(L.dCreatePlane(stat, 0.0, 0.0, 1.0, 0.0), "plane")
    for i in range(4):
        g = L.dCreateBox(stat, 1.5, 1.5, 0.5)
        L.dGeomSetPosition(g, float(rng.uniform(-2, 2)), float(rng.uniform(-2, 2)), 0.1)
        tag(g, f"s{i}")
    for i in range(3):
        g = L.dCreateBox(sub, 1.0, 1.0, 1.0)
        L.dGeomSetPosition(g, float(rng.uniform(-2, 2)), float(rng.uniform(-2, 2)), 0.6)
        tag(g, f"n{i}")
    log = []
    buf = (C.c_byte * (128 * 8))()

    def near(data, o1, o2):
        log.append((names[o1], names[o2]))
        if L.dGeomIsSpace(o1) or L.dGeomIsSpace(o2):
            L.dSpaceCollide2(o1, o2, None, cb)
            return
        log.append(("contacts", L.dCollide(o1, o2, 8, buf, 128)))

    cb = CB(near)
    singles = [g for g, n in names.items() if n.startswith("d")]
    for label, a, b in [("dyn x stat", dyn, stat), ("stat x dyn", stat, dyn), ("geom x stat", singles[0], stat),
                        ("sub x geom", sub, singles[1]), ("geom x geom", singles[2], singles[3])]:
        log.append(("call", label))
        L.dSpaceCollide2(a, b, None, cb)
    L.dSpaceSetSublevel(sub, 1)
    log.append(("call", "sub (sublevel 1) x dyn"))
    L.dSpaceCollide2(sub, dyn, None, cb)
    L.dSpaceDestroy(dyn)
    L.dSpaceDestroy(stat)
    L.dWorldDestroy(w)
    return log


def test_space_collide2_matches_reference_build(odemi_mod):
    assert po.have_ref("single"), "the reference single build (oracle/_ref) is needed"
    ref = _run(_declare(po.ref_lib("single")))
    got = _run(_declare(odemi_mod.lib()))
    say("dSpaceCollide2 callback log:", len(ref), "entries;", sum(1 for e in ref if e[0] == "contacts"),
        "dCollide calls")
    assert got == ref


def test_dropin_static_space_frame(odemi_mod):
    """The ODE frame of an application with a static terrain space
    (dSpaceCollide2 of the car's space against the plane and ramps, then
    dSpaceCollide of the car's space; demos/raycar/ode.cpp:118-123) through
    the unchanged scene driver linked against this library: the callback
    pairs of every step equal the reference double build's, and the
    trajectory stays within the reference's FP32 envelope."""
    from test_golden import CASES
    from test_gpu_parity import _envelope
    fac, n, every, diag = CASES["vehicle_spaces"]
    sc = fac()
    g = golden("vehicle_spaces")
    lib = po.load(os.path.join(ROOT, "tests", "_build", "libscene_odemi.so"))
    r = po.run(lib, sc, n, every, diag_steps=diag, max_pairs=100000, max_contacts=100000)
    assert r["status"] == 0
    assert np.array_equal(r["pair_step_count"], g["pair_step_count"])
    assert np.array_equal(r["pairs"], g["pairs"])
    assert np.array_equal(r["contact_step_count"], g["contact_step_count"])
    env = _envelope(sc, n, every)
    err = np.abs(r["traj"][:, :, :3] - g["traj"][:, :, :3]).max(axis=(1, 2))
    say("static-space frame: pairs", len(r["pairs"]), "max traj err", float(err.max()))
    assert np.all(err <= 4 * env + 1e-3)


def _declare_rays(L):
    _declare(L)
    for name, res, args in [
            ("dCreateRay", V, [V, F]), ("dGeomRaySet", None, [V, F, F, F, F, F, F]),
            ("dGeomRayGet", None, [V, C.POINTER(F), C.POINTER(F)]), ("dGeomRaySetLength", None, [V, F]),
            ("dGeomRayGetLength", F, [V]), ("dGeomRaySetParams", None, [V, I, I]),
            ("dGeomRayGetParams", None, [V, C.POINTER(I), C.POINTER(I)]),
            ("dGeomRaySetClosestHit", None, [V, I]), ("dGeomRayGetClosestHit", I, [V]),
            ("dCreateCapsule", V, [V, F, F]), ("dGeomSetQuaternion", None, [V, C.POINTER(F)]),
            ("dGeomDestroy", None, [V]), ("dGeomGetAABB", None, [V, C.POINTER(F)]),
            ("dGeomGetBody", V, [V])]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _wheel_rays(L):
    """A car's wheel sensing (demos/raycar/car.cpp:351-380, 451-480): one ray
    geom outside every space, set per wheel with dGeomRaySet, collided with
    dSpaceCollide2 against a static terrain space (plane, boxes, capsules and
    a nested space); the callback skips the car's own body, descends into
    spaces and keeps the nearest contact of dCollide.  Returns the callback
    log, the nearest contact per wheel and the ray getters' answers."""
    L.dInitODE2(0)
    w = L.dWorldCreate()
    stat = L.dHashSpaceCreate(None)
    sub = L.dSimpleSpaceCreate(stat)
    names = {}

    def tag(g, name):
        names[g] = name
        L.dGeomSetData(g, C.c_void_p(len(names)))
        return g

    names[stat], names[sub] = "stat", "sub"
    rng = np.random.default_rng(31)
    tag(L.dCreatePlane(stat, 0.0, 0.0, 1.0, 0.0), "plane")
    for i in range(6):
        g = tag(L.dCreateBox(stat, *map(float, rng.uniform(0.5, 2.0, 3))), f"s{i}")
        L.dGeomSetPosition(g, *map(float, [*rng.uniform(-3, 3, 2), rng.uniform(-0.2, 0.4)]))
        q = rng.normal(size=4)
        L.dGeomSetQuaternion(g, (F * 4)(*map(float, q / np.linalg.norm(q))))
    for i in range(3):
        g = tag(L.dCreateCapsule(sub, float(rng.uniform(0.2, 0.5)), float(rng.uniform(0.5, 2.0))), f"c{i}")
        L.dGeomSetPosition(g, *map(float, [*rng.uniform(-3, 3, 2), 0.3]))
        q = rng.normal(size=4)
        L.dGeomSetQuaternion(g, (F * 4)(*map(float, q / np.linalg.norm(q))))
    car = L.dBodyCreate(w)
    L.dBodySetPosition(car, 0.0, 0.0, 1.0)
    chassis = tag(L.dCreateBox(stat, 1.0, 2.0, 0.4), "chassis")   # the car's own geom in the terrain space
    L.dGeomSetBody(chassis, car)
    ray = tag(L.dCreateRay(None, 1.0), "ray")
    log, nearest, getters = [], [], []
    buf = (C.c_byte * (128 * 20))()
    state = {}

    def near(data, o1, o2):
        log.append((names[o1], names[o2]))
        if L.dGeomGetBody(o1) == car or L.dGeomGetBody(o2) == car:
            return
        if L.dGeomIsSpace(o1) or L.dGeomIsSpace(o2):
            L.dSpaceCollide2(o1, o2, None, cb)
            return
        n = L.dCollide(o1, o2, 10, buf, 128)
        log.append(("contacts", n))
        raw = np.frombuffer(bytes(buf), dtype=np.uint8)
        for i in range(n):
            c = raw[128 * i:128 * i + 128]
            if state["dbl"]:
                v = np.frombuffer(c[:72].tobytes(), dtype=np.float64)
                pos, nrm = v[0:3], v[4:7]
            else:
                v = np.frombuffer(c[:36].tobytes(), dtype=np.float32).astype(np.float64)
                pos, nrm = v[0:3], v[4:7]
            dist = float(np.linalg.norm(pos - state["origin"]))
            if dist < state["best"][0]:
                state["best"] = (dist, pos.copy(), nrm.copy())

    cb = CB(near)
    state["dbl"] = False
    for frame in range(12):
        origin_c = np.array([*rng.uniform(-2.5, 2.5, 2), rng.uniform(0.6, 1.8)])
        up = rng.normal(size=3) * 0.25 + np.array([0, 0, 1.0])
        for wheel in range(4):
            pos = origin_c + np.array([(-1) ** wheel * 0.8, (-1) ** (wheel // 2) * 1.2, 0.0])
            state["origin"] = pos
            state["best"] = (1e4, None, None)
            L.dGeomRaySet(ray, *map(float, pos), *map(float, -up))
            L.dGeomRaySetClosestHit(ray, 1)
            length = float(rng.uniform(0.5, 2.5))
            L.dGeomRaySetLength(ray, length)
            L.dGeomRaySetParams(ray, 0, 0)
            log.append(("frame", frame, wheel))
            L.dSpaceCollide2(stat, ray, None, cb)
            nearest.append(state["best"])
            s, d = (F * 4)(), (F * 4)()
            fc, bc = I(), I()
            L.dGeomRayGetParams(ray, C.byref(fc), C.byref(bc))
            L.dGeomRayGet(ray, s, d)
            getters.append((L.dGeomRayGetLength(ray), L.dGeomRayGetClosestHit(ray), fc.value, bc.value,
                            *s[:3], *d[:3]))
    L.dGeomDestroy(ray)
    L.dSpaceDestroy(stat)
    L.dWorldDestroy(w)
    return log, nearest, getters


def test_wheel_rays_match_reference_build(odemi_mod):
    """Raycar's wheel rays through this library and through the reference's
    single-precision build with the same calls: identical callback logs
    (pairs in order, dCollide counts), identical hits and misses, nearest
    contacts within the FP32 tolerance, and the ray getters' answers."""
    assert po.have_ref("single"), "the reference single build (oracle/_ref) is needed"
    ref_log, ref_near, ref_get = _wheel_rays(_declare_rays(po.ref_lib("single")))
    got_log, got_near, got_get = _wheel_rays(_declare_rays(odemi_mod.lib()))
    hits = sum(1 for b in ref_near if b[1] is not None)
    say("wheel rays: callback entries", len(ref_log), "wheel casts", len(ref_near), "hits", hits)
    assert got_log == ref_log
    ref_get, got_get = np.array(ref_get), np.array(got_get)
    assert np.array_equal(got_get[:, :4], ref_get[:, :4])            # length, closest hit, params
    np.testing.assert_allclose(got_get[:, 4:], ref_get[:, 4:], atol=1e-6)   # start, unit direction
    assert hits >= 12
    for (dr, pr, nr), (dg, pg, ng) in zip(ref_near, got_near):
        assert (pr is None) == (pg is None)
        if pr is not None:
            assert abs(dr - dg) < 2e-4
            np.testing.assert_allclose(pg, pr, atol=2e-4)
            np.testing.assert_allclose(ng, nr, atol=2e-4)
