"""Pins the CPU restatement (oracle/) before it is trusted as the checker:
known-answer tests of the reference, closed forms, invariants and brute force
on small inputs, each chosen so that a plausible mistake fails it."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import pyoracle as po

DP = C.POINTER(C.c_double)


def lib(p="f64"):
    L = po.oracle_lib(p)
    L.orc_rand_test.restype = C.c_int
    L.orc_rand_ints.restype = C.c_ulong
    L.orc_rand_ints.argtypes = [C.c_ulong, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.orc_collide.restype = C.c_int
    L.orc_collide.argtypes = [C.c_int, DP, DP, DP, C.c_int, DP, DP, DP, C.c_int, DP]
    L.orc_aabb.argtypes = [C.c_int, DP, DP, DP, DP]
    L.orc_plane_space.argtypes = [DP, DP, DP]
    return L


def d(a):
    a = np.ascontiguousarray(a, np.float64)
    return a, a.ctypes.data_as(DP)


def rot(q):
    w, x, y, z = q / np.linalg.norm(q)
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    R4 = np.zeros((3, 4))
    R4[:, :3] = R
    return R4.reshape(12), R


def collide(t1, p1, q1, s1, t2, p2, q2, s2, maxc=4, prec="f64"):
    L = lib(prec)
    R1, _ = rot(np.asarray(q1, float))
    R2, _ = rot(np.asarray(q2, float))
    bufs = [d(p1), d(R1), d(s1), d(p2), d(R2), d(s2)]
    out = np.zeros((max(maxc, 1), 7))
    n = L.orc_collide(t1, bufs[0][1], bufs[1][1], bufs[2][1], t2, bufs[3][1], bufs[4][1], bufs[5][1], maxc,
                      out.ctypes.data_as(DP))
    return out[:n]


SPHERE, BOX, PLANE = 0, 1, 2
ID = [1.0, 0, 0, 0]


def test_rand_known_answer():
    # dTestRand: seed 0 -> 0x3c6ef35f, 0x47502932, 0xd1ccf6e9, 0xaaf95334, 0x6252e503 (ode/src/misc.cpp:52-62)
    assert lib().orc_rand_test() == 1


def test_rand_int_matches_lcg_closed_form():
    # dRandInt(n) = xor-folded LCG value mod n (misc.cpp:66-117), restated here independently
    bounds = np.array([1, 2, 3, 4, 5, 16, 17, 200, 256, 257, 5000, 65536, 70000], np.int32)
    out = np.zeros_like(bounds)
    lib().orc_rand_ints(12345, len(bounds), bounds.ctypes.data_as(C.POINTER(C.c_int)),
                        out.ctypes.data_as(C.POINTER(C.c_int)))
    s = 12345
    for n, got in zip(bounds.tolist(), out.tolist()):
        s = (1664525 * s + 1013904223) & 0xFFFFFFFF
        r = s
        if n <= 16:
            r ^= r >> 16; r ^= r >> 8; r ^= r >> 4
            if n <= 2:
                r ^= r >> 2; r ^= r >> 1
            elif n <= 4:
                r ^= r >> 2
        elif n <= 256:
            r ^= r >> 16; r ^= r >> 8
        elif n <= 65536:
            r ^= r >> 16
        assert got == r % n


def test_plane_space_orthonormal():
    # testPlaneSpace invariant (ode/demo/demo_ode.cpp:219-235)
    rng = np.random.default_rng(0)
    L = lib()
    for _ in range(200):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        p = np.zeros(3); q = np.zeros(3)
        L.orc_plane_space(d(n)[1], p.ctypes.data_as(DP), q.ctypes.data_as(DP))
        M = np.stack([n, p, q])
        assert np.allclose(M @ M.T, np.eye(3), atol=1e-12)
        assert np.allclose(np.cross(n, p), q, atol=1e-12)


def test_sphere_sphere_closed_form():
    rng = np.random.default_rng(1)
    for _ in range(100):
        p1, p2 = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        r1, r2 = rng.uniform(0.1, 1, 2)
        c = collide(SPHERE, p1, ID, [r1, 0, 0, 0], SPHERE, p2, ID, [r2, 0, 0, 0])
        dist = np.linalg.norm(p1 - p2)
        if dist > r1 + r2:
            assert len(c) == 0
            continue
        n = (p1 - p2) / dist
        assert np.allclose(c[0, 3:6], n) and math.isclose(c[0, 6], r1 + r2 - dist, abs_tol=1e-12)
        assert np.allclose(c[0, :3], p1 + n * 0.5 * (r2 - r1 - dist))


def test_sphere_plane_and_reversed_dispatch():
    rng = np.random.default_rng(2)
    for _ in range(100):
        n = rng.normal(size=3); n /= np.linalg.norm(n)
        off = rng.uniform(-1, 1)
        p = rng.uniform(-2, 2, 3); r = rng.uniform(0.1, 1)
        c = collide(SPHERE, p, ID, [r, 0, 0, 0], PLANE, [0, 0, 0], ID, [*n, off])
        depth = off - n @ p + r
        if depth < 0:
            assert len(c) == 0
            continue
        assert np.allclose(c[0, 3:6], n) and math.isclose(c[0, 6], depth, abs_tol=1e-12)
        assert np.allclose(c[0, :3], p - r * n)
        # plane first: the normal is negated (collision_kernel.cpp:319-333)
        c2 = collide(PLANE, [0, 0, 0], ID, [*n, off], SPHERE, p, ID, [r, 0, 0, 0])
        assert np.allclose(c2[0, 3:6], -n) and np.allclose(c2[0, :3], c[0, :3])


def test_box_plane_against_vertex_enumeration():
    # the deepest contact is the deepest box vertex; every contact is a box
    # vertex whose depth is its distance below the plane
    rng = np.random.default_rng(3)
    n_checked = 0
    for _ in range(300):
        q = rng.normal(size=4)
        side = rng.uniform(0.2, 2, 3)
        p = np.array([0, 0, rng.uniform(-0.5, 1.0)])
        c = collide(BOX, p, q, [*side, 0], PLANE, [0, 0, 0], ID, [0, 0, 1, 0], maxc=4)
        _, R = rot(q)
        verts = np.array([p + R @ (np.array(s) * side / 2) for s in np.ndindex(2, 2, 2)]) if False else \
            np.array([p + R @ (np.array([sx, sy, sz]) * side / 2) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
        depths = -verts[:, 2]
        if depths.max() < 0:
            assert len(c) == 0
            continue
        n_checked += 1
        assert math.isclose(c[0, 6], depths.max(), abs_tol=1e-9)
        for row in c:
            k = np.argmin(np.linalg.norm(verts - row[:3], axis=1))
            assert np.linalg.norm(verts[k] - row[:3]) < 1e-9
            assert math.isclose(row[6], depths[k], abs_tol=1e-9) and row[6] >= -1e-12
            assert np.allclose(row[3:6], [0, 0, 1])
    assert n_checked > 50


def test_sphere_box_closed_form():
    rng = np.random.default_rng(4)
    for _ in range(300):
        q = rng.normal(size=4)
        side = rng.uniform(0.3, 2, 3)
        _, R = rot(q)
        ps = rng.uniform(-1.5, 1.5, 3)
        r = rng.uniform(0.05, 0.6)
        c = collide(SPHERE, ps, ID, [r, 0, 0, 0], BOX, [0, 0, 0], q, [*side, 0])
        local = R.T @ ps
        clamped = np.clip(local, -side / 2, side / 2)
        if np.all(np.abs(local) < side / 2):
            face = np.argmin(side / 2 - np.abs(local))
            assert len(c) == 1 and math.isclose(c[0, 6], side[face] / 2 - abs(local[face]) + r, abs_tol=1e-9)
            assert np.allclose(c[0, 3:6], np.sign(local[face]) * R[:, face])
            continue
        dist = np.linalg.norm(local - clamped)
        if dist > r:
            assert len(c) == 0
            continue
        assert math.isclose(c[0, 6], r - dist, abs_tol=1e-9)
        assert np.allclose(c[0, :3], R @ clamped, atol=1e-9)


def _sat_depth(p1, R1, s1, p2, R2, s2):
    """Brute-force minimum penetration over the 15 separating axes."""
    axes = [R1[:, i] for i in range(3)] + [R2[:, i] for i in range(3)]
    axes += [np.cross(R1[:, i], R2[:, j]) for i in range(3) for j in range(3)]
    best = np.inf
    for a in axes:
        l = np.linalg.norm(a)
        if l < 1e-9:
            continue
        a = a / l
        r1 = sum(abs(a @ R1[:, i]) * s1[i] / 2 for i in range(3))
        r2 = sum(abs(a @ R2[:, i]) * s2[i] / 2 for i in range(3))
        pen = r1 + r2 - abs(a @ (p2 - p1))
        if pen < 0:
            return -1.0
        best = min(best, pen)
    return best


def test_box_box_against_brute_force_sat():
    rng = np.random.default_rng(5)
    hits = 0
    for _ in range(400):
        q1, q2 = rng.normal(size=4), rng.normal(size=4)
        s1, s2 = rng.uniform(0.3, 1.5, 3), rng.uniform(0.3, 1.5, 3)
        p1 = np.zeros(3); p2 = rng.uniform(-1.2, 1.2, 3)
        c = collide(BOX, p1, q1, [*s1, 0], BOX, p2, q2, [*s2, 0], maxc=8)
        _, R1 = rot(q1)
        _, R2 = rot(q2)
        pen = _sat_depth(p1, R1, s1, p2, R2, s2)
        if pen < 0:
            assert len(c) == 0
            continue
        if len(c) == 0:
            continue  # the face clipping may cull every point for grazing contacts
        hits += 1
        assert 1 <= len(c) <= 8
        assert np.allclose(np.linalg.norm(c[:, 3:6], axis=1), 1)
        assert np.all(c[:, 6] >= -1e-9)
        # the reported maximum depth is at least the true minimum penetration
        # along the chosen axis (edge axes are preferred only within 1.05x)
        assert c[:, 6].max() <= pen * 1.06 + 1e-9 or len(c) > 1
        # the normal points into box 1: moving box 1 along it separates
        n = c[0, 3:6]
        assert n @ (p1 - p2) >= -1e-9 or len(c) > 1
    assert hits > 100


def test_aabb_closed_form():
    rng = np.random.default_rng(6)
    L = lib()
    for _ in range(100):
        q = rng.normal(size=4)
        side = rng.uniform(0.1, 2, 3)
        p = rng.uniform(-5, 5, 3)
        R4, R = rot(q)
        out = np.zeros(6)
        L.orc_aabb(BOX, d(p)[1], d(R4)[1], d([*side, 0])[1], out.ctypes.data_as(DP))
        verts = np.array([p + R @ (np.array([sx, sy, sz]) * side / 2) for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)])
        assert np.allclose(out[0::2], verts.min(0)) and np.allclose(out[1::2], verts.max(0))


def test_tiny_lcp_contact_against_closed_form():
    # one unit box resting on the plane: after a step the normal impulses
    # carry its weight, the velocity normal to the plane stays ~0
    from odemi import scenes
    sc = scenes.box_stack(1, gap=-1e-4)
    r = po.run(lib(), sc, 1, 1, diag_steps=1, max_pairs=16, max_contacts=16, feedback=True)
    fz = r["feedback"][:, 2].sum()
    assert math.isclose(fz, 9.81 * 1.0, rel_tol=0.05), fz
    assert abs(r["traj"][1][0, 9]) < 0.05


def test_free_fall_closed_form():
    # no contacts: v = g t, x = x0 + sum of h v (semi-implicit Euler)
    from odemi import scenes
    sc = scenes.box_stack(1)
    sc.body_pos[0, 2] = 50.0
    h, n = sc.dt, 25
    r = po.run(lib(), sc, n, n)
    vz = -9.81 * h * n
    z = 50.0 + sum(-9.81 * h * (k + 1) * h for k in range(n))
    assert math.isclose(r["traj"][1][0, 9], vz, rel_tol=1e-12)
    assert math.isclose(r["traj"][1][0, 2], z, rel_tol=1e-12)


def test_coloured_order_converges_to_the_reference_orders_solution():
    """The coloured-order restatement (the order of the MI355X coloured
    mode) and the reference order solve the same LCP: with the solver run to
    convergence, one step from the same state gives the same contact
    impulses (row order only changes the iterates)."""
    from odemi import scenes
    lib64 = po.oracle_lib("f64")
    rng = np.random.default_rng(12)
    for seed in (1, 2):
        sc = scenes.pile_world(rng, 24, 2, seed=seed)
        # advance to a state with resting contacts, then one converged step each way
        st = po.run(lib64, sc, 60, 60)["traj"][-1]
        sc.iterations = 100000
        sc.body_pos, sc.body_q = st[:, :3].copy(), st[:, 3:7].copy()
        sc.body_lvel, sc.body_avel = st[:, 7:10].copy(), st[:, 10:13].copy()
        try:
            lib64.orc_set_colored(1)
            tc = po.tape_run(lib64, sc, 1)
        finally:
            lib64.orc_set_colored(0)
        tr = po.tape_run(lib64, sc, 1)
        assert tc["nk"][0] == tr["nk"][0] > 0
        # contacts come in different orders: match them by (g1, g2, position)
        def keyed(con):
            out = {}
            for c in con:
                g = (min(int(c[10]), int(c[11])), max(int(c[10]), int(c[11])))
                out.setdefault(g, []).append(c)
            return out
        kc, kr = keyed(tc["contacts"]), keyed(tr["contacts"])
        assert kc.keys() == kr.keys()
        scale = max(1.0, float(np.abs(tr["contacts"][:, 7]).max()))
        for g in kr:
            for c in kc[g]:
                d = [np.abs(c[:3] - r[:3]).max() for r in kr[g]]
                r = kr[g][int(np.argmin(d))]
                assert min(d) < 1e-9
                # the normal may be flipped when the pair order differs: compare the
                # normal impulse (sign-free) and the friction impulse magnitude
                assert abs(c[7] - r[7]) <= 1e-6 * scale, (g, c[7], r[7])
                assert abs(np.hypot(c[8], c[9]) - np.hypot(r[8], r[9])) <= 1e-6 * scale
        # and the bodies end in the same state
        fc, fr = tc["states"][-1], tr["states"][-1]
        assert np.abs(fc[:, 7:13] - fr[:, 7:13]).max() <= 1e-6 * max(1.0, np.abs(fr[:, 7:13]).max())


def test_warm_start_changes_only_the_starting_point():
    """Warm start (coloured mode) starts each solve from the previous step's
    impulses instead of zero.  The frictionless LCP (regularised by cfm, so
    its solution is unique) is unchanged: solved to 1e-10, warm and cold runs
    agree to round-off after the warm-started steps, and at 1e-6 the warm run
    needs fewer sweeps on a resting pile.  A start whose fc does not match
    its impulses converges to a different solution and fails the first
    check."""
    import ctypes as C
    from odemi import scenes
    lib64 = po.oracle_lib("f64")
    lib64.orc_set_colored_solver.argtypes = [C.c_double, C.c_int, C.c_int, C.c_int]
    lib64.orc_set_warm_start.argtypes = [C.c_double]
    lib64.orc_colored_sweeps.restype = C.c_longlong
    lib64.orc_colored_unconverged.restype = C.c_longlong
    rng = np.random.default_rng(21)
    sc = scenes.pile_world(rng, 24, 2, seed=3)
    sc.mu = 0.0
    st = po.run(lib64, sc, 80, 80)["traj"][-1]
    sc.body_pos, sc.body_q = st[:, :3].copy(), st[:, 3:7].copy()
    sc.body_lvel, sc.body_avel = st[:, 7:10].copy(), st[:, 10:13].copy()
    out = {}
    try:
        lib64.orc_set_colored(1)
        for tol, steps, cap in ((1e-10, 3, 1000000), (1e-6, 15, 400000)):
            lib64.orc_set_colored_solver(tol, cap, 10, 1)
            for warm in (0.0, 1.0):
                lib64.orc_set_warm_start(warm)
                lib64.orc_colored_sweeps()
                lib64.orc_colored_unconverged()
                t = po.run(lib64, sc, steps, steps)["traj"][-1]
                out[tol, warm] = (t, lib64.orc_colored_sweeps(), lib64.orc_colored_unconverged())
    finally:
        lib64.orc_set_colored(0)
        lib64.orc_set_colored_solver(0.0, 0, 10, 0)
        lib64.orc_set_warm_start(0.0)
    (tc, _, uc), (tw, _, uw) = out[1e-10, 0.0], out[1e-10, 1.0]
    assert uc == uw == 0
    assert np.abs(tw[:, :3] - tc[:, :3]).max() < 1e-9
    assert np.abs(tw[:, 7:13] - tc[:, 7:13]).max() < 1e-7
    sw_c, sw_w = out[1e-6, 0.0][1], out[1e-6, 1.0][1]
    assert out[1e-6, 0.0][2] == out[1e-6, 1.0][2] == 0
    assert sw_w < 0.85 * sw_c, (sw_w, sw_c)


RAY, CAPSULE = 4, 3


def _ray_fixture():
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rays.npz"))
    return z["cases"], z["count"], z["contacts"]


def _case(row):
    t1, t2 = int(row[0]), int(row[1])
    return t1, row[2:5], row[5:9], row[9:13], t2, row[13:16], row[16:20], row[20:24]


def test_ray_colliders_match_the_reference_build():
    """The restatement's ray colliders (ray-sphere, -box, -capsule, -plane in
    both argument orders, ray.cpp:196-497) against the reference's own
    dCollide in double precision (tests/golden/rays.npz, written by
    make_rays.py): the same hit or miss in every case, contacts equal to
    rounding."""
    cases, count, contacts = _ray_fixture()
    for i, row in enumerate(cases):
        t1, p1, q1, s1, t2, p2, q2, s2 = _case(row)
        got = collide(t1, p1, q1, s1, t2, p2, q2, s2, 4, "f64")
        assert len(got) == count[i], (i, t1, t2, len(got), count[i])
        if count[i]:
            np.testing.assert_allclose(got[0, :7], contacts[i], rtol=0, atol=1e-9, err_msg=str(i))


def test_ray_contacts_lie_on_the_surface_hit():
    """What the geometry fixes for every reference ray hit: the contact is
    start + depth * direction with 0 <= depth <= length, it lies on the
    surface of the shape (sphere: |x - c| = r; plane: n.x = d; box: on a face;
    capsule: at distance r from its segment), and the normal is a unit vector
    (the surface normal, reversed when the ray starts inside)."""
    cases, count, contacts = _ray_fixture()
    n_hit = 0
    for i, row in enumerate(cases):
        if not count[i]:
            continue
        t1, p1, q1, s1, t2, p2, q2, s2 = _case(row)
        if t1 == RAY:
            pr, qr, length, t, po_, qo, s = p1, q1, s1[0], t2, p2, q2, s2
        else:
            pr, qr, length, t, po_, qo, s = p2, q2, s2[0], t1, p1, q1, s1
        _, Rr = rot(qr)
        _, Ro = rot(qo)
        x, nrm, depth = contacts[i, :3], contacts[i, 3:6], contacts[i, 6]
        assert -1e-12 <= depth <= length + 1e-12
        np.testing.assert_allclose(x, pr + depth * Rr[:, 2], atol=1e-9)
        assert abs(np.linalg.norm(nrm) - 1) < 1e-9
        if t == 0:
            assert abs(np.linalg.norm(x - po_) - s[0]) < 1e-9
        elif t == 2:
            assert abs(np.dot(s[:3], x) - s[3]) < 1e-9
            assert abs(abs(np.dot(nrm, s[:3])) - 1) < 1e-9
        elif t == 1:
            loc = Ro.T @ (x - po_)
            h = np.asarray(s[:3]) / 2
            assert np.all(np.abs(loc) <= h + 1e-9)
            assert np.min(h - np.abs(loc)) < 1e-9
        else:
            ax = Ro[:, 2]
            k = np.clip(np.dot(x - po_, ax), -s[1] / 2, s[1] / 2)
            assert abs(np.linalg.norm(x - (po_ + k * ax)) - s[0]) < 1e-9
        n_hit += 1
    assert n_hit > 300


def _cyl_fixture():
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "cylinder_colliders.npz"))
    return z["cases"], z["count"], z["contacts"]


def _cyl_case(row):
    t1, t2, maxc = int(row[0]), int(row[1]), int(row[2])
    return t1, row[3:6], row[6:10], row[10:14], t2, row[14:17], row[17:21], row[21:25], maxc


def test_cylinder_colliders_match_the_reference_build():
    """The restatement's cylinder colliders (cylinder-sphere, -plane, -box,
    ray-cylinder, both argument orders, maxc 1-8; cylinder-cylinder and
    capsule-cylinder without a collider) against the reference's own dCollide
    in double precision (tests/golden/cylinder_colliders.npz, written by
    make_cylinders.py): the same number of contacts in every case and the
    same contacts to rounding."""
    cases, count, contacts = _cyl_fixture()
    for i, row in enumerate(cases):
        t1, p1, q1, s1, t2, p2, q2, s2, maxc = _cyl_case(row)
        got = collide(t1, p1, q1, s1, t2, p2, q2, s2, maxc, "f64")
        assert len(got) == count[i], (i, t1, t2, maxc, len(got), count[i])
        if count[i]:
            np.testing.assert_allclose(got[:, :7], contacts[i, :count[i]], rtol=0, atol=1e-9, err_msg=str(i))


def test_cylinder_contacts_against_the_geometry():
    """What the geometry fixes for the reference's cylinder contacts: a
    sphere's contact depth equals the sphere radius minus the distance from
    its centre to the cylinder surface (side, disc or rim, computed here from
    the closest point of the solid cylinder), a plane contact's depth is the
    point's depth below the plane, and every normal is a unit vector."""
    cases, count, contacts = _cyl_fixture()
    n_sph = n_pl = 0
    for i, row in enumerate(cases):
        if not count[i]:
            continue
        t1, p1, q1, s1, t2, p2, q2, s2, maxc = _cyl_case(row)
        for k in range(count[i]):
            assert abs(np.linalg.norm(contacts[i, k, 3:6]) - 1) < 1e-6
        other = t2 if t1 == 5 else t1
        pc, qc, sc = (p1, q1, s1) if t1 == 5 else (p2, q2, s2)
        po_, so = (p2, s2) if t1 == 5 else (p1, s1)
        if other == 0:
            _, Rc = rot(qc)
            loc = Rc.T @ (po_ - pc)
            rad, hl = sc[0], sc[1] / 2
            rr = np.hypot(loc[0], loc[1])
            inside = rr <= rad and abs(loc[2]) <= hl
            if inside:
                continue
            cl = np.array([loc[0], loc[1], 0.0]) * (min(rr, rad) / rr if rr > 0 else 0.0)
            cl[2] = np.clip(loc[2], -hl, hl)
            dist = np.linalg.norm(loc - cl)
            assert abs(contacts[i, 0, 6] - (so[0] - dist)) < 1e-9, i
            n_sph += 1
        elif other == 2:
            for k in range(count[i]):
                x = contacts[i, k, :3]
                assert abs(contacts[i, k, 6] - (so[3] - np.dot(so[:3], x))) < 1e-9
            n_pl += 1
    assert n_sph > 60 and n_pl > 60


def _blcp(A, b, lo, hi, prec="f64"):
    L = lib(prec)
    L.orc_blcp.restype = C.c_int
    L.orc_blcp.argtypes = [C.c_int, DP, DP, DP, DP, DP]
    n = len(b)
    bufs = [d(A), d(b), d(lo), d(hi)]
    x = np.zeros(n)
    r = L.orc_blcp(n, *[q[1] for q in bufs], x.ctypes.data_as(DP))
    return r, x


def test_exact_lcp_against_enumeration():
    """The dWorldStep restatement's box-constrained LCP solver on random
    positive definite problems (n = 1..6, some rows unbounded, some with
    lo = 0, some with finite two-sided bounds): its answer satisfies the
    complementarity conditions, and equals the one solution found by
    enumerating all 3^n assignments of each row to its lower bound, its
    upper bound or the free set (each solved by a linear solve)."""
    import itertools
    rng = np.random.default_rng(5)
    for trial in range(150):
        n = int(rng.integers(1, 7))
        M = rng.normal(size=(n, n))
        A = M @ M.T + 0.05 * np.eye(n)
        b = rng.normal(size=n) * 3
        kind = rng.integers(3, size=n)
        lo = np.where(kind == 0, -np.inf, np.where(kind == 1, 0.0, -rng.uniform(0.1, 2, n)))
        hi = np.where(kind == 0, np.inf, np.where(kind == 1, np.inf, rng.uniform(0.1, 2, n)))
        r, x = _blcp(A, b, lo, hi)
        assert r >= 0, trial
        w = A @ x - b
        tol = 1e-9 * (1 + np.abs(b).max())
        assert np.all(x >= lo - tol) and np.all(x <= hi + tol)
        for i in range(n):
            if x[i] <= lo[i] + tol:
                assert w[i] >= -1e-7
            elif x[i] >= hi[i] - tol:
                assert w[i] <= 1e-7
            else:
                assert abs(w[i]) <= 1e-7
        sols = []
        for assign in itertools.product(range(3), repeat=n):
            if any((a == 1 and not np.isfinite(lo[i])) or (a == 2 and not np.isfinite(hi[i]))
                   for i, a in enumerate(assign)):
                continue
            xs = np.where(np.array(assign) == 1, lo, np.where(np.array(assign) == 2, hi, 0.0))
            fr = [i for i in range(n) if assign[i] == 0]
            fixed = [i for i in range(n) if assign[i] != 0]
            if fr:
                rhs = b[fr] - A[np.ix_(fr, fixed)] @ xs[fixed] if fixed else b[fr]
                xs[fr] = np.linalg.solve(A[np.ix_(fr, fr)], rhs)
            ws = A @ xs - b
            ok = all((assign[i] == 0 and lo[i] - 1e-9 <= xs[i] <= hi[i] + 1e-9) or
                     (assign[i] == 1 and ws[i] >= -1e-9) or (assign[i] == 2 and ws[i] <= 1e-9) for i in range(n))
            if ok:
                sols.append(xs)
        assert sols, trial
        for s in sols:
            np.testing.assert_allclose(s, sols[0], atol=1e-7)
        np.testing.assert_allclose(x, sols[0], atol=1e-7)


def test_coloured_order_caps_colours_at_the_workgroup_width():
    """The coloured mode's colouring (Jones-Plassmann rounds, first fit)
    caps a colour at the workgroup width, so a colour is one pass of the
    block.  On 144 boxes resting on a plane (four plane contacts each):
    every colour is a proper colouring (no body twice); width 256 gives four
    colours of 144, width 64 none above 64 and fewer 64-lane passes; rounds
    rise and, inside a round, constraint ids rise (the assignment order);
    and each constraint took the lowest colour its bodies had not taken and
    that was not full at its turn (first fit)."""
    from odemi import scenes
    L = po.oracle_lib("f64")
    L.orc_colour_log.restype = C.c_longlong
    L.orc_colour_log.argtypes = [C.c_void_p, C.c_longlong]
    sc = scenes.box_floor(12)
    logs = {}
    try:
        L.orc_set_colored(1)
        L.orc_set_colour_log(1)
        for width in (256, 64):
            L.orc_set_colour_width(width)
            po.run(L, sc, 1, 20)
            n = L.orc_colour_log(None, 0)
            buf = np.zeros(5 * n, np.int32)
            L.orc_colour_log(buf.ctypes.data, n)
            logs[width] = buf.reshape(n, 5)
    finally:
        L.orc_set_colour_log(0)
        L.orc_set_colour_width(256)
        L.orc_set_colored(0)
    passes = {}
    for width, log in logs.items():
        cid, b1, b2, col, rnd = (log[:, k] for k in range(5))
        assert len(log) == 4 * 144
        for c in np.unique(col):
            bodies = np.concatenate([b1[col == c], b2[(col == c) & (b2 >= 0)]])
            assert len(bodies) == len(np.unique(bodies))
        sizes = np.bincount(col)
        assert sizes.max() <= width
        passes[width] = int(sum((s + 63) // 64 for s in sizes))
        order = rnd.astype(np.int64) * (1 << 32) + cid
        assert np.all(np.diff(order) > 0)
        used = {}
        cnt = np.zeros(64, int)
        for i in range(len(log)):
            taken = used.get(b1[i], set()) | (used.get(b2[i], set()) if b2[i] >= 0 else set())
            c = int(col[i])
            assert c not in taken and cnt[c] < width
            assert all(k in taken or cnt[k] >= width for k in range(c))
            cnt[c] += 1
            for b in (b1[i], b2[i]):
                if b >= 0:
                    used.setdefault(b, set()).add(c)
    assert sorted(np.bincount(logs[256][:, 3])) == [144] * 4
    assert passes[64] < passes[256]
