"""Writes cylinder_colliders.npz: cylinder contacts from the reference ODE 0.12
double build (oracle/_ref/libode_ref_double.so, built from the reference sources by
oracle/ref.mk), through its own dCollide, for seeded cylinders against
spheres, boxes, planes and rays in both argument orders
(collision_cylinder_sphere.cpp, collision_cylinder_plane.cpp,
collision_cylinder_box.cpp, ray.cpp:500-656; the colliders table of
collision_kernel.cpp:196-213).  Cylinder-cylinder and capsule-cylinder pairs
have no collider in this configuration (no libccd): both give no contacts,
which the cases include.

The cases cover: spheres at the side, at a disc and at a rim; a cylinder
standing on or lying on a plane and tilted (the parallel-disc branch and the
deepest-point branch); boxes against the side (the clipped side segment) and
against a disc (the box face clipped to the disc's 8-gon), each with
maxc 1, 2, 4 and 8; rays through the caps and the side.  Each row of `cases`
holds  t1, t2, maxc, p1[3], q1[4], s1[4], p2[3], q2[4], s2[4]  and
`contacts[i]` the reference's contacts (up to 8) of case i: pos[3],
normal[3], depth, with `count[i]` of them.

Run from the repository root:  python tests/golden/make_cylinders.py
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from make_rays import ContactGeom, _geom as _geom_basic, _lib, quat_for_dir, rot  # noqa: E402

SPHERE, BOX, PLANE, CAPSULE, RAY, CYLINDER = 0, 1, 2, 3, 4, 5


def _geom(L, t, s, p, q):
    if t != CYLINDER:
        return _geom_basic(L, t, s, p, q)
    g = L.dCreateCylinder(None, s[0], s[1])
    L.dGeomSetPosition(g, *map(float, p))
    R4 = np.zeros((3, 4))
    R4[:, :3] = rot(np.asarray(q, float))
    L.dGeomSetRotation(g, np.ascontiguousarray(R4.reshape(12)).ctypes.data_as(C.POINTER(C.c_double)))
    return g


def cylinder_cases(seed=53):
    """(t_other, maxc, cyl p, cyl q, cyl size, other p, other q, other size)"""
    rng = np.random.default_rng(seed)
    out = []

    def unit_q():
        q = rng.normal(size=4)
        return q / np.linalg.norm(q)

    for trial in range(150):          # spheres: side, disc and rim regions
        rad, ln = rng.uniform(0.2, 0.7), rng.uniform(0.3, 1.5)
        qc = unit_q()
        R = rot(qc)
        kind = trial % 3
        perp = np.cross(R[:, 2], rng.normal(size=3))
        perp /= np.linalg.norm(perp)
        r2 = rng.uniform(0.1, 0.6)
        if kind == 0:       # side
            ps = R[:, 2] * rng.uniform(-ln / 2, ln / 2) + perp * (rad + r2 * rng.uniform(-0.9, 0.9))
        elif kind == 1:     # disc
            ps = R[:, 2] * np.sign(rng.normal()) * (ln / 2 + r2 * rng.uniform(-0.9, 0.9)) + perp * rng.uniform(0, rad)
        else:               # rim
            ps = (R[:, 2] * np.sign(rng.normal()) * (ln / 2 + r2 * rng.uniform(0.05, 0.6))
                  + perp * (rad + r2 * rng.uniform(0.05, 0.6)))
        pc = rng.uniform(-0.3, 0.3, 3)
        out.append((SPHERE, 1, pc, qc, [rad, ln, 0, 0], pc + ps, np.array([1.0, 0, 0, 0]), [r2, 0, 0, 0]))
    for trial in range(120):          # planes
        rad, ln = rng.uniform(0.2, 0.7), rng.uniform(0.3, 1.5)
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        kind = trial % 3
        if kind == 0:       # a disc parallel to the plane (standing)
            qc = quat_for_dir(n * np.sign(rng.normal()))
        elif kind == 1:     # lying on its side
            qc = quat_for_dir(np.cross(n, rng.normal(size=3)))
        else:
            qc = unit_q()
        R = rot(qc)
        ext = np.abs(R[:, 2] @ n) * ln / 2 + rad * np.sqrt(max(0.0, 1 - (R[:, 2] @ n) ** 2))
        pc = n * (ext - rng.uniform(0.0, 0.2)) + rng.uniform(-1, 1, 3) - n * (rng.uniform(-1, 1, 3) @ n)
        d = float(rng.uniform(-0.2, 0.2))
        pc = pc + n * d
        maxc = [1, 2, 3, 4, 8][trial % 5]
        out.append((PLANE, maxc, pc, qc, [rad, ln, 0, 0], np.zeros(3), np.array([1.0, 0, 0, 0]), [*n, d]))
    for trial in range(240):          # boxes
        rad, ln = rng.uniform(0.2, 0.7), rng.uniform(0.3, 1.5)
        qb = unit_q()
        Rb = rot(qb)
        side = rng.uniform(0.4, 1.6, 3)
        kind = trial % 4
        face = rng.integers(3)
        sgn = np.sign(rng.normal())
        nrm = Rb[:, face] * sgn
        if kind == 0:       # disc on a box face
            qc = quat_for_dir(-nrm * np.sign(rng.normal()))
            off = ln / 2
        elif kind == 1:     # side on a box face
            qc = quat_for_dir(np.cross(nrm, rng.normal(size=3)))
            off = rad
        elif kind == 2:     # tilted disc on a face (the 0.9 boundary region)
            qc = quat_for_dir(-nrm + rng.normal(size=3) * 0.3)
            off = ln / 2
        else:
            qc = unit_q()
            off = rad
        pc = nrm * (side[face] / 2 + off - rng.uniform(0.0, 0.15)) + Rb @ rng.uniform(-0.3, 0.3, 3) * 0.5
        maxc = [1, 2, 4, 8][trial % 4]
        out.append((BOX, maxc, pc, qc, [rad, ln, 0, 0], np.zeros(3), qb, [*side, 0]))
    for trial in range(120):          # rays
        rad, ln = rng.uniform(0.2, 0.7), rng.uniform(0.3, 1.5)
        qc = unit_q()
        R = rot(qc)
        kind = trial % 4
        pc = rng.uniform(-0.3, 0.3, 3)
        if kind == 0:       # inside, along the axis (the caps branch)
            perp = np.cross(R[:, 2], rng.normal(size=3))
            perp /= np.linalg.norm(perp)
            pr = pc + perp * rng.uniform(0, 0.9) * rad + R[:, 2] * rng.uniform(-ln, ln)
            d = R[:, 2] * np.sign(rng.normal())
        elif kind == 1:
            pr = pc + rng.uniform(-0.1, 0.1, 3)
            d = rng.normal(size=3)
        else:
            pr = pc + rng.normal(size=3) * 1.5
            d = pc + rng.uniform(-0.3, 0.3, 3) - pr
        out.append((RAY, 1, pc, qc, [rad, ln, 0, 0], pr, quat_for_dir(d), [rng.uniform(0.3, 3.0), 0, 0, 0]))
    for trial in range(20):           # pairs without a collider
        t = CYLINDER if trial % 2 == 0 else CAPSULE
        pc = rng.uniform(-0.1, 0.1, 3)
        out.append((t, 4, pc, unit_q(), [0.4, 1.0, 0, 0], pc + rng.uniform(-0.2, 0.2, 3), unit_q(), [0.4, 1.0, 0, 0]))
    return out


def make():
    L = _lib()
    vp = C.c_void_p
    L.dCreateCylinder.restype = vp
    L.dCreateCylinder.argtypes = [vp, C.c_double, C.c_double]
    rows, cnt, con = [], [], []
    buf = (ContactGeom * 8)()
    for (t, maxc, pc, qc, sc, po_, qo, so) in cylinder_cases():
        for order in (0, 1):
            a = (CYLINDER, pc, qc, sc)
            b = (t, po_, qo, so)
            x, y = (a, b) if order == 0 else (b, a)
            g1 = _geom(L, x[0], x[3], x[1], x[2])
            g2 = _geom(L, y[0], y[3], y[1], y[2])
            n = L.dCollide(g1, g2, maxc, buf, C.sizeof(ContactGeom))
            L.dGeomDestroy(g1)
            L.dGeomDestroy(g2)
            rows.append([x[0], y[0], maxc, *x[1], *x[2], *x[3], *y[1], *y[2], *y[3]])
            cnt.append(n)
            c = np.zeros((8, 7))
            for i in range(n):
                c[i, :3] = buf[i].pos[:3]
                c[i, 3:6] = buf[i].normal[:3]
                c[i, 6] = buf[i].depth
            con.append(c)
    cnt = np.array(cnt)
    np.savez_compressed(os.path.join(HERE, "cylinder_colliders.npz"), cases=np.array(rows), count=cnt,
                        contacts=np.array(con))
    print("cylinder cases", len(cnt), "with contacts", int((cnt > 0).sum()), "contacts", int(cnt.sum()),
          "by count", np.bincount(cnt))


if __name__ == "__main__":
    make()
