"""Writes rays.npz: ray contacts from the reference ODE 0.12 double build
(oracle/_ref/libode_ref_double.so, built from the reference sources by
oracle/ref.mk), through its own dCollide, for seeded rays against spheres,
boxes, capsules and planes in both argument orders (ray.cpp:196-497 and the
colliders table of collision_kernel.cpp:181-194).

The cases cover a ray starting outside and inside each shape, rays parallel
to a capsule's axis (the end-cap branch, ray.cpp:382-400), rays too short to
reach, and rays pointing away.  Each row of `cases` holds
  t1, t2, p1[3], q1[4], s1[4], p2[3], q2[4], s2[4]
and `contacts[i]` the reference's contacts of case i (at most 1 for a ray):
pos[3], normal[3], depth, with `count[i]` of them.

Run from the repository root:  python tests/golden/make_rays.py
"""
import ctypes as C
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import pyoracle as po  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
SPHERE, BOX, PLANE, CAPSULE, RAY = 0, 1, 2, 3, 4


class ContactGeom(C.Structure):
    # dContactGeom of the double build (collision.h:70-80)
    _fields_ = [("pos", C.c_double * 4), ("normal", C.c_double * 4), ("depth", C.c_double),
                ("g1", C.c_void_p), ("g2", C.c_void_p), ("side1", C.c_int), ("side2", C.c_int)]


def rot(q):
    w, x, y, z = q / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def quat_for_dir(d):
    """A unit quaternion whose rotation's column 2 is the unit vector d."""
    d = d / np.linalg.norm(d)
    z = np.array([0.0, 0.0, 1.0])
    c = float(np.dot(z, d))
    if c < -0.999999:
        return np.array([0.0, 1.0, 0.0, 0.0])
    ax = np.cross(z, d)
    q = np.array([1.0 + c, *ax])
    return q / np.linalg.norm(q)


def ray_cases(seed=41):
    """(t_other, ray p, ray q, length, other p, other q, other size) per case."""
    rng = np.random.default_rng(seed)
    out = []
    for t in (SPHERE, BOX, CAPSULE, PLANE):
        for trial in range(120):
            if t == SPHERE:
                s = [rng.uniform(0.2, 0.8), 0, 0, 0]
            elif t == BOX:
                s = [*rng.uniform(0.3, 1.5, 3), 0]
            elif t == CAPSULE:
                s = [rng.uniform(0.15, 0.5), rng.uniform(0.3, 1.5), 0, 0]
            else:
                n = rng.normal(size=3)
                n /= np.linalg.norm(n)
                s = [*n, rng.uniform(-0.5, 0.5)]
            po_ = rng.uniform(-0.3, 0.3, 3) if t != PLANE else np.zeros(3)
            qo = rng.normal(size=4) if t != PLANE else np.array([1.0, 0, 0, 0])
            kind = trial % 6
            if kind == 0:        # from inside (near the centre)
                pr = po_ + rng.uniform(-0.1, 0.1, 3)
                d = rng.normal(size=3)
            elif kind == 1 and t == CAPSULE:   # parallel to the capsule axis
                ax = rot(qo)[:, 2]
                perp = np.cross(ax, rng.normal(size=3))
                perp /= np.linalg.norm(perp)
                off = perp * rng.uniform(0.0, 0.9) * s[0]
                sgn = 1.0 if trial % 12 < 6 else -1.0
                pr = po_ + off + ax * (sgn * rng.uniform(0.0, 1.5))
                d = -sgn * ax if trial % 4 else sgn * ax
            elif kind == 5:      # pointing away
                pr = po_ + rng.normal(size=3) * 2.0
                d = pr - po_
            else:                # from outside towards a point near the centre
                pr = po_ + rng.normal(size=3) * 1.5
                d = (po_ + rng.uniform(-0.4, 0.4, 3)) - pr
            length = rng.uniform(0.3, 3.5)
            out.append((t, pr, quat_for_dir(d), length, po_, qo, s))
    return out


def _lib():
    L = po.ref_lib("double")
    vp = C.c_void_p
    L.dCreateRay.restype = vp; L.dCreateRay.argtypes = [vp, C.c_double]
    L.dCreateSphere.restype = vp; L.dCreateSphere.argtypes = [vp, C.c_double]
    L.dCreateBox.restype = vp; L.dCreateBox.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    L.dCreateCapsule.restype = vp; L.dCreateCapsule.argtypes = [vp, C.c_double, C.c_double]
    L.dCreatePlane.restype = vp; L.dCreatePlane.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double]
    L.dGeomSetPosition.argtypes = [vp, C.c_double, C.c_double, C.c_double]
    L.dGeomSetRotation.argtypes = [vp, C.POINTER(C.c_double)]
    L.dGeomDestroy.argtypes = [vp]
    L.dCollide.restype = C.c_int
    L.dCollide.argtypes = [vp, vp, C.c_int, C.POINTER(ContactGeom), C.c_int]
    L.dInitODE2.argtypes = [C.c_uint]
    L.dInitODE2(0)
    return L


def _geom(L, t, s, p, q):
    if t == SPHERE:
        g = L.dCreateSphere(None, s[0])
    elif t == BOX:
        g = L.dCreateBox(None, s[0], s[1], s[2])
    elif t == CAPSULE:
        g = L.dCreateCapsule(None, s[0], s[1])
    elif t == RAY:
        g = L.dCreateRay(None, s[0])
    else:
        return L.dCreatePlane(None, s[0], s[1], s[2], s[3])
    L.dGeomSetPosition(g, *map(float, p))
    R4 = np.zeros((3, 4))
    R4[:, :3] = rot(np.asarray(q, float))
    L.dGeomSetRotation(g, np.ascontiguousarray(R4.reshape(12)).ctypes.data_as(C.POINTER(C.c_double)))
    return g


def make():
    L = _lib()
    rows, cnt, con = [], [], []
    buf = (ContactGeom * 4)()
    for (t, pr, qr, length, po_, qo, s) in ray_cases():
        for order in (0, 1):
            a = (RAY, pr, qr, [length, 0, 0, 0])
            b = (t, po_, qo, s)
            x, y = (a, b) if order == 0 else (b, a)
            g1 = _geom(L, x[0], x[3], x[1], x[2])
            g2 = _geom(L, y[0], y[3], y[1], y[2])
            n = L.dCollide(g1, g2, 4, buf, C.sizeof(ContactGeom))
            L.dGeomDestroy(g1)
            L.dGeomDestroy(g2)
            rows.append([x[0], y[0], *x[1], *x[2], *x[3], *y[1], *y[2], *y[3]])
            cnt.append(n)
            c = np.zeros(7)
            if n:
                c[:3] = buf[0].pos[:3]
                c[3:6] = buf[0].normal[:3]
                c[6] = buf[0].depth
            con.append(c)
    cnt = np.array(cnt)
    np.savez_compressed(os.path.join(HERE, "rays.npz"), cases=np.array(rows), count=cnt, contacts=np.array(con))
    print("rays cases", len(cnt), "hits", int((cnt > 0).sum()))


if __name__ == "__main__":
    make()
