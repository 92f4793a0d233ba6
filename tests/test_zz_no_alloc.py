"""dCollideNoAlloc against dCollide through the C ABI: the same contacts, bit
for bit, on every collider pair type, on consecutive calls whose second one
asks for fewer contacts (the shared scratch array keeps the last call's
contents), and the refusal of a request above the array's 8 contacts."""
import ctypes as C

import numpy as np
import pytest

from conftest import say

pytestmark = pytest.mark.gpu

V, F = C.c_void_p, C.c_float


def _api():
    import odemi
    L = odemi.lib()
    for name, res, args in [
            ("dInitODE2", C.c_int, [C.c_uint]),
            ("dCreateBox", V, [V, F, F, F]), ("dCreateSphere", V, [V, F]), ("dCreatePlane", V, [V, F, F, F, F]),
            ("dGeomSetPosition", None, [V, F, F, F]), ("dGeomSetRotation", None, [V, C.POINTER(F)]),
            ("dCollide", C.c_int, [V, V, C.c_int, V, C.c_int]),
            ("dCollideNoAlloc", C.c_int, [V, V, C.c_int, V, C.c_int]), ("dGeomDestroy", None, [V])]:
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    R4 = np.zeros((3, 4))
    R4[:, :3] = R
    return (F * 12)(*R4.reshape(-1).tolist())


def _pairs(L):
    """Overlapping pairs of every type combination: (g1, g2, contacts expected > 0)."""
    rng = np.random.default_rng(4)
    out = []
    box = lambda s, p: (L.dCreateBox(None, *s), p)
    sph = lambda r, p: (L.dCreateSphere(None, r), p)
    made = [box((1.0, 1.2, 0.8), (0.0, 0.0, 0.0)), box((0.9, 1.0, 1.1), (0.3, 0.2, 0.85)),   # box-box
            sph(0.5, (0.0, 0.0, 3.0)), box((1.0, 1.0, 1.0), (0.2, 0.1, 3.6)),                # sphere-box
            sph(0.4, (5.0, 0.0, 0.0)), sph(0.5, (5.5, 0.2, 0.3))]                            # sphere-sphere
    for g, p in made:
        L.dGeomSetPosition(g, *p)
        L.dGeomSetRotation(g, _rot(rng))
    plane = L.dCreatePlane(None, 0.0, 0.0, 1.0, 0.4)
    flat = L.dCreateBox(None, 2.0, 2.0, 0.5)                                                 # box-plane, 4 points
    L.dGeomSetPosition(flat, 8.0, 0.0, 0.5)
    ball = L.dCreateSphere(None, 0.3)                                                       # sphere-plane
    L.dGeomSetPosition(ball, 10.0, 0.0, 0.5)
    g = [m[0] for m in made]
    out = [(g[0], g[1]), (g[2], g[3]), (g[4], g[5]), (flat, plane), (ball, plane)]
    return out, g + [plane, flat, ball]


def _call(fn, g1, g2, maxc, nbuf=8):
    buf = (C.c_byte * (64 * nbuf))()
    C.memset(buf, 0x5A, 64 * nbuf)          # sentinel: rows past the count stay untouched
    n = fn(g1, g2, maxc, buf, 64)
    return n, bytes(buf)


def test_no_alloc_collide_matches_dcollide(capfd):
    L = _api()
    L.dInitODE2(0)
    pairs, geoms = _pairs(L)
    for k, (g1, g2) in enumerate(pairs):
        for maxc in (8, 3):                  # second call: a smaller request
            say(f"dCollideNoAlloc pair {k}, {maxc} contacts requested")
            n_ref, b_ref = _call(L.dCollide, g1, g2, maxc)
            n, b = _call(L.dCollideNoAlloc, g1, g2, maxc)
            assert n_ref > 0 and n == n_ref, (k, maxc, n, n_ref)
            assert b == b_ref, (k, maxc)
    # consecutive calls, the second on a pair with fewer contacts: nothing of
    # the first call's contacts shows through
    n1, _ = _call(L.dCollideNoAlloc, *pairs[3], 8)
    n2, b2 = _call(L.dCollideNoAlloc, *pairs[4], 8)
    n2r, b2r = _call(L.dCollide, *pairs[4], 8)
    assert n1 > n2 == n2r == 1 and b2 == b2r
    # a request above the scratch array's 8 contacts is refused
    capfd.readouterr()
    n, _ = _call(L.dCollideNoAlloc, *pairs[0], 9, nbuf=9)
    assert n == 0
    assert "at most 8" in capfd.readouterr().err
    for g in geoms:
        L.dGeomDestroy(g)
