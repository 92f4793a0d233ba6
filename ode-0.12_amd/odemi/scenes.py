"""Seeded synthetic workloads for the ODE step path.

Each generator returns ``Scene`` objects (one per world) built from a seed:

* ``box_stack``      -- 10 unit boxes stacked on a z=0 plane with 1 mm gaps.
* ``random_pile``    -- one world of 4096 boxes/spheres dropped in a column.
* ``uneven_worlds``  -- many independent piles of Poisson-distributed size with
  ball/hinge chains.
* ``vehicle_worlds`` -- four-wheeled cars on hinge2 suspensions over ramps.
* ``diff_piles``     -- small piles for the adjoint.
* ``joint_zoo``      -- hinge stops and motors, ball, hinge2 steering stops,
  spin motor and suspension.
* ``surface_pile``   -- a pile with every contact surface option (mu2,
  bounce, soft ERP/CFM), or frictionless.

Bodies, geoms and joints are listed in creation order; the order matters for
the reference's list-order effects (ode/src/ode.cpp:62-68 prepends bodies and
joints, ode/src/collision_kernel.h:166-171 prepends geoms to a space).
Masses follow dMassSetBoxTotal / dMassSetSphereTotal (ode/src/mass.cpp:118-221).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

SPHERE, BOX, PLANE, CAPSULE, CYLINDER = 0, 1, 2, 3, 5
BALL, HINGE, HINGE2 = 1, 2, 3

# contact surface mode bits (include/ode/contact.h:33-49)
MU2, FDIR1, BOUNCE, SOFT_ERP, SOFT_CFM = 0x001, 0x002, 0x004, 0x008, 0x010
APPROX1 = 0x3000

# joint_param columns
(JP_LO, JP_HI, JP_VEL, JP_FMAX, JP_FUDGE, JP_BOUNCE, JP_CFM, JP_STOP_ERP,
 JP_STOP_CFM, JP_SUSP_ERP, JP_SUSP_CFM, JP_VEL2, JP_FMAX2, JP_ERP) = range(14)
NJP = 14


@dataclass
class Scene:
    """One world: parameters plus per-body/geom/joint arrays (float64)."""
    gravity: tuple = (0.0, 0.0, -9.81)
    erp: float = 0.2
    cfm: float = 1e-5
    iterations: int = 20
    sor_w: float = 1.3
    max_vel: float = math.inf
    min_depth: float = 0.0
    dt: float = 0.01
    seed: int = 0
    hash_minlevel: int = -3
    hash_maxlevel: int = 10
    surf_mode: int = APPROX1
    mu: float = 0.5
    mu2: float = 0.0
    bounce: float = 0.0
    bounce_vel: float = 0.0
    soft_erp: float = 0.0
    soft_cfm: float = 0.0
    max_contacts: int = 4
    skip_connected: int = 1
    body_pos: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    body_q: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
    body_lvel: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    body_avel: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    body_mass: np.ndarray = field(default_factory=lambda: np.zeros((0,)))
    body_I: np.ndarray = field(default_factory=lambda: np.zeros((0, 9)))
    geom_type: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    geom_body: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    geom_param: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
    geom_pos: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    geom_q: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
    joint_type: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    joint_b1: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    joint_b2: np.ndarray = field(default_factory=lambda: np.zeros((0,), np.int32))
    joint_anchor: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    joint_axis1: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    joint_axis2: np.ndarray = field(default_factory=lambda: np.zeros((0, 3)))
    joint_param: np.ndarray = field(default_factory=lambda: np.zeros((0, NJP)))
    step_ctrl: np.ndarray | None = None   # (nsteps, nj, 4): lostop, histop, torque1, torque2
    geom_space: np.ndarray | None = None  # per geom 0 (dynamic space) or 1 (static terrain space)
    world_step: int = 0                   # 1: dWorldStep (exact LCP) instead of dWorldQuickStep

    @property
    def nb(self) -> int:
        return int(self.body_pos.shape[0])

    @property
    def ng(self) -> int:
        return int(self.geom_type.shape[0])

    @property
    def nj(self) -> int:
        return int(self.joint_type.shape[0])


class _Builder:
    """Accumulates bodies/geoms/joints in creation order."""

    def __init__(self):
        self.b_pos, self.b_q, self.b_mass, self.b_I = [], [], [], []
        self.b_lvel, self.b_avel = [], []
        self.g_type, self.g_body, self.g_param, self.g_pos, self.g_q = [], [], [], [], []
        self.j_type, self.j_b1, self.j_b2 = [], [], []
        self.j_anchor, self.j_axis1, self.j_axis2, self.j_param = [], [], [], []

    def body(self, pos, mass, I, q=(1.0, 0.0, 0.0, 0.0), lvel=(0.0, 0.0, 0.0), avel=(0.0, 0.0, 0.0)) -> int:
        self.b_pos.append(pos)
        self.b_q.append(q)
        self.b_mass.append(mass)
        self.b_I.append(np.asarray(I, np.float64).reshape(9))
        self.b_lvel.append(lvel)
        self.b_avel.append(avel)
        return len(self.b_pos) - 1

    def box_body(self, pos, sides, density, **kw) -> int:
        lx, ly, lz = sides
        m = lx * ly * lz * density
        I = np.diag([m / 12.0 * (ly * ly + lz * lz), m / 12.0 * (lx * lx + lz * lz), m / 12.0 * (lx * lx + ly * ly)])
        b = self.body(pos, m, I, **kw)
        self.geom(BOX, (lx, ly, lz, 0.0), b)
        return b

    def sphere_body(self, pos, r, density, **kw) -> int:
        m = (4.0 / 3.0) * math.pi * r * r * r * density
        I = np.eye(3) * (0.4 * m * r * r)
        b = self.body(pos, m, I, **kw)
        self.geom(SPHERE, (r, 0.0, 0.0, 0.0), b)
        return b

    def capsule_body(self, pos, radius, length, density, **kw) -> int:
        """A capsule along its body z axis; mass and inertia of
        dMassSetCapsule (mass.cpp:141-162)."""
        m1 = math.pi * radius * radius * length * density
        m2 = (4.0 / 3.0) * math.pi * radius ** 3 * density
        ia = m1 * (0.25 * radius * radius + length * length / 12.0) + \
            m2 * (0.4 * radius * radius + 0.375 * radius * length + 0.25 * length * length)
        ib = (m1 * 0.5 + m2 * 0.4) * radius * radius
        b = self.body(pos, m1 + m2, np.diag([ia, ia, ib]), **kw)
        self.geom(CAPSULE, (radius, length, 0.0, 0.0), b)
        return b

    def cylinder_body(self, pos, radius, length, density, **kw) -> int:
        """A solid cylinder along its body z axis; mass and inertia of
        dMassSetCylinder (mass.cpp:173-198)."""
        m = math.pi * radius * radius * length * density
        i = m * (0.25 * radius * radius + length * length / 12.0)
        b = self.body(pos, m, np.diag([i, i, m * 0.5 * radius * radius]), **kw)
        self.geom(CYLINDER, (radius, length, 0.0, 0.0), b)
        return b

    def geom(self, gtype, param, body=-1, pos=(0.0, 0.0, 0.0), q=(1.0, 0.0, 0.0, 0.0)) -> int:
        self.g_type.append(gtype)
        self.g_body.append(body)
        self.g_param.append(param)
        self.g_pos.append(pos)
        self.g_q.append(q)
        return len(self.g_type) - 1

    def joint(self, jtype, b1, b2, anchor, axis1=(1.0, 0.0, 0.0), axis2=(0.0, 1.0, 0.0), **params) -> int:
        p = np.full(NJP, np.nan)
        for k, v in params.items():
            p[globals()["JP_" + k.upper()]] = v
        self.j_type.append(jtype)
        self.j_b1.append(b1)
        self.j_b2.append(b2)
        self.j_anchor.append(anchor)
        self.j_axis1.append(axis1)
        self.j_axis2.append(axis2)
        self.j_param.append(p)
        return len(self.j_type) - 1

    def scene(self, **kw) -> Scene:
        f = lambda x, shape: np.asarray(x, np.float64).reshape(shape)
        i = lambda x: np.asarray(x, np.int32).reshape(-1)
        return Scene(
            body_pos=f(self.b_pos, (-1, 3)), body_q=f(self.b_q, (-1, 4)),
            body_lvel=f(self.b_lvel, (-1, 3)), body_avel=f(self.b_avel, (-1, 3)),
            body_mass=f(self.b_mass, (-1,)), body_I=f(self.b_I, (-1, 9)),
            geom_type=i(self.g_type), geom_body=i(self.g_body), geom_param=f(self.g_param, (-1, 4)),
            geom_pos=f(self.g_pos, (-1, 3)), geom_q=f(self.g_q, (-1, 4)),
            joint_type=i(self.j_type), joint_b1=i(self.j_b1), joint_b2=i(self.j_b2),
            joint_anchor=f(self.j_anchor, (-1, 3)), joint_axis1=f(self.j_axis1, (-1, 3)),
            joint_axis2=f(self.j_axis2, (-1, 3)), joint_param=f(self.j_param, (-1, NJP)),
            **kw)


def joint_zoo(seed: int = 5) -> Scene:
    """Joint limits, stops and motors over a plane: a powered hinge to the
    world with stops (motor 2 rad/s, fmax 3, stops +-0.4 rad), a hinge with
    bouncy soft stops below it, a ball joint, and a wheel on a hinge2 with a
    locked steering axis (stops 0.1..0.1), a spin motor (vel2 8 rad/s, fmax2
    15) and a soft suspension; h=0.01, mu=0.8."""
    b = _Builder()
    b.geom(PLANE, (0.0, 0.0, 1.0, 0.0))
    a = b.box_body((0.0, 0.0, 1.6), (0.6, 0.2, 0.2), 1.0)
    c = b.box_body((0.0, 0.0, 1.1), (0.2, 0.2, 0.6), 1.0)
    d = b.sphere_body((0.0, 0.0, 0.65), 0.15, 1.0)
    b.joint(HINGE, a, -1, (0.0, 0.0, 1.6), axis1=(0.0, 0.0, 1.0), lo=-0.4, hi=0.4, vel=2.0, fmax=3.0)
    b.joint(HINGE, c, a, (0.0, 0.0, 1.4), axis1=(1.0, 0.0, 0.0), lo=-0.3, hi=0.25, bounce=0.5, stop_erp=0.6,
            stop_cfm=1e-4)
    b.joint(BALL, d, c, (0.0, 0.0, 0.8))
    ch = b.box_body((2.0, 0.0, 0.6), (1.0, 0.6, 0.2), 2.0)
    wh = b.sphere_body((2.0, 0.45, 0.35), 0.3, 1.0)
    b.joint(HINGE2, ch, wh, (2.0, 0.45, 0.35), axis1=(0.0, 0.0, 1.0), axis2=(0.0, 1.0, 0.0), lo=0.1, hi=0.1,
            vel2=8.0, fmax2=15.0, susp_erp=0.4, susp_cfm=0.02)
    return b.scene(dt=0.01, mu=0.8, seed=seed)


def surface_pile(seed: int = 13, mu: float = 0.7, mu2: float = 0.3) -> Scene:
    """A 40-body pile with every contact surface option of the batch policy:
    dContactMu2 (mu along the first friction direction, mu2 along the
    second), dContactBounce (0.3 above 0.05 m/s), dContactSoftERP (0.5) and
    dContactSoftCFM (1e-4), dContactApprox1; mu = 0 gives frictionless
    contacts (one row each)."""
    sc = pile_world(np.random.default_rng(seed), 40, 0, seed=seed)
    sc.surf_mode = APPROX1 | MU2 | BOUNCE | SOFT_ERP | SOFT_CFM
    sc.mu, sc.mu2 = mu, mu2
    sc.bounce, sc.bounce_vel = 0.3, 0.05
    sc.soft_erp, sc.soft_cfm = 0.5, 1e-4
    return sc


def capsule_pile(n: int = 30, seed: int = 19) -> Scene:
    """Capsules (radius U[0.1,0.3], length U[0.2,1.0], random orientation)
    with spheres and boxes dropped on a plane, spawned without overlap of
    their bounding spheres; h = 0.01, mu = 0.6.  Every capsule collider of the
    reference (capsule-sphere, -box, -capsule, -plane) meets its pairs."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    b.geom(PLANE, (0.0, 0.0, 1.0, 0.0))
    clear = _Clearance()
    half = 0.3 * math.sqrt(n)
    for i in range(n):
        dens = rng.uniform(0.5, 2.0)
        kind = i % 3
        if kind == 0:
            r, ln = rng.uniform(0.1, 0.3), rng.uniform(0.2, 1.0)
            rad = r + 0.5 * ln
        elif kind == 1:
            r = rng.uniform(0.1, 0.4)
            rad = r
        else:
            he = rng.uniform(0.1, 0.4, size=3)
            rad = float(np.linalg.norm(he))
        for _ in range(200):
            pos = (rng.uniform(-half, half), rng.uniform(-half, half), rng.uniform(0.4, 0.4 + n / 8.0))
            if clear.free([pos], [rad]):
                break
        clear.add([pos], [rad])
        q = rng.normal(size=4)
        q = tuple(q / np.linalg.norm(q))
        if kind == 0:
            b.capsule_body(pos, r, ln, dens, q=q)
        elif kind == 1:
            b.sphere_body(pos, r, dens)
        else:
            b.box_body(pos, tuple(2.0 * he), dens, q=q)
    return b.scene(dt=0.01, mu=0.6, seed=seed)


def with_world_step(sc: Scene) -> Scene:
    """The scene stepped with dWorldStep (the island's LCP solved exactly)
    instead of dWorldQuickStep."""
    sc.world_step = 1
    return sc


def cylinder_pile(n: int = 32, seed: int = 31) -> Scene:
    """Cylinders (radius U[0.15,0.35], length U[0.2,0.8]; half of them upright
    or on their side, the rest at random orientations) with spheres and boxes
    dropped on a plane, spawned without overlap of their bounding spheres;
    h = 0.01, mu = 0.6.  Every cylinder collider of the reference without
    libccd (cylinder-sphere, -box, -plane) meets its pairs; two cylinders have
    no collider there and pass through each other, here as in the reference."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    b.geom(PLANE, (0.0, 0.0, 1.0, 0.0))
    clear = _Clearance()
    half = 0.3 * math.sqrt(n)
    for i in range(n):
        dens = rng.uniform(0.5, 2.0)
        kind = i % 3
        if kind == 0:
            r, ln = rng.uniform(0.15, 0.35), rng.uniform(0.2, 0.8)
            rad = math.hypot(r, 0.5 * ln)
        elif kind == 1:
            r = rng.uniform(0.1, 0.4)
            rad = r
        else:
            he = rng.uniform(0.1, 0.4, size=3)
            rad = float(np.linalg.norm(he))
        for _ in range(200):
            pos = (rng.uniform(-half, half), rng.uniform(-half, half), rng.uniform(0.4, 0.4 + n / 8.0))
            if clear.free([pos], [rad]):
                break
        clear.add([pos], [rad])
        q = rng.normal(size=4)
        q = tuple(q / np.linalg.norm(q))
        if kind == 0:
            if i % 4 == 0:
                q = (1.0, 0.0, 0.0, 0.0)                       # upright: a disc on the plane
            elif i % 4 == 2:
                q = (math.sqrt(0.5), math.sqrt(0.5), 0.0, 0.0)  # on its side
            b.cylinder_body(pos, r, ln, dens, q=q)
        elif kind == 1:
            b.sphere_body(pos, r, dens)
        else:
            b.box_body(pos, tuple(2.0 * he), dens, q=q)
    return b.scene(dt=0.01, mu=0.6, seed=seed)


def box_floor(n: int = 12, seed: int = 1) -> Scene:
    """n x n boxes 0.5 m on a side, density 1, resting 1 mm deep on a z=0
    plane with 0.1 m between them; h = 0.01, mu = 0.5.  Every box has four
    contacts with the plane from the first step, so a graph colouring has
    four colours of n*n members each unless their size is capped."""
    b = _Builder()
    b.geom(PLANE, (0.0, 0.0, 1.0, 0.0))
    for k in range(n * n):
        b.box_body((0.6 * (k % n), 0.6 * (k // n), 0.249), (0.5, 0.5, 0.5), 1.0)
    return b.scene(dt=0.01, mu=0.5, seed=seed)


def box_stack(n: int = 10, gap: float = 1e-3, seed: int = 0) -> Scene:
    """10 boxes 1x1x1 m, density 1, on a z=0 plane with 1 mm gaps; h=0.01,
    20 iterations, ERP 0.2, CFM 1e-5, mu=0.5 (Coulomb friction through findex)."""
    b = _Builder()
    b.geom(PLANE, (0.0, 0.0, 1.0, 0.0))
    for i in range(n):
        b.box_body((0.0, 0.0, 0.5 + gap + i * (1.0 + gap)), (1.0, 1.0, 1.0), 1.0)
    return b.scene(dt=0.01, mu=0.5, seed=seed)


class _Clearance:
    """Bounding spheres of the bodies placed so far, binned in a uniform grid,
    for spawning without initial interpenetration (rejection of candidate
    poses that overlap)."""

    CELL = 2.0   # > twice the largest bounding radius (a box of half-extents 0.5: 0.87)

    def __init__(self):
        self.cells: dict = {}

    def _key(self, p):
        return (math.floor(p[0] / self.CELL), math.floor(p[1] / self.CELL), math.floor(p[2] / self.CELL))

    def free(self, centres, radii) -> bool:
        for p, r in zip(centres, radii):
            kx, ky, kz = self._key(p)
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        for (q, s) in self.cells.get((kx + dx, ky + dy, kz + dz), ()):
                            ex, ey, ez = p[0] - q[0], p[1] - q[1], p[2] - q[2]
                            if ex * ex + ey * ey + ez * ez <= (r + s) * (r + s):
                                return False
        return True

    def add(self, centres, radii):
        for p, r in zip(centres, radii):
            self.cells.setdefault(self._key(p), []).append((tuple(p), r))


def _pile_bodies(b: _Builder, rng: np.random.Generator, nb: int, half_w: float, z_lo: float, z_hi: float,
                 x0: float = 0.0, y0: float = 0.0, clear: _Clearance | None = None):
    """Even index: box with half-extents U[0.1,0.5]^3; odd: sphere r~U[0.1,0.5];
    density U[0.5,2]; position uniform in the column; identity orientation.
    With `clear`, a body whose bounding sphere would overlap one placed before
    draws its position again (up to 200 times, then keeps the last draw)."""
    for i in range(nb):
        dens = rng.uniform(0.5, 2.0)
        if clear is None:
            # draw order of the fixtures: density, position, size
            pos = (x0 + rng.uniform(-half_w, half_w), y0 + rng.uniform(-half_w, half_w), rng.uniform(z_lo, z_hi))
        if i % 2 == 0:
            he = rng.uniform(0.1, 0.5, size=3)
            rad = float(np.linalg.norm(he))
        else:
            r = rng.uniform(0.1, 0.5)
            rad = r
        if clear is not None:
            for _ in range(200):
                pos = (x0 + rng.uniform(-half_w, half_w), y0 + rng.uniform(-half_w, half_w),
                       rng.uniform(z_lo, z_hi))
                if clear.free([pos], [rad]):
                    break
            clear.add([pos], [rad])
        if i % 2 == 0:
            b.box_body(pos, tuple(2.0 * he), dens)
        else:
            b.sphere_body(pos, r, dens)


def random_pile(nb: int = 4096, seed: int = 12345) -> Scene:
    """One world: nb bodies in a 20x20x40 m column (z in [0.5,40.5]) over a
    plane; mu=0.6, h=0.005, 20 iterations, default hash levels -3..10."""
    rng = np.random.default_rng(seed)
    b = _Builder()
    b.geom(PLANE, (0.0, 0.0, 1.0, 0.0))
    _pile_bodies(b, rng, nb, 10.0, 0.5, 40.5)
    return b.scene(dt=0.005, mu=0.6, seed=seed)


def _chain(b: _Builder, rng: np.random.Generator, n_links: int, start, link_len=0.3, link_w=0.1, pitch=0.27):
    """Horizontal chain of boxes along x, joined alternately by ball and hinge
    joints at the middle of the overlap of neighbouring links (hinge axis y).
    Neighbours overlap by link_len - pitch so that their bounding boxes
    intersect by a clear margin rather than touching exactly (an exact touch
    is decided by the last bit of the arithmetic)."""
    x0, y0, z0 = start
    links = []
    for k in range(n_links):
        links.append(b.box_body((x0 + pitch * (k + 0.5), y0, z0), (link_len, link_w, link_w), 1.0))
    for k in range(n_links - 1):
        anchor = (x0 + pitch * (k + 1), y0, z0)
        if k % 2 == 0:
            b.joint(BALL, links[k], links[k + 1], anchor)
        else:
            b.joint(HINGE, links[k], links[k + 1], anchor, axis1=(0.0, 1.0, 0.0))


def pile_world(rng: np.random.Generator, nb: int, n_chains: int, chain_len_lo: int = 4, chain_len_hi: int = 12,
               dt: float = 0.01, mu: float = 0.6, seed: int = 0, spawn_clear: bool = False) -> Scene:
    """One small world: `nb` pile bodies in a column of the same body density as
    the 4096-body pile (footprint half-width 0.35*sqrt(nb), height nb/8 m),
    followed by `n_chains` link chains; chain links are extra bodies.

    spawn_clear: no body or chain link starts inside another (rejection
    sampling on bounding spheres).  Uniform spawns interpenetrate by up to a
    metre; the contact ERP then throws bodies apart at tens of m/s, and a
    thin box spun up that way grows its spin through the explicit gyroscopic
    term of the integrator until it overflows -- the reference's own FP64
    build does that on 4 of the 4096 uniform-spawn worlds within 200 steps."""
    b = _Builder()
    b.geom(PLANE, (0.0, 0.0, 1.0, 0.0))
    half_w = 0.35 * math.sqrt(nb)
    clear = _Clearance() if spawn_clear else None
    _pile_bodies(b, rng, nb, half_w, 0.5, 0.5 + nb / 8.0, clear=clear)
    link_r = math.sqrt(0.15 ** 2 + 2 * 0.05 ** 2)
    for _ in range(n_chains):
        n_links = int(rng.integers(chain_len_lo, chain_len_hi + 1))
        for _ in range(200 if spawn_clear else 1):
            start = (rng.uniform(-half_w, half_w - 0.3 * n_links), rng.uniform(-half_w, half_w),
                     rng.uniform(0.5, 0.5 + nb / 8.0))
            cen = [(start[0] + 0.27 * (k + 0.5), start[1], start[2]) for k in range(n_links)]
            if clear is None or clear.free(cen, [link_r] * n_links):
                break
        if clear is not None:
            clear.add(cen, [link_r] * n_links)
        _chain(b, rng, n_links, start)
    return b.scene(dt=dt, mu=mu, seed=seed)


def uneven_worlds(n_worlds: int = 4096, seed: int = 7, mean: float = 64.0, lo: int = 8, hi: int = 256,
                  max_chains: int = 16, dt: float = 0.005, sor_w: float = 1.0,
                  spawn_clear: bool = True) -> list[Scene]:
    """Independent worlds with body count clamp(Poisson(64), 8, 256) plus
    U{0..16} ball/hinge chains of U{4..12} links; the piles' step parameters
    are those of the 4096-body pile: h=0.005, 20 iterations, mu=0.6.  The
    relaxation factor is 1.0: these worlds are solved in the coloured order,
    where over-relaxation (ODE's default 1.3) lets some chain-heavy worlds
    gain energy from step to step (the CPU restatement of the coloured
    order shows the same).  Bodies and chains spawn without interpenetration
    (see pile_world)."""
    rng = np.random.default_rng(seed)
    worlds = []
    for w in range(n_worlds):
        nb = int(np.clip(rng.poisson(mean), lo, hi))
        nch = int(rng.integers(0, max_chains + 1))
        sc = pile_world(rng, nb, nch, dt=dt, seed=int(rng.integers(0, 2 ** 31)), spawn_clear=spawn_clear)
        sc.sor_w = sor_w
        worlds.append(sc)
    return worlds


def vehicle_world(rng: np.random.Generator, nsteps: int, n_ramps: int = 32, seed: int = 0,
                  static_space: bool = False) -> Scene:
    """A 4x2x0.5 m, 1000 kg chassis on four r=0.35 m, 20 kg sphere wheels with
    hinge2 joints (axis1 = z steering, axis2 = y spin, as demo_buggy.cpp:
    271-278); suspension ERP/CFM from k=3e4 N/m, c=2e3 N s/m at h=0.01; 32
    static ramp boxes; per-step steer stop U[-0.5,0.5] rad on the front wheels
    and wheel torque U[-500,500] N m; mu=1.0.  static_space: the plane and
    the ramps sit in a static space of their own."""
    h = 0.01
    kp, kd = 3e4, 2e3
    susp_erp = h * kp / (h * kp + kd)
    susp_cfm = 1.0 / (h * kp + kd)
    b = _Builder()
    b.geom(PLANE, (0.0, 0.0, 1.0, 0.0))
    z0 = 1.0
    lx, ly, lz = 4.0, 2.0, 0.5
    m = 1000.0
    I = np.diag([m / 12.0 * (ly * ly + lz * lz), m / 12.0 * (lx * lx + lz * lz), m / 12.0 * (lx * lx + ly * ly)])
    chassis = b.body((0.0, 0.0, z0), m, I)
    b.geom(BOX, (lx, ly, lz, 0.0), chassis)
    wheels = []
    r = 0.35
    for wx, wy in ((1.5, 1.2), (1.5, -1.2), (-1.5, 1.2), (-1.5, -1.2)):
        mw = 20.0
        wb = b.body((wx, wy, z0 - 0.3), mw, np.eye(3) * (0.4 * mw * r * r))
        b.geom(SPHERE, (r, 0.0, 0.0, 0.0), wb)
        wheels.append(wb)
    for k, wb in enumerate(wheels):
        wx, wy = (1.5, 1.5, -1.5, -1.5)[k], (1.2, -1.2, 1.2, -1.2)[k]
        b.joint(HINGE2, chassis, wb, (wx, wy, z0 - 0.3), axis1=(0.0, 0.0, 1.0), axis2=(0.0, 1.0, 0.0),
                susp_erp=susp_erp, susp_cfm=susp_cfm, lo=0.0, hi=0.0)
    for _ in range(n_ramps):
        ang = rng.uniform(-0.3, 0.3)
        yaw = rng.uniform(-math.pi, math.pi)
        q = _quat_mul((math.cos(yaw / 2), 0.0, 0.0, math.sin(yaw / 2)), (math.cos(ang / 2), math.sin(ang / 2), 0.0, 0.0))
        b.geom(BOX, (rng.uniform(2.0, 6.0), rng.uniform(2.0, 6.0), rng.uniform(0.2, 0.8), 0.0), -1,
               pos=(rng.uniform(-40, 40), rng.uniform(-40, 40), 0.0), q=q)
    sc = b.scene(dt=h, mu=1.0, seed=seed)
    ctrl = np.zeros((nsteps, sc.nj, 4))
    steer = rng.uniform(-0.5, 0.5, size=nsteps)
    torque = rng.uniform(-500.0, 500.0, size=nsteps)
    for k in range(4):
        front = k < 2
        ctrl[:, k, 0] = steer if front else 0.0
        ctrl[:, k, 1] = steer if front else 0.0
        ctrl[:, k, 3] = torque
    sc.step_ctrl = ctrl
    if static_space:
        # the plane and the ramps in a static terrain space, collided against
        # the car's space with dSpaceCollide2 (the raycar frame)
        sc.geom_space = np.where(sc.geom_body < 0, 1, 0).astype(np.int32)
    return sc


def vehicle_worlds(n_worlds: int = 8192, nsteps: int = 1000, seed: int = 11) -> list[Scene]:
    rng = np.random.default_rng(seed)
    return [vehicle_world(rng, nsteps, seed=int(rng.integers(0, 2 ** 31))) for _ in range(n_worlds)]


def diff_piles(n_worlds: int = 1024, seed: int = 5, lo: int = 32, hi: int = 128) -> list[Scene]:
    """Worlds of U{32..128} boxes/spheres (the uneven-worlds pile distribution,
    no chains, spawned without interpenetration), h=0.01, with restitution
    enabled (dContactBounce)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_worlds):
        nb = int(rng.integers(lo, hi + 1))
        sc = pile_world(rng, nb, 0, seed=int(rng.integers(0, 2 ** 31)), spawn_clear=True)
        sc.surf_mode = APPROX1 | BOUNCE
        sc.sor_w = 1.0
        sc.bounce = 0.2
        sc.bounce_vel = 0.1
        out.append(sc)
    return out


def _quat_mul(a, b):
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return (aw * bw - ax * bx - ay * by - az * bz,
            aw * bx + ax * bw + ay * bz - az * by,
            aw * by + ay * bw + az * bx - ax * bz,
            aw * bz + az * bw + ax * by - ay * bx)
