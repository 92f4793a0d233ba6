// orc_world.hpp -- CPU restatement of the ODE 0.12 step path for one world:
// hash-space broadphase with the reference's callback order, contact joints,
// ball/hinge/hinge2 rows, islands, SOR-LCP with the seeded row shuffle, and
// the integrator.
//
// TEST INFRASTRUCTURE (oracle).  Only tests/, the smoke check and the CPU
// baseline leg of bench.py may load code built from this directory.  Every
// block cites the reference code whose behaviour it restates; list orders
// (prepend on create/attach, move-to-front of moved geoms) are modelled
// explicitly because pair, island and row order follow from them.
#pragma once
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdint>
#include <vector>

#include "orc_collide.hpp"
#include "orc_math.hpp"

namespace orc {

// ---------------------------------------------------------------- RNG
// Linear congruential generator and the xor-folded bounded draw:
// ode/src/misc.cpp:33-37 (dRand) and 66-117 (dRandInt).
struct Rng {
  unsigned long seed = 0;
  unsigned long next() {
    seed = (1664525UL * seed + 1013904223UL) & 0xffffffffUL;
    return seed;
  }
  int below(int n) {
    unsigned long un = (unsigned long)n;
    unsigned long r = next();
    if (un <= 0x10UL) {
      r ^= r >> 16; r ^= r >> 8; r ^= r >> 4;
      if (un <= 0x2UL) { r ^= r >> 2; r ^= r >> 1; }
      else if (un <= 0x4UL) { r ^= r >> 2; }
    } else if (un <= 0x100UL) {
      r ^= r >> 16; r ^= r >> 8;
    } else if (un <= 0x10000UL) {
      r ^= r >> 16;
    }
    return (int)(r % un);
  }
};

// ---------------------------------------------------------------- objects
enum { BODY_GYRO = 256 };
enum { J_CONTACT = 0, J_BALL = 1, J_HINGE = 2, J_HINGE2 = 3 };
enum { JF_REVERSE = 2 };
// contact surface mode bits: include/ode/contact.h:33-49
enum {
  CM_MU2 = 0x001, CM_FDIR1 = 0x002, CM_BOUNCE = 0x004, CM_SOFT_ERP = 0x008, CM_SOFT_CFM = 0x010,
  CM_MOTION1 = 0x020, CM_MOTION2 = 0x040, CM_MOTIONN = 0x080, CM_SLIP1 = 0x100, CM_SLIP2 = 0x200,
  CM_APPROX1_1 = 0x1000, CM_APPROX1_2 = 0x2000
};

template <class R> struct Body {
  R pos[4], q[4], Rm[12], lvel[4], avel[4], facc[4], tacc[4];
  R mass, invMass;
  R I[12], invI[12];
  int flags = BODY_GYRO;
  std::vector<int> geoms;   // geoms attached, newest first
  std::vector<int> jlist;   // attached joints, newest first (dJointAttach prepends)
  int tag = 0;
};

template <class R> struct Geom {
  int type, body;
  R param[4];
  R pos[4], Rm[12];         // pose of a static geom
  R aabb[6];
  unsigned long cat = ~0UL, col = ~0UL;
};

// limit/motor block: ode/src/joints/joint.h:172-190
template <class R> struct LimMot {
  R vel, fmax, lostop, histop, fudge, normal_cfm, stop_erp, stop_cfm, bounce;
  int limit = 0;
  R limit_err = 0;
  void init(R erp, R cfm) {  // joint.cpp:456-469
    vel = 0; fmax = 0; lostop = -INFINITY; histop = INFINITY; fudge = 1;
    normal_cfm = cfm; stop_erp = erp; stop_cfm = cfm; bounce = 0; limit = 0; limit_err = 0;
  }
  // joint.cpp:538-553
  int test_limit(R angle) {
    if (angle <= lostop) { limit = 1; limit_err = angle - lostop; return 1; }
    if (angle >= histop) { limit = 2; limit_err = angle - histop; return 1; }
    limit = 0;
    return 0;
  }
};

template <class R> struct Surface {
  int mode;
  R mu, mu2, bounce, bounce_vel, soft_erp, soft_cfm, motion1, motion2, motionN, slip1, slip2;
};

template <class R> struct Joint {
  int type;
  int node0 = -1, node1 = -1;   // first and second body (second may be -1)
  int flags = 0;
  int the_m = 0;
  long fb = -1;                 // feedback slot or -1
  // contact
  Surface<R> surf;
  R cpos[3], cnormal[3], cdepth, fdir1[3];
  // ball / hinge / hinge2
  R anchor1[4] = {0, 0, 0, 0}, anchor2[4] = {0, 0, 0, 0};
  R axis1[4] = {1, 0, 0, 0}, axis2[4] = {1, 0, 0, 0};
  R qrel[4] = {0, 0, 0, 0};
  R erp, cfm;                   // ball
  LimMot<R> limot, limot2;
  R susp_erp, susp_cfm;
  R v1[4] = {1, 0, 0, 0}, v2[4] = {0, 1, 0, 0};
  R c0 = 0, s0 = 0;
  R lam[6] = {0, 0, 0, 0, 0, 0};   // impulses of the rows after the last step
  int nrows = 0;
  R used_erp = 0;               // the Info2 erp of the joint's rows in the last step
  long long wkey = -1;          // contact: (g1, g2, index within the pair), the warm-start key
};

template <class R> struct WorldParams {
  R gravity[3] = {0, 0, 0};
  R erp = R(0.2), cfm = R(1e-5);
  int iterations = 20;
  R w = R(1.3);
  R max_vel = INFINITY, min_depth = 0;
  int hash_min = -3, hash_max = 10;
  int colored = 0;   // 1: graph-coloured Gauss-Seidel order (no shuffle), as the MI355X coloured mode
  int colour_width = 256;   // coloured mode: most members of a colour (the MI355X workgroup width)
  int exact = 0;     // 1: dWorldStep -- the island's LCP solved exactly (step.cpp, lcp.cpp)
  // coloured mode only: iterate until the LCP residual <= tol (checked every
  // check_every sweeps, at most max_iterations); accel = 1 adds the
  // nonsmooth nonlinear conjugate-gradient step between sweeps
  R tol = 0;
  int max_iterations = 0, check_every = 10, accel = 0;
  // coloured mode only: warm start; the solve starts from warm times the
  // impulses of the previous step (contacts matched by geom pair and index
  // within the pair, permanent joints by joint); 0 = start from zero
  R warm = 0;
  long long sweeps = 0;   // sweeps run (statistics)
  long long unconverged = 0;   // tolerance-mode solves that ended at the cap
};

// row buffers of one island (quickstep.cpp:711-743)
template <class R> struct Rows {
  std::vector<R> J, c, cfm, lo, hi, rhs, lambda;
  std::vector<int> findex, jb;
};

template <class R> struct World {
  WorldParams<R> prm;
  std::vector<Body<R>> bodies;
  std::vector<Geom<R>> geoms;
  std::vector<Joint<R>> joints;  // permanent joints, then this step's contacts
  int n_perm = 0;
  std::vector<int> space;        // geom ids in space list order (front first)
  Rng rng;
  // feedback output: 12 reals per contact joint with fb >= 0
  std::vector<R> fbout;

  // ------------------------------------------------------------ setup
  int add_body() {  // ode.cpp:248-294 (world body list prepends)
    Body<R> b;
    zero(b.pos, 4); zero(b.q, 4); b.q[0] = 1; identity3(b.Rm);
    zero(b.lvel, 4); zero(b.avel, 4); zero(b.facc, 4); zero(b.tacc, 4);
    b.mass = 1; b.invMass = 1; identity3(b.I); identity3(b.invI);
    bodies.push_back(b);
    return (int)bodies.size() - 1;
  }
  void set_mass(int bi, R mass, const R *I9) {  // ode.cpp:492-509
    Body<R> &b = bodies[bi];
    zero(b.I, 12);
    for (int i = 0; i < 3; i++) for (int j = 0; j < 3; j++) b.I[4 * i + j] = I9[3 * i + j];
    b.mass = mass;
    if (!invert_pd3(b.I, b.invI)) identity3(b.invI);
    b.invMass = recip(mass);
  }
  void set_quat(int bi, const R *q) {  // ode.cpp:380-396
    Body<R> &b = bodies[bi];
    for (int k = 0; k < 4; k++) b.q[k] = q[k];
    normalize4(b.q);
    quat_to_rot(b.Rm, b.q);
  }
  // geoms are prepended to the space list (collision_kernel.h:166-171)
  int add_geom(int type, const R *param) {
    Geom<R> g;
    g.type = type;
    g.body = -1;
    for (int k = 0; k < 4; k++) g.param[k] = param[k];
    if (type == G_PLANE) {  // plane.cpp:49-76 unit normal
      R l = g.param[0] * g.param[0] + g.param[1] * g.param[1] + g.param[2] * g.param[2];
      if (l > 0) {
        l = rsqrt(l);
        for (int k = 0; k < 4; k++) g.param[k] *= l;
      } else { g.param[0] = 1; g.param[1] = 0; g.param[2] = 0; g.param[3] = 0; }
    }
    zero(g.pos, 4);
    identity3(g.Rm);
    geoms.push_back(g);
    int id = (int)geoms.size() - 1;
    space.insert(space.begin(), id);
    return id;
  }
  void geom_set_body(int gi, int bi) {
    geoms[gi].body = bi;
    bodies[bi].geoms.insert(bodies[bi].geoms.begin(), gi);
  }
  const R *geom_pos(int gi) const {
    const Geom<R> &g = geoms[gi];
    return g.body >= 0 ? bodies[g.body].pos : g.pos;
  }
  const R *geom_rot(int gi) const {
    const Geom<R> &g = geoms[gi];
    return g.body >= 0 ? bodies[g.body].Rm : g.Rm;
  }
  GeomShape<R> shape(int gi) const {
    GeomShape<R> s;
    s.type = geoms[gi].type;
    s.pos = geom_pos(gi);
    s.Rm = geom_rot(gi);
    for (int k = 0; k < 4; k++) s.param[k] = geoms[gi].param[k];
    return s;
  }

  // dJointAttach: ode.cpp:1348-1399.  A missing first body swaps the two
  // and marks the joint reversed; node lists prepend.
  void attach(int ji, int b1, int b2) {
    Joint<R> &j = joints[ji];
    if (b1 < 0) { b1 = b2; b2 = -1; j.flags |= JF_REVERSE; }
    else j.flags &= ~JF_REVERSE;
    j.node0 = b1;
    j.node1 = b2;
    if (b1 >= 0) bodies[b1].jlist.insert(bodies[b1].jlist.begin(), ji);
    if (b2 >= 0) bodies[b2].jlist.insert(bodies[b2].jlist.begin(), ji);
  }
  int other_body(int ji, int b) const {
    const Joint<R> &j = joints[ji];
    return (j.node0 == b) ? j.node1 : j.node0;
  }

  // anchors relative to the bodies: joint.cpp:260-288
  void set_anchors(Joint<R> &j, R x, R y, R z) {
    if (j.node0 >= 0) {
      const Body<R> &b1 = bodies[j.node0];
      R q[4] = {x - b1.pos[0], y - b1.pos[1], z - b1.pos[2], 0};
      mtv(j.anchor1, b1.Rm, q);
      if (j.node1 >= 0) {
        const Body<R> &b2 = bodies[j.node1];
        R q2[4] = {x - b2.pos[0], y - b2.pos[1], z - b2.pos[2], 0};
        mtv(j.anchor2, b2.Rm, q2);
      } else { j.anchor2[0] = x; j.anchor2[1] = y; j.anchor2[2] = z; }
    }
    j.anchor1[3] = 0;
    j.anchor2[3] = 0;
  }
  // axes relative to the bodies: joint.cpp:293-324
  void set_axes(Joint<R> &j, R x, R y, R z, bool a1, bool a2) {
    if (j.node0 < 0) return;
    R q[4] = {x, y, z, 0};
    normalize3(q);
    if (a1) { mtv(j.axis1, bodies[j.node0].Rm, q); j.axis1[3] = 0; }
    if (a2) {
      if (j.node1 >= 0) mtv(j.axis2, bodies[j.node1].Rm, q);
      else { j.axis2[0] = x; j.axis2[1] = y; j.axis2[2] = z; }
      j.axis2[3] = 0;
    }
  }
  // hinge qrel: hinge.cpp:377-395
  void hinge_qrel(Joint<R> &j) {
    if (j.node0 < 0) return;
    const R *q1 = bodies[j.node0].q;
    if (j.node1 >= 0) qmul1(j.qrel, q1, bodies[j.node1].q);
    else { j.qrel[0] = q1[0]; j.qrel[1] = -q1[1]; j.qrel[2] = -q1[2]; j.qrel[3] = -q1[3]; }
  }
  // hinge2 axis data: hinge2.cpp:107-115
  void hinge2_axis_info(const Joint<R> &j, R *ax1, R *ax2, R *axc, R &sn, R &cs) const {
    mv(ax1, bodies[j.node0].Rm, j.axis1);
    mv(ax2, bodies[j.node1].Rm, j.axis2);
    cross3(axc, ax1, ax2);
    sn = std::sqrt(axc[0] * axc[0] + axc[1] * axc[1] + axc[2] * axc[2]);
    cs = dot3(ax1, ax2);
  }
  // hinge2.cpp:186-211
  void hinge2_make_v1v2(Joint<R> &j) {
    if (j.node0 < 0) return;
    R ax1[4], ax2[4], v[4];
    mv(ax1, bodies[j.node0].Rm, j.axis1);
    mv(ax2, bodies[j.node1].Rm, j.axis2);
    if ((ax1[0] == 0 && ax1[1] == 0 && ax1[2] == 0) || (ax2[0] == 0 && ax2[1] == 0 && ax2[2] == 0) ||
        (ax1[0] == ax2[0] && ax1[1] == ax2[1] && ax1[2] == ax2[2]))
      return;
    R k = dot3(ax1, ax2);
    for (int i = 0; i < 3; i++) ax2[i] -= k * ax1[i];
    ax2[3] = 0;
    normalize3(ax2);
    cross3(v, ax1, ax2);
    mtv(j.v1, bodies[j.node0].Rm, ax2);
    mtv(j.v2, bodies[j.node0].Rm, v);
  }
  // hinge2.cpp:34-43
  R hinge2_angle(const Joint<R> &j) const {
    R a1[3], a2[3];
    mv(a1, bodies[j.node1].Rm, j.axis2);
    mtv(a2, bodies[j.node0].Rm, a1);
    R x = dot3(j.v1, a2), y = dot3(j.v2, a2);
    return -std::atan2(y, x);
  }
  // hinge angle from the relative rotation: joint.cpp:381-450
  R hinge_angle(const Joint<R> &j) const {
    R qr[4];
    const R *q1 = bodies[j.node0].q;
    if (j.node1 >= 0) {
      R qq[4];
      qmul1(qq, q1, bodies[j.node1].q);
      qmul2(qr, qq, j.qrel);
    } else {
      qmul3(qr, q1, j.qrel);
    }
    R cost2 = qr[0];
    R sint2 = std::sqrt(qr[1] * qr[1] + qr[2] * qr[2] + qr[3] * qr[3]);
    R th = (dot3(qr + 1, j.axis1) >= 0) ? 2 * std::atan2(sint2, cost2) : 2 * std::atan2(sint2, -cost2);
    if (th > M_PI) th -= (R)(2 * M_PI);
    return -th;
  }
  void add_torque(int b, R x, R y, R z) { bodies[b].tacc[0] += x; bodies[b].tacc[1] += y; bodies[b].tacc[2] += z; }
  void add_force(int b, R x, R y, R z) { bodies[b].facc[0] += x; bodies[b].facc[1] += y; bodies[b].facc[2] += z; }
  // dJointAddHinge2Torques: hinge2.cpp:393-410
  void hinge2_add_torques(int ji, R t1, R t2) {
    const Joint<R> &j = joints[ji];
    if (j.node0 < 0 || j.node1 < 0) return;
    R a1[3], a2[3];
    mv(a1, bodies[j.node0].Rm, j.axis1);
    mv(a2, bodies[j.node1].Rm, j.axis2);
    R t[3];
    for (int i = 0; i < 3; i++) t[i] = a1[i] * t1 + a2[i] * t2;
    add_torque(j.node0, t[0], t[1], t[2]);
    add_torque(j.node1, -t[0], -t[1], -t[2]);
  }

  // ------------------------------------------------------------ broadphase
  // Hash space collide: ode/src/collision_space.cpp:420-583 with
  // collideAABBs (collision_space_internal.h:48-82).  Emits callback pairs
  // (g1,g2) in the reference's order.  All geoms are enabled here.
  static int find_level(const R *b) {  // collision_space.cpp:331-351
    for (int k = 0; k < 6; k++)
      if (std::isinf(b[k])) return INT_MAX;
    R q = b[1] - b[0];
    if (b[3] - b[2] > q) q = b[3] - b[2];
    if (b[5] - b[4] > q) q = b[5] - b[4];
    int lvl;
    std::frexp(q, &lvl);
    return lvl;
  }
  bool aabb_pass(int a, int b) const {
    const Geom<R> &g1 = geoms[a], &g2 = geoms[b];
    if (g1.body == g2.body && g1.body >= 0) return false;
    if (((g1.cat & g2.col) || (g2.cat & g1.col)) == 0) return false;
    const R *b1 = g1.aabb, *b2 = g2.aabb;
    if (b1[0] > b2[1] || b1[1] < b2[0] || b1[2] > b2[3] || b1[3] < b2[2] || b1[4] > b2[5] || b1[5] < b2[4])
      return false;
    return true;
  }
  // A second top-level space (geom ids in list order), collided against the
  // first with dSpaceCollide2 (collision_space.cpp:768-837): the space with
  // fewer geoms is iterated outside, each of its geoms tested against every
  // geom of the other in list order (dxHashSpace::collide2,
  // collision_space.cpp:585-603); the pair keeps the (first space, second
  // space) argument order.  Appends to out.
  std::vector<int> space2;
  void collide2_pairs(std::vector<std::pair<int, int>> &out) {
    for (int gi : space) geom_aabb(shape(gi), geoms[gi].aabb);
    for (int gi : space2) geom_aabb(shape(gi), geoms[gi].aabb);
    if (space.size() < space2.size()) {
      for (int a : space)
        for (int b : space2)
          if (aabb_pass(b, a)) out.push_back({a, b});
    } else {
      for (int b : space2)
        for (int a : space)
          if (aabb_pass(a, b)) out.push_back({a, b});
    }
  }
  // moves a geom from the first space's list to the front of the second's
  // (a geom created in the second space: list prepend, collision_kernel.h:166-171)
  void move_to_space2(int gi) {
    auto it = std::find(space.begin(), space.end(), gi);
    if (it != space.end()) space.erase(it);
    space2.insert(space2.begin(), gi);
  }
  void hash_pairs(std::vector<std::pair<int, int>> &out) {
    out.clear();
    for (int gi : space) geom_aabb(shape(gi), geoms[gi].aabb);
    if (space.size() < 2) return;
    struct Box { int geom, level, db[6], index; };
    std::vector<Box> main_list;  // in first_aabb order (reverse of the space list)
    std::vector<int> big;        // big_boxes order (reverse of the space list)
    int maxlevel = prm.hash_min - 1;
    int n = 0;
    std::vector<Box> tmp;
    for (int gi : space) {
      int lvl = find_level(geoms[gi].aabb);
      if (lvl < prm.hash_min) lvl = prm.hash_min;
      if (lvl <= prm.hash_max) {
        Box bx;
        bx.geom = gi;
        bx.level = lvl;
        if (lvl > maxlevel) maxlevel = lvl;
        R cell = (R)std::ldexp(1.0, lvl);
        for (int k = 0; k < 6; k++) bx.db[k] = (int)std::floor(geoms[gi].aabb[k] / cell);
        bx.index = n++;
        tmp.push_back(bx);
      } else {
        big.insert(big.begin(), gi);
      }
    }
    main_list.assign(tmp.rbegin(), tmp.rend());
    // tested bitmap and a hash table of cell nodes (collision_space.cpp:478-516)
    int rowsize = (n + 7) >> 3;
    std::vector<unsigned char> tested((size_t)n * rowsize, 0);
    static const long primes[31] = {1L, 2L, 3L, 7L, 13L, 31L, 61L, 127L, 251L, 509L, 1021L, 2039L, 4093L,
                                     8191L, 16381L, 32749L, 65521L, 131071L, 262139L, 524287L, 1048573L,
                                     2097143L, 4194301L, 8388593L, 16777213L, 33554393L, 67108859L,
                                     134217689L, 268435399L, 536870909L, 1073741789L};
    int pi = 0;
    for (; pi < 31; pi++) if (primes[pi] >= 8 * n) break;
    if (pi >= 31) pi = 30;
    long sz = primes[pi];
    struct Node { int next, x, y, z, box; };
    std::vector<Node> nodes;
    std::vector<int> table(sz, -1);
    auto vaddr = [](int level, int x, int y, int z) -> unsigned long {
      return (unsigned long)(level * 1000 + x * 100 + y * 10 + z);
    };
    for (int bi = 0; bi < (int)main_list.size(); bi++) {
      const Box &bx = main_list[bi];
      for (int x = bx.db[0]; x <= bx.db[1]; x++)
        for (int y = bx.db[2]; y <= bx.db[3]; y++)
          for (int z = bx.db[4]; z <= bx.db[5]; z++) {
            unsigned long h = vaddr(bx.level, x, y, z) % (unsigned long)sz;
            nodes.push_back({table[h], x, y, z, bi});
            table[h] = (int)nodes.size() - 1;
          }
    }
    for (int bi = 0; bi < (int)main_list.size(); bi++) {
      const Box &bx = main_list[bi];
      int db[6];
      for (int k = 0; k < 6; k++) db[k] = bx.db[k];
      for (int level = bx.level; level <= maxlevel; level++) {
        for (int x = db[0]; x <= db[1]; x++)
          for (int y = db[2]; y <= db[3]; y++)
            for (int z = db[4]; z <= db[5]; z++) {
              unsigned long h = vaddr(level, x, y, z) % (unsigned long)sz;
              for (int ni = table[h]; ni >= 0; ni = nodes[ni].next) {
                const Node &nd = nodes[ni];
                if (nd.box == bi) continue;
                const Box &ob = main_list[nd.box];
                if (ob.level != level || nd.x != x || nd.y != y || nd.z != z) continue;
                size_t idx;
                unsigned char mask;
                if (bx.index <= ob.index) { idx = (size_t)bx.index * rowsize + (ob.index >> 3); mask = (unsigned char)(1 << (ob.index & 7)); }
                else { idx = (size_t)ob.index * rowsize + (bx.index >> 3); mask = (unsigned char)(1 << (bx.index & 7)); }
                if ((tested[idx] & mask) == 0 && aabb_pass(bx.geom, ob.geom)) out.push_back({bx.geom, ob.geom});
                tested[idx] |= mask;
              }
            }
        for (int k = 0; k < 6; k++) db[k] >>= 1;
      }
    }
    for (const Box &bx : main_list)
      for (int g2 : big)
        if (aabb_pass(bx.geom, g2)) out.push_back({bx.geom, g2});
    for (size_t a = 0; a < big.size(); a++)
      for (size_t b = a + 1; b < big.size(); b++)
        if (aabb_pass(big[a], big[b])) out.push_back({big[a], big[b]});
  }

  // dAreConnectedExcluding(b1, b2, contact): ode.cpp:1529-1536
  bool connected_non_contact(int b1, int b2) const {
    for (int ji : bodies[b1].jlist)
      if (joints[ji].type != J_CONTACT && other_body(ji, b1) == b2) return true;
    return false;
  }

  int add_contact_joint(const Surface<R> &s, const ContactGeom<R> &cg, int b1, int b2) {
    Joint<R> j;
    j.type = J_CONTACT;
    j.surf = s;
    for (int k = 0; k < 3; k++) { j.cpos[k] = cg.pos[k]; j.cnormal[k] = cg.normal[k]; j.fdir1[k] = 0; }
    j.cdepth = cg.depth;
    int slot = 0;
    const long long pk = ((long long)cg.g1 << 32) | ((long long)cg.g2 << 8);
    if ((int)joints.size() > n_perm && (joints.back().wkey & ~255LL) == pk) slot = (int)(joints.back().wkey & 255) + 1;
    j.wkey = pk | (slot & 255);
    joints.push_back(j);
    int ji = (int)joints.size() - 1;
    attach(ji, b1, b2);
    return ji;
  }
  // dJointGroupEmpty: ode.cpp:1305-1332 (contacts are the newest entries of
  // every body list, so removing them keeps the rest in order)
  void empty_contacts() {
    for (Body<R> &b : bodies) {
      size_t k = 0;
      while (k < b.jlist.size() && b.jlist[k] >= n_perm) k++;
      b.jlist.erase(b.jlist.begin(), b.jlist.begin() + k);
    }
    joints.resize(n_perm);
  }

  // ------------------------------------------------------------ rows
  struct Info2 {
    R fps, erp;
    R *J;       // rows of 12 (l1 a1 l2 a2)
    R *c, *cfm, *lo, *hi;
    int *findex;
  };
  void row_set_ball(const Joint<R> &j, Info2 &in, const R *anc1, const R *anc2) {  // joint.cpp:85-126
    R *J = in.J;
    J[0] = 1; J[12 + 1] = 1; J[24 + 2] = 1;
    R a1[3], a2[3];
    mv(a1, bodies[j.node0].Rm, anc1);
    // -[a1]x into the angular block of body 1
    J[3 + 1] = a1[2]; J[3 + 2] = -a1[1];
    J[12 + 3 + 0] = -a1[2]; J[12 + 3 + 2] = a1[0];
    J[24 + 3 + 0] = a1[1]; J[24 + 3 + 1] = -a1[0];
    if (j.node1 >= 0) {
      J[6] = -1; J[12 + 7] = -1; J[24 + 8] = -1;
      mv(a2, bodies[j.node1].Rm, anc2);
      J[9 + 1] = -a2[2]; J[9 + 2] = a2[1];
      J[12 + 9 + 0] = a2[2]; J[12 + 9 + 2] = -a2[0];
      J[24 + 9 + 0] = -a2[1]; J[24 + 9 + 1] = a2[0];
    }
    R k = in.fps * in.erp;
    const R *p1 = bodies[j.node0].pos;
    if (j.node1 >= 0) {
      const R *p2 = bodies[j.node1].pos;
      for (int i = 0; i < 3; i++) in.c[i] = k * (a2[i] + p2[i] - a1[i] - p1[i]);
    } else {
      for (int i = 0; i < 3; i++) in.c[i] = k * (anc2[i] - a1[i] - p1[i]);
    }
  }
  void row_set_ball2(const Joint<R> &j, Info2 &in, const R *anc1, const R *anc2, const R *axis, R erp1) {  // joint.cpp:134-196
    R *J = in.J;
    R q1[3], q2[3];
    plane_space(axis, q1, q2);
    for (int i = 0; i < 3; i++) { J[i] = axis[i]; J[12 + i] = q1[i]; J[24 + i] = q2[i]; }
    R a1[3], a2[3];
    mv(a1, bodies[j.node0].Rm, anc1);
    cross3(J + 3, a1, axis);
    cross3(J + 12 + 3, a1, q1);
    cross3(J + 24 + 3, a1, q2);
    if (j.node1 >= 0) {
      for (int i = 0; i < 3; i++) { J[6 + i] = -axis[i]; J[12 + 6 + i] = -q1[i]; J[24 + 6 + i] = -q2[i]; }
      mv(a2, bodies[j.node1].Rm, anc2);
      const R *qs[3] = {axis, q1, q2};
      for (int r = 0; r < 3; r++) {
        R t[3];
        cross3(t, a2, qs[r]);
        for (int i = 0; i < 3; i++) J[12 * r + 9 + i] = -t[i];
      }
    }
    R k1 = in.fps * erp1;
    R k = in.fps * in.erp;
    for (int i = 0; i < 3; i++) a1[i] += bodies[j.node0].pos[i];
    R d[3];
    if (j.node1 >= 0) {
      for (int i = 0; i < 3; i++) a2[i] += bodies[j.node1].pos[i];
      for (int i = 0; i < 3; i++) d[i] = a2[i] - a1[i];
    } else {
      for (int i = 0; i < 3; i++) d[i] = anc2[i] - a1[i];
    }
    in.c[0] = k1 * dot3(axis, d);
    in.c[1] = k * dot3(q1, d);
    in.c[2] = k * dot3(q2, d);
  }
  // dxJointLimitMotor::addLimot for rotational axes: joint.cpp:556-733
  int row_limot(const Joint<R> &j, LimMot<R> &lm, Info2 &in, int row, const R *ax1) {
    int powered = lm.fmax > 0;
    if (!(powered || lm.limit)) return 0;
    R *J = in.J + 12 * row;
    for (int i = 0; i < 3; i++) J[3 + i] = ax1[i];
    if (j.node1 >= 0) for (int i = 0; i < 3; i++) J[9 + i] = -ax1[i];
    if (lm.limit && lm.lostop == lm.histop) powered = 0;
    if (powered) {
      in.cfm[row] = lm.normal_cfm;
      if (!lm.limit) {
        in.c[row] = lm.vel;
        in.lo[row] = -lm.fmax;
        in.hi[row] = lm.fmax;
      } else {
        R fm = lm.fmax;
        if (lm.vel > 0 || (lm.vel == 0 && lm.limit == 2)) fm = -fm;
        if ((lm.limit == 1 && lm.vel > 0) || (lm.limit == 2 && lm.vel < 0)) fm *= lm.fudge;
        add_torque(j.node0, -fm * ax1[0], -fm * ax1[1], -fm * ax1[2]);
        if (j.node1 >= 0) add_torque(j.node1, fm * ax1[0], fm * ax1[1], fm * ax1[2]);
      }
    }
    if (lm.limit) {
      R k = in.fps * lm.stop_erp;
      in.c[row] = -k * lm.limit_err;
      in.cfm[row] = lm.stop_cfm;
      if (lm.lostop == lm.histop) {
        in.lo[row] = -INFINITY;
        in.hi[row] = INFINITY;
      } else {
        if (lm.limit == 1) { in.lo[row] = 0; in.hi[row] = INFINITY; }
        else { in.lo[row] = -INFINITY; in.hi[row] = 0; }
        if (lm.bounce > 0) {
          R v = dot3(bodies[j.node0].avel, ax1);
          if (j.node1 >= 0) v -= dot3(bodies[j.node1].avel, ax1);
          if (lm.limit == 1) {
            if (v < 0) { R nc = -lm.bounce * v; if (nc > in.c[row]) in.c[row] = nc; }
          } else {
            if (v > 0) { R nc = -lm.bounce * v; if (nc < in.c[row]) in.c[row] = nc; }
          }
        }
      }
    }
    return 1;
  }
  // getInfo1: contact.cpp:47-71, ball.cpp:48-52, hinge.cpp:53-74, hinge2.cpp:80-100
  int info1(Joint<R> &j) {
    switch (j.type) {
      case J_CONTACT: {
        int m = 1;
        if (j.surf.mu < 0) j.surf.mu = 0;
        if (j.surf.mode & CM_MU2) {
          if (j.surf.mu > 0 || j.surf.mu2 > 0) m++;
          if (j.surf.mu2 < 0) j.surf.mu2 = 0;
          if (j.surf.mu2 > 0) m++;
        } else if (j.surf.mu > 0) m += 2;
        j.the_m = m;
        return m;
      }
      case J_BALL: return 3;
      case J_HINGE: {
        int m = (j.limot.fmax > 0) ? 6 : 5;
        if ((j.limot.lostop >= -M_PI || j.limot.histop <= M_PI) && j.limot.lostop <= j.limot.histop) {
          if (j.limot.test_limit(hinge_angle(j))) m = 6;
        }
        return m;
      }
      case J_HINGE2: {
        int m = 4;
        j.limot.limit = 0;
        if ((j.limot.lostop >= -M_PI || j.limot.histop <= M_PI) && j.limot.lostop <= j.limot.histop)
          j.limot.test_limit(hinge2_angle(j));
        if (j.limot.limit || j.limot.fmax > 0) m++;
        j.limot2.limit = 0;
        if (j.limot2.fmax > 0) m++;
        return m;
      }
    }
    return 0;
  }
  // getInfo2: contact.cpp:75-256, ball.cpp:56-64, hinge.cpp:76-149,
  // hinge2.cpp:127-180
  void info2(Joint<R> &j, Info2 &in) {
    R *J = in.J;
    if (j.type == J_CONTACT) {
      R n[3];
      for (int k = 0; k < 3; k++) n[k] = (j.flags & JF_REVERSE) ? -j.cnormal[k] : j.cnormal[k];
      const Body<R> &b1 = bodies[j.node0];
      R c1[3], c2[3] = {0, 0, 0};
      for (int k = 0; k < 3; k++) c1[k] = j.cpos[k] - b1.pos[k];
      for (int k = 0; k < 3; k++) J[k] = n[k];
      cross3(J + 3, c1, n);
      if (j.node1 >= 0) {
        const Body<R> &b2 = bodies[j.node1];
        for (int k = 0; k < 3; k++) c2[k] = j.cpos[k] - b2.pos[k];
        for (int k = 0; k < 3; k++) J[6 + k] = -n[k];
        cross3(J + 9, c2, n);
        for (int k = 0; k < 3; k++) J[9 + k] = -J[9 + k];
      }
      R erp = in.erp;
      if (j.surf.mode & CM_SOFT_ERP) erp = j.surf.soft_erp;
      R k = in.fps * erp;
      R depth = j.cdepth - prm.min_depth;
      if (depth < 0) depth = 0;
      if (j.surf.mode & CM_SOFT_CFM) in.cfm[0] = j.surf.soft_cfm;
      R motionN = (j.surf.mode & CM_MOTIONN) ? j.surf.motionN : R(0);
      in.c[0] = k * depth + motionN;
      if (in.c[0] > prm.max_vel) in.c[0] = prm.max_vel;
      if (j.surf.mode & CM_BOUNCE) {
        R out = dot3(J, b1.lvel) + dot3(J + 3, b1.avel);
        if (j.node1 >= 0) out += dot3(J + 6, bodies[j.node1].lvel) + dot3(J + 9, bodies[j.node1].avel);
        out -= motionN;
        if (j.surf.bounce_vel >= 0 && (-out) > j.surf.bounce_vel) {
          R nc = -j.surf.bounce * out + motionN;
          if (nc > in.c[0]) in.c[0] = nc;
        }
      }
      in.lo[0] = 0;
      in.hi[0] = INFINITY;
      R t1[3], t2[3];
      if (j.the_m >= 2) {
        if (j.surf.mode & CM_FDIR1) {
          for (int k2 = 0; k2 < 3; k2++) t1[k2] = j.fdir1[k2];
          cross3(t2, n, t1);
        } else {
          plane_space(n, t1, t2);
        }
        R *Jr = J + 12;
        for (int k2 = 0; k2 < 3; k2++) Jr[k2] = t1[k2];
        cross3(Jr + 3, c1, t1);
        if (j.node1 >= 0) {
          for (int k2 = 0; k2 < 3; k2++) Jr[6 + k2] = -t1[k2];
          cross3(Jr + 9, c2, t1);
          for (int k2 = 0; k2 < 3; k2++) Jr[9 + k2] = -Jr[9 + k2];
        }
        if (j.surf.mode & CM_MOTION1) in.c[1] = j.surf.motion1;
        in.lo[1] = -j.surf.mu;
        in.hi[1] = j.surf.mu;
        if (j.surf.mode & CM_APPROX1_1) in.findex[1] = 0;
        if (j.surf.mode & CM_SLIP1) in.cfm[1] = j.surf.slip1;
      }
      if (j.the_m >= 3) {
        R *Jr = J + 24;
        for (int k2 = 0; k2 < 3; k2++) Jr[k2] = t2[k2];
        cross3(Jr + 3, c1, t2);
        if (j.node1 >= 0) {
          for (int k2 = 0; k2 < 3; k2++) Jr[6 + k2] = -t2[k2];
          cross3(Jr + 9, c2, t2);
          for (int k2 = 0; k2 < 3; k2++) Jr[9 + k2] = -Jr[9 + k2];
        }
        if (j.surf.mode & CM_MOTION2) in.c[2] = j.surf.motion2;
        if (j.surf.mode & CM_MU2) { in.lo[2] = -j.surf.mu2; in.hi[2] = j.surf.mu2; }
        else { in.lo[2] = -j.surf.mu; in.hi[2] = j.surf.mu; }
        if (j.surf.mode & CM_APPROX1_2) in.findex[2] = 0;
        if (j.surf.mode & CM_SLIP2) in.cfm[2] = j.surf.slip2;
      }
    } else if (j.type == J_BALL) {
      in.erp = j.erp;
      in.cfm[0] = in.cfm[1] = in.cfm[2] = j.cfm;
      row_set_ball(j, in, j.anchor1, j.anchor2);
    } else if (j.type == J_HINGE) {
      row_set_ball(j, in, j.anchor1, j.anchor2);
      R ax1[3], p[3], q[3];
      mv(ax1, bodies[j.node0].Rm, j.axis1);
      plane_space(ax1, p, q);
      for (int i = 0; i < 3; i++) { J[36 + 3 + i] = p[i]; J[48 + 3 + i] = q[i]; }
      if (j.node1 >= 0)
        for (int i = 0; i < 3; i++) { J[36 + 9 + i] = -p[i]; J[48 + 9 + i] = -q[i]; }
      R ax2[3], b[3];
      if (j.node1 >= 0) mv(ax2, bodies[j.node1].Rm, j.axis2);
      else for (int i = 0; i < 3; i++) ax2[i] = j.axis2[i];
      cross3(b, ax1, ax2);
      R k = in.fps * in.erp;
      in.c[3] = k * dot3(b, p);
      in.c[4] = k * dot3(b, q);
      row_limot(j, j.limot, in, 5, ax1);
    } else if (j.type == J_HINGE2) {
      R ax1[4], ax2[4], qv[4], sn, cs;
      hinge2_axis_info(j, ax1, ax2, qv, sn, cs);
      qv[3] = 0;
      normalize3(qv);
      row_set_ball2(j, in, j.anchor1, j.anchor2, ax1, j.susp_erp);
      for (int i = 0; i < 3; i++) J[36 + 3 + i] = qv[i];
      if (j.node1 >= 0) for (int i = 0; i < 3; i++) J[36 + 9 + i] = -qv[i];
      R k = in.fps * in.erp;
      in.c[3] = k * (j.c0 * sn - j.s0 * cs);
      int row = 4 + row_limot(j, j.limot, in, 4, ax1);
      row_limot(j, j.limot2, in, row, ax2);
      in.cfm[0] = j.susp_cfm;
    }
  }

  // ------------------------------------------------------------ stepper
  // dxQuickStepper for one island: ode/src/quickstep.cpp:592-1025 with
  // SOR_LCP (342-584) and compute_invM_JT (117-136).
  void quickstep_island(const std::vector<int> &ib, const std::vector<int> &ij, R h) {
    const R hinv = recip(h);
    const int nb = (int)ib.size();
    for (int i = 0; i < nb; i++) bodies[ib[i]].tag = i;
    std::vector<R> invI(12 * (size_t)nb);
    for (int i = 0; i < nb; i++) {
      Body<R> &b = bodies[ib[i]];
      R tmp[12];
      mmt(tmp, b.invI, b.Rm);
      mm(&invI[12 * i], b.Rm, tmp);
      if (b.flags & BODY_GYRO) {
        R Iw[12], t3[3];
        mmt(tmp, b.I, b.Rm);
        mm(Iw, b.Rm, tmp);
        mv(t3, Iw, b.avel);
        R cr[3];
        cross3(cr, b.avel, t3);
        for (int k = 0; k < 3; k++) b.tacc[k] -= cr[k];
      }
    }
    for (int k = 0; k < 3; k++) {
      if (prm.gravity[k] == 0) continue;
      for (int i = 0; i < nb; i++) {
        Body<R> &b = bodies[ib[i]];
        b.facc[k] += b.mass * prm.gravity[k];
      }
    }
    std::vector<int> act;
    std::vector<int> jm;
    for (int ji : ij) {
      int m = info1(joints[ji]);
      if (m > 0) { act.push_back(ji); jm.push_back(m); }
    }
    int m = 0;
    for (int x : jm) m += x;
    std::vector<R> cforce(6 * (size_t)nb, 0);
    std::vector<R> lambda(m, 0);
    std::vector<R> Jcopy;
    std::vector<long> fbslots;
    if (m > 0) {
      std::vector<R> J(12 * (size_t)m, 0), c(m, 0), cfm(m, prm.cfm), lo(m, -INFINITY), hi(m, INFINITY), rhs(m);
      std::vector<int> findex(m, -1), jb(2 * (size_t)m);
      Info2 in;
      in.fps = hinv;
      // set once per island: a ball joint overwrites it for the joints that
      // follow (quickstep.cpp:777-779, ball.cpp:58)
      in.erp = prm.erp;
      int ofs = 0;
      for (size_t a = 0; a < act.size(); a++) {
        Joint<R> &jt = joints[act[a]];
        in.J = &J[12 * (size_t)ofs];
        in.c = &c[ofs]; in.cfm = &cfm[ofs]; in.lo = &lo[ofs]; in.hi = &hi[ofs]; in.findex = &findex[ofs];
        info2(jt, in);
        jt.used_erp = in.erp;
        if (jt.fb >= 0) Jcopy.insert(Jcopy.end(), J.begin() + 12 * (size_t)ofs, J.begin() + 12 * (size_t)(ofs + jm[a]));
        for (int r = 0; r < jm[a]; r++)
          if (findex[ofs + r] != -1) findex[ofs + r] += ofs;
        int b1 = bodies[jt.node0].tag;
        int b2 = jt.node1 >= 0 ? bodies[jt.node1].tag : -1;
        for (int r = 0; r < jm[a]; r++) { jb[2 * (ofs + r)] = b1; jb[2 * (ofs + r) + 1] = b2; }
        ofs += jm[a];
      }
      // rhs = c/h - J (v/h + M^-1 f)   (quickstep.cpp:832-854)
      std::vector<R> tmp1(6 * (size_t)nb);
      for (int i = 0; i < nb; i++) {
        Body<R> &b = bodies[ib[i]];
        for (int k = 0; k < 3; k++) tmp1[6 * i + k] = b.facc[k] * b.invMass + b.lvel[k] * hinv;
        mv(&tmp1[6 * i + 3], &invI[12 * i], b.tacc);
        for (int k = 0; k < 3; k++) tmp1[6 * i + 3 + k] += b.avel[k] * hinv;
      }
      for (int i = 0; i < m; i++) {
        int b1 = jb[2 * i], b2 = jb[2 * i + 1];
        R s = 0;
        for (int k = 0; k < 6; k++) s += J[12 * i + k] * tmp1[6 * b1 + k];
        if (b2 != -1) for (int k = 0; k < 6; k++) s += J[12 * i + 6 + k] * tmp1[6 * b2 + k];
        rhs[i] = s;
      }
      for (int i = 0; i < m; i++) rhs[i] = c[i] * hinv - rhs[i];
      for (int i = 0; i < m; i++) cfm[i] *= hinv;
      if (prm.colored) {
        // constraint ids of the coloured mode: contacts in creation order,
        // then the permanent joints in creation order
        const int nk = (int)joints.size() - n_perm;
        std::vector<int> row0(act.size()), cid(act.size());
        int o = 0;
        for (size_t a = 0; a < act.size(); a++) {
          row0[a] = o;
          o += jm[a];
          cid[a] = act[a] >= n_perm ? act[a] - n_perm : nk + act[a];
        }
        if (prm.warm != 0 && warm_ok) {
          for (size_t a = 0; a < act.size(); a++) {
            const Joint<R> &jt = joints[act[a]];
            const R *prev = nullptr;
            int np = 0;
            if (act[a] >= n_perm) {
              auto f = std::lower_bound(warm_prev.begin(), warm_prev.end(), std::make_pair(jt.wkey, std::array<R, 3>{}),
                                        [](const auto &x, const auto &y) { return x.first < y.first; });
              if (f != warm_prev.end() && f->first == jt.wkey) { prev = f->second.data(); np = 3; }
            } else if (jt.nrows == jm[a]) {
              prev = jt.lam;
              np = 6;
            }
            for (int r = 0; r < jm[a] && r < np; r++) lambda[row0[a] + r] = prm.warm * prev[r];
          }
        }
        sor_colored(m, nb, J, jb, ib, invI, lambda, cforce, rhs, lo, hi, cfm, findex, row0, jm, cid);
      } else if (prm.exact) {
        lcp_exact(m, nb, J, jb, ib, invI, lambda, cforce, rhs, lo, hi, cfm, findex);
      } else {
        sor_lcp(m, nb, J, jb, ib, invI, lambda, cforce, rhs, lo, hi, cfm, findex);
      }
      for (int i = 0; i < nb; i++) {
        Body<R> &b = bodies[ib[i]];
        for (int k = 0; k < 3; k++) { b.lvel[k] += h * cforce[6 * i + k]; b.avel[k] += h * cforce[6 * i + 3 + k]; }
      }
      // joint feedback = J^T lambda per body (quickstep.cpp:918-957)
      size_t jc = 0;
      int lo_ofs = 0;
      for (size_t a = 0; a < act.size(); a++) {
        Joint<R> &jt = joints[act[a]];
        jt.nrows = jm[a];
        for (int r = 0; r < jm[a] && r < 6; r++) jt.lam[r] = lambda[lo_ofs + r];
        if (jt.fb >= 0) {
          R d1[6] = {0, 0, 0, 0, 0, 0}, d2[6] = {0, 0, 0, 0, 0, 0};
          for (int r = 0; r < jm[a]; r++) {
            R s = lambda[lo_ofs + r];
            for (int k = 0; k < 6; k++) d1[k] += Jcopy[jc + 12 * r + k] * s;
            for (int k = 0; k < 6; k++) d2[k] += Jcopy[jc + 12 * r + 6 + k] * s;
          }
          R *o = &fbout[12 * (size_t)jt.fb];
          for (int k = 0; k < 6; k++) o[k] = d1[k];
          if (jt.node1 >= 0) for (int k = 0; k < 6; k++) o[6 + k] = d2[k];
          jc += 12 * (size_t)jm[a];
        }
        lo_ofs += jm[a];
      }
    }
    // v += h M^-1 f_ext (quickstep.cpp:960-975)
    for (int i = 0; i < nb; i++) {
      Body<R> &b = bodies[ib[i]];
      R k = h * b.invMass;
      for (int j = 0; j < 3; j++) { b.lvel[j] += k * b.facc[j]; b.tacc[j] *= h; }
      R t[3];
      mv(t, &invI[12 * i], b.tacc);
      for (int j = 0; j < 3; j++) b.avel[j] += t[j];
    }
    for (int i = 0; i < nb; i++) step_body(ib[i], h);
    for (int i = 0; i < nb; i++) {
      Body<R> &b = bodies[ib[i]];
      zero(b.facc, 3);
      zero(b.tacc, 3);
    }
  }

  // ---------------------------------------------------------- dWorldStep
  // The island's LCP of step.cpp:420-710 -- A = J M^-1 J^T + cfm/h, b = rhs,
  // bounds lo/hi, friction rows bounded by findex -- solved as dSolveLCP
  // (lcp.cpp:784-1100) defines the answer: first the LCP of the rows without
  // a findex (friction impulses zero), then every friction row's bounds fixed
  // at +-|hi * lambda_normal| from that solution (0 when the normal impulse is
  // 0) and the LCP of all rows solved with those bounds (lcp.cpp:826-842).
  // With cfm > 0, A is positive definite and each of the two box-constrained
  // LCPs has exactly one solution, which this restatement computes by block
  // principal pivoting (Judice & Pires 1994) with a least-index single pivot
  // as the fallback that guarantees termination (Murty 1974), on a dense
  // Cholesky factorisation of the free rows.  Dantzig's pivoting in the
  // reference reaches the same solution in a different sequence of
  // arithmetic.

  // x for the rows of `free` (indices into 0..m-1): A_FF x_F = rhs_F by
  // Cholesky; false if A_FF is not positive definite
  static bool chol_solve(int m, const std::vector<R> &A, const std::vector<int> &fr, std::vector<R> &rhs) {
    const int n = (int)fr.size();
    std::vector<R> L((size_t)n * n, 0);
    for (int j = 0; j < n; j++) {
      R d = A[(size_t)fr[j] * m + fr[j]];
      for (int k = 0; k < j; k++) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
      if (!(d > 0)) return false;
      d = std::sqrt(d);
      L[(size_t)j * n + j] = d;
      for (int i = j + 1; i < n; i++) {
        R a = A[(size_t)fr[i] * m + fr[j]];
        for (int k = 0; k < j; k++) a -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
        L[(size_t)i * n + j] = a / d;
      }
    }
    for (int i = 0; i < n; i++) {
      R a = rhs[i];
      for (int k = 0; k < i; k++) a -= L[(size_t)i * n + k] * rhs[k];
      rhs[i] = a / L[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; i--) {
      R a = rhs[i];
      for (int k = i + 1; k < n; k++) a -= L[(size_t)k * n + i] * rhs[k];
      rhs[i] = a / L[(size_t)i * n + i];
    }
    return true;
  }

  // box-constrained LCP over the rows with in[i]: x in [lo,hi], w = A x - b,
  // x = lo => w >= 0, x = hi => w <= 0, else w = 0; rows outside `in` stay at
  // x = 0.  st: 0 free, 1 at lo, 2 at hi (start and result).  Returns the
  // number of pivoting rounds, or -1 when the cap is reached.
  static int blcp(int m, const std::vector<R> &A, const std::vector<R> &b, const std::vector<R> &lo,
                  const std::vector<R> &hi, const std::vector<char> &in, std::vector<R> &x, std::vector<int> &st) {
    int best = m + 1, backup = 3;
    for (int round = 0; round < 50 * m + 50; round++) {
      std::vector<int> fr;
      for (int i = 0; i < m; i++) {
        if (!in[i]) { x[i] = 0; continue; }
        if (st[i] == 0) fr.push_back(i);
        else x[i] = st[i] == 1 ? lo[i] : hi[i];
      }
      std::vector<R> r(fr.size());
      for (size_t a = 0; a < fr.size(); a++) {
        R v = b[fr[a]];
        for (int j = 0; j < m; j++)
          if (in[j] && st[j] != 0) v -= A[(size_t)fr[a] * m + j] * x[j];
        r[a] = v;
      }
      if (!chol_solve(m, A, fr, r)) return -1;
      for (size_t a = 0; a < fr.size(); a++) x[fr[a]] = r[a];
      std::vector<int> bad;
      for (int i = 0; i < m; i++) {
        if (!in[i]) continue;
        if (st[i] == 0) {
          if (x[i] < lo[i] || x[i] > hi[i]) bad.push_back(i);
        } else {
          R w = -b[i];
          for (int j = 0; j < m; j++)
            if (in[j]) w += A[(size_t)i * m + j] * x[j];
          if ((st[i] == 1 && w < 0) || (st[i] == 2 && w > 0)) bad.push_back(i);
        }
      }
      if (bad.empty()) return round;
      auto flip = [&](int i) {
        if (st[i] == 0) st[i] = x[i] < lo[i] ? 1 : 2;
        else st[i] = 0;
      };
      if ((int)bad.size() < best) {
        best = (int)bad.size();
        backup = 3;
        for (int i : bad) flip(i);
      } else if (backup > 0) {
        backup--;
        for (int i : bad) flip(i);
      } else {
        flip(bad.back());
      }
    }
    return -1;
  }

  void lcp_exact(int m, int nb, std::vector<R> &J, const std::vector<int> &jb, const std::vector<int> &ib,
                 const std::vector<R> &invI, std::vector<R> &lambda, std::vector<R> &fc, const std::vector<R> &b,
                 std::vector<R> lo, std::vector<R> hi, const std::vector<R> &cfm, const std::vector<int> &findex) {
    std::vector<R> iMJ(12 * (size_t)m, 0);
    for (int i = 0; i < m; i++) {
      int b1 = jb[2 * i], b2 = jb[2 * i + 1];
      R k1 = bodies[ib[b1]].invMass;
      for (int k = 0; k < 3; k++) iMJ[12 * i + k] = k1 * J[12 * i + k];
      mv(&iMJ[12 * i + 3], &invI[12 * b1], &J[12 * i + 3]);
      if (b2 != -1) {
        R k2 = bodies[ib[b2]].invMass;
        for (int k = 0; k < 3; k++) iMJ[12 * i + 6 + k] = k2 * J[12 * i + 6 + k];
        mv(&iMJ[12 * i + 9], &invI[12 * b2], &J[12 * i + 9]);
      }
    }
    // A_ij = J_i . M^-1 J_j^T over the bodies the two rows share
    std::vector<R> A((size_t)m * m, 0);
    for (int i = 0; i < m; i++)
      for (int j = 0; j < m; j++) {
        R s = 0;
        for (int si = 0; si < 2; si++) {
          const int bi = jb[2 * i + si];
          if (bi < 0) continue;
          for (int sj = 0; sj < 2; sj++)
            if (jb[2 * j + sj] == bi)
              for (int k = 0; k < 6; k++) s += J[12 * i + 6 * si + k] * iMJ[12 * j + 6 * sj + k];
        }
        A[(size_t)i * m + j] = s + (i == j ? cfm[i] : R(0));
      }
    std::vector<R> x(m, 0);
    std::vector<int> st(m, 0);
    std::vector<char> in(m);
    // rows with lo = hi (= 0) are held there whatever w is (lcp.cpp:832-838)
    for (int i = 0; i < m; i++) in[i] = findex[i] < 0 && lo[i] != hi[i];
    lcp_rounds_last = blcp(m, A, b, lo, hi, in, x, st);
    bool fric = false;
    for (int i = 0; i < m; i++) fric |= findex[i] >= 0;
    if (fric && lcp_rounds_last >= 0) {
      for (int i = 0; i < m; i++) {
        if (findex[i] < 0) continue;
        const R xn = x[findex[i]];
        if (xn == 0) {
          lo[i] = hi[i] = 0;
          in[i] = 0;       // held at zero (lcp.cpp:832-835 with w of either sign)
        } else {
          hi[i] = std::fabs(hi[i] * xn);
          lo[i] = -hi[i];
          in[i] = 1;
          st[i] = 0;
        }
      }
      lcp_rounds_last = blcp(m, A, b, lo, hi, in, x, st);
    }
    for (int i = 0; i < m; i++) lambda[i] = x[i];
    std::fill(fc.begin(), fc.end(), R(0));
    for (int i = 0; i < m; i++) {
      const int b1 = jb[2 * i], b2 = jb[2 * i + 1];
      for (int k = 0; k < 6; k++) fc[6 * (size_t)b1 + k] += iMJ[12 * i + k] * x[i];
      if (b2 != -1)
        for (int k = 0; k < 6; k++) fc[6 * (size_t)b2 + k] += iMJ[12 * i + 6 + k] * x[i];
    }
  }
  int lcp_rounds_last = 0;   // pivoting rounds of the last exact solve (-1: cap reached)

  void sor_lcp(int m, int nb, std::vector<R> &J, const std::vector<int> &jb, const std::vector<int> &ib,
               const std::vector<R> &invI, std::vector<R> &lambda, std::vector<R> &fc, std::vector<R> &b,
               const std::vector<R> &lo, const std::vector<R> &hi, const std::vector<R> &cfm,
               const std::vector<int> &findex) {
    std::vector<R> iMJ(12 * (size_t)m, 0);
    for (int i = 0; i < m; i++) {
      int b1 = jb[2 * i], b2 = jb[2 * i + 1];
      R k1 = bodies[ib[b1]].invMass;
      for (int k = 0; k < 3; k++) iMJ[12 * i + k] = k1 * J[12 * i + k];
      mv(&iMJ[12 * i + 3], &invI[12 * b1], &J[12 * i + 3]);
      if (b2 != -1) {
        R k2 = bodies[ib[b2]].invMass;
        for (int k = 0; k < 3; k++) iMJ[12 * i + 6 + k] = k2 * J[12 * i + 6 + k];
        mv(&iMJ[12 * i + 9], &invI[12 * b2], &J[12 * i + 9]);
      }
    }
    std::fill(fc.begin(), fc.end(), R(0));
    std::vector<R> Ad(m);
    for (int i = 0; i < m; i++) {
      R s = 0;
      for (int k = 0; k < 6; k++) s += iMJ[12 * i + k] * J[12 * i + k];
      if (jb[2 * i + 1] != -1) for (int k = 6; k < 12; k++) s += iMJ[12 * i + k] * J[12 * i + k];
      Ad[i] = prm.w / (s + cfm[i]);
    }
    // unscaled rows and rhs, for the residual after the iterations
    std::vector<R> J0, b0;
    if (track_residual) { J0 = J; b0 = b; }
    for (int i = 0; i < m; i++) {
      R a = Ad[i];
      for (int k = 0; k < 12; k++) J[12 * i + k] *= a;
      b[i] *= a;
      Ad[i] = a * cfm[i];
    }
    // unbounded-friction-free rows first, findex rows reversed at the tail
    std::vector<int> order(m);
    {
      int head = 0, tail = m - 1;
      for (int i = 0; i < m; i++) {
        if (findex[i] == -1) order[head++] = i;
        else order[tail--] = i;
      }
    }
    for (int it = 0; it < prm.iterations; it++) {
      if ((it & 7) == 0) {
        for (int i = 1; i < m; i++) {
          int s = rng.below(i + 1);
          std::swap(order[i], order[s]);
        }
      }
      for (int oi = 0; oi < m; oi++) {
        int r = order[oi];
        int b1 = jb[2 * r], b2 = jb[2 * r + 1];
        R *f1 = &fc[6 * (size_t)b1];
        R *f2 = b2 != -1 ? &fc[6 * (size_t)b2] : nullptr;
        R old = lambda[r];
        const R *Jr = &J[12 * (size_t)r];
        R delta = b[r] - old * Ad[r];
        delta -= f1[0] * Jr[0] + f1[1] * Jr[1] + f1[2] * Jr[2] + f1[3] * Jr[3] + f1[4] * Jr[4] + f1[5] * Jr[5];
        if (f2) delta -= f2[0] * Jr[6] + f2[1] * Jr[7] + f2[2] * Jr[8] + f2[3] * Jr[9] + f2[4] * Jr[10] + f2[5] * Jr[11];
        R lo_a, hi_a;
        if (findex[r] != -1) {
          hi_a = std::fabs(hi[r] * lambda[findex[r]]);
          lo_a = -hi_a;
        } else { hi_a = hi[r]; lo_a = lo[r]; }
        R nl = old + delta;
        if (nl < lo_a) { delta = lo_a - old; lambda[r] = lo_a; }
        else if (nl > hi_a) { delta = hi_a - old; lambda[r] = hi_a; }
        else lambda[r] = nl;
        const R *M = &iMJ[12 * (size_t)r];
        for (int k = 0; k < 6; k++) f1[k] += delta * M[k];
        if (f2) for (int k = 0; k < 6; k++) f2[k] += delta * M[6 + k];
      }
    }
    if (track_residual) {
      // max_i |lambda_i - clamp(lambda_i - (A lambda - b)_i / A_ii)| with the
      // unscaled A = J M^-1 J^T + cfm/h and b (not part of the reference,
      // which computes no residual)
      for (int r = 0; r < m; r++) {
        int b1 = jb[2 * r], b2 = jb[2 * r + 1];
        const R *Jr = &J0[12 * (size_t)r], *M = &iMJ[12 * (size_t)r];
        R s = 0, d = 0;
        for (int k = 0; k < 6; k++) { s += Jr[k] * fc[6 * (size_t)b1 + k]; d += Jr[k] * M[k]; }
        if (b2 != -1)
          for (int k = 0; k < 6; k++) { s += Jr[6 + k] * fc[6 * (size_t)b2 + k]; d += Jr[6 + k] * M[6 + k]; }
        R y = s + cfm[r] * lambda[r] - b0[r];
        R lo_a = lo[r], hi_a = hi[r];
        if (findex[r] != -1) { hi_a = std::fabs(hi[r] * lambda[findex[r]]); lo_a = -hi_a; }
        R dd = y / (d + cfm[r]), p = lambda[r] - dd, e;
        if (p < lo_a) e = lambda[r] - lo_a;
        else if (p > hi_a) e = hi_a - lambda[r];
        else e = dd;
        step_res = std::max(step_res, std::fabs(e));
      }
    }
  }
  bool track_residual = false;   // reference order: record the residual of each solve
  R step_res = 0;                // largest residual over this step's islands

  // Graph-coloured Gauss-Seidel over the island's constraints (the order of
  // the MI355X coloured mode, not a reference order): Jones-Plassmann
  // colouring with priority hash(i) = (i * 2654435761) mod 2^31 of the
  // constraint's id cid (contacts, then joints), colours swept in increasing order,
  // each constraint's rows relaxed in row order with the per-row update of
  // quickstep.cpp:484-582; no shuffle.
  void sor_colored(int m, int nb, std::vector<R> &J, const std::vector<int> &jb, const std::vector<int> &ib,
                   const std::vector<R> &invI, std::vector<R> &lambda, std::vector<R> &fc, std::vector<R> &b,
                   const std::vector<R> &lo, const std::vector<R> &hi, const std::vector<R> &cfm,
                   const std::vector<int> &findex, const std::vector<int> &row0, const std::vector<int> &jm,
                   const std::vector<int> &cid) {
    const int nc = (int)row0.size();
    std::vector<R> iMJ(12 * (size_t)m, 0);
    for (int i = 0; i < m; i++) {
      int b1 = jb[2 * i], b2 = jb[2 * i + 1];
      R k1 = bodies[ib[b1]].invMass;
      for (int k = 0; k < 3; k++) iMJ[12 * i + k] = k1 * J[12 * i + k];
      mv(&iMJ[12 * i + 3], &invI[12 * b1], &J[12 * i + 3]);
      if (b2 != -1) {
        R k2 = bodies[ib[b2]].invMass;
        for (int k = 0; k < 3; k++) iMJ[12 * i + 6 + k] = k2 * J[12 * i + 6 + k];
        mv(&iMJ[12 * i + 9], &invI[12 * b2], &J[12 * i + 9]);
      }
    }
    // fc = inv(M) J^T lambda of the starting impulses
    std::fill(fc.begin(), fc.end(), R(0));
    for (int i = 0; i < m; i++) {
      if (lambda[i] == 0) continue;
      int b1 = jb[2 * i], b2 = jb[2 * i + 1];
      for (int q = 0; q < 6; q++) fc[6 * (size_t)b1 + q] += lambda[i] * iMJ[12 * i + q];
      if (b2 != -1) for (int q = 0; q < 6; q++) fc[6 * (size_t)b2 + q] += lambda[i] * iMJ[12 * i + 6 + q];
    }
    std::vector<R> Ad(m);
    for (int i = 0; i < m; i++) {
      R s = 0;
      for (int k = 0; k < 6; k++) s += iMJ[12 * i + k] * J[12 * i + k];
      if (jb[2 * i + 1] != -1) for (int k = 6; k < 12; k++) s += iMJ[12 * i + k] * J[12 * i + k];
      Ad[i] = prm.w / (s + cfm[i]);
    }
    // colouring: Jones-Plassmann rounds; each round's winners, in
    // constraint-id order, take the lowest colour below 64 that neither of
    // their bodies has and that has fewer than colour_width members (first
    // fit, capped at the MI355X workgroup width); past that, a colour of
    // their own per round
    std::vector<int> col(nc, -1), ccnt(64, 0), win;
    if (colour_log) colour_log->clear();
    std::vector<long long> bmax(nb);
    std::vector<unsigned long long> used_mask(nb, 0ull);
    int ncol = 0, round = 0;
    auto prio = [](int i) { return (long long)(((unsigned)i * 2654435761u) & 0x7fffffffu); };
    while (true) {
      std::fill(bmax.begin(), bmax.end(), -1LL);
      int left = 0;
      for (int i = 0; i < nc; i++) {
        if (col[i] != -1 || jm[i] <= 0) continue;
        left++;
        int r = row0[i];
        bmax[jb[2 * r]] = std::max(bmax[jb[2 * r]], prio(cid[i]));
        if (jb[2 * r + 1] != -1) bmax[jb[2 * r + 1]] = std::max(bmax[jb[2 * r + 1]], prio(cid[i]));
      }
      if (!left) break;
      win.clear();
      for (int i = 0; i < nc; i++) {
        if (col[i] != -1 || jm[i] <= 0) continue;
        int r = row0[i];
        const int b1 = jb[2 * r], b2 = jb[2 * r + 1];
        if (bmax[b1] == prio(cid[i]) && (b2 == -1 || bmax[b2] == prio(cid[i]))) win.push_back(i);
      }
      std::sort(win.begin(), win.end(), [&](int x, int y) { return cid[x] < cid[y]; });
      for (int i : win) {
        int r = row0[i];
        const int b1 = jb[2 * r], b2 = jb[2 * r + 1];
        const unsigned long long used = used_mask[b1] | (b2 != -1 ? used_mask[b2] : 0ull);
        int cc = 0;
        while (cc < 64 && ((used >> cc) & 1ull || ccnt[cc] >= prm.colour_width)) cc++;
        if (cc == 64) {
          cc = 64 + round;
        } else {
          ccnt[cc]++;
          used_mask[b1] |= 1ull << cc;
          if (b2 != -1) used_mask[b2] |= 1ull << cc;
        }
        col[i] = cc;
        ncol = std::max(ncol, cc + 1);
        if (colour_log) colour_log->insert(colour_log->end(), {cid[i], b1, b2, cc, round});
      }
      round++;
    }
    auto sweep = [&]() {
      for (int g = 0; g < ncol; g++)
        for (int i = 0; i < nc; i++) {
          if (col[i] != g) continue;
          for (int k = 0; k < jm[i]; k++) {
            int r = row0[i] + k;
            int b1 = jb[2 * r], b2 = jb[2 * r + 1];
            R *f1 = &fc[6 * (size_t)b1];
            R *f2 = b2 != -1 ? &fc[6 * (size_t)b2] : nullptr;
            const R *Jr = &J[12 * (size_t)r];
            R s = f1[0] * Jr[0] + f1[1] * Jr[1] + f1[2] * Jr[2] + f1[3] * Jr[3] + f1[4] * Jr[4] + f1[5] * Jr[5];
            if (f2) s += f2[0] * Jr[6] + f2[1] * Jr[7] + f2[2] * Jr[8] + f2[3] * Jr[9] + f2[4] * Jr[10] + f2[5] * Jr[11];
            R old = lambda[r];
            R delta = Ad[r] * (b[r] - old * cfm[r] - s);
            R lo_a, hi_a;
            if (findex[r] != -1) {
              hi_a = std::fabs(hi[r] * lambda[findex[r]]);
              lo_a = -hi_a;
            } else { hi_a = hi[r]; lo_a = lo[r]; }
            R nl = old + delta;
            if (nl < lo_a) { delta = lo_a - old; nl = lo_a; }
            else if (nl > hi_a) { delta = hi_a - old; nl = hi_a; }
            else if (prm.tol > 0) delta = nl - old;   // fc follows the stored impulse
            lambda[r] = nl;
            const R *M = &iMJ[12 * (size_t)r];
            for (int q = 0; q < 6; q++) f1[q] += delta * M[q];
            if (f2) for (int q = 0; q < 6; q++) f2[q] += delta * M[6 + q];
          }
        }
    };
    // LCP residual max_i |lambda_i - clamp(lambda_i - (A lambda - b)_i / A_ii)|
    auto residual = [&]() {
      R worst = 0;
      for (int r = 0; r < m; r++) {
        int b1 = jb[2 * r], b2 = jb[2 * r + 1];
        const R *Jr = &J[12 * (size_t)r];
        const R *f1 = &fc[6 * (size_t)b1];
        R s = 0;
        for (int q = 0; q < 6; q++) s += Jr[q] * f1[q];
        if (b2 != -1) for (int q = 0; q < 6; q++) s += Jr[6 + q] * fc[6 * (size_t)b2 + q];
        R y = s + cfm[r] * lambda[r] - b[r];
        R lo_a = lo[r], hi_a = hi[r];
        if (findex[r] != -1) { hi_a = std::fabs(hi[r] * lambda[findex[r]]); lo_a = -hi_a; }
        // |lambda - clamp(lambda - d)| without the cancellation of
        // lambda - (lambda - d): |d| when the step stays inside the bounds
        R d = y * (Ad[r] / prm.w);
        R p = lambda[r] - d, e;
        if (p < lo_a) e = lambda[r] - lo_a;
        else if (p > hi_a) e = hi_a - lambda[r];
        else e = d;
        worst = std::max(worst, std::fabs(e));
      }
      return worst;
    };
    const bool use_tol = prm.tol > 0;
    last_res = 0;
    const int maxit = use_tol ? prm.max_iterations : prm.iterations;
    std::vector<R> lprev(m, 0), dir(m, 0), pend(m, 0);
    R rr_prev = 0, res_prev = INFINITY;
    bool accel_off = false;
    for (int it = 0; it < maxit; it++) {
      if (prm.accel && !accel_off) {
        // the pending momentum beta p_prev of the previous sweep
        for (int r = 0; r < m; r++) {
            R dl = pend[r];
            if (dl == 0) continue;
            lambda[r] += dl;
            int b1 = jb[2 * r], b2 = jb[2 * r + 1];
            const R *M = &iMJ[12 * (size_t)r];
            for (int q = 0; q < 6; q++) fc[6 * (size_t)b1 + q] += dl * M[q];
            if (b2 != -1) for (int q = 0; q < 6; q++) fc[6 * (size_t)b2 + q] += dl * M[6 + q];
          }
        for (int r = 0; r < m; r++) lprev[r] = lambda[r];
      }
      sweep();
      prm.sweeps++;
      if (prm.accel && !accel_off) {
        // nonsmooth nonlinear CG: d = PGS(l) - l, beta = |d|^2 / |d_prev|^2,
        // restart when beta > 1; l += beta p_prev (before the next sweep),
        // p = d + beta p_prev
        R rr = 0;
        for (int r = 0; r < m; r++) { R d = lambda[r] - lprev[r]; rr += d * d; }
        R beta = rr_prev > 0 ? rr / rr_prev : 0;
        if (!(beta <= 1)) beta = 0;
        for (int r = 0; r < m; r++) {
          pend[r] = beta * dir[r];
          dir[r] = (lambda[r] - lprev[r]) + pend[r];
        }
        rr_prev = rr;
      }
      if (use_tol && (it + 1) % prm.check_every == 0) {
        const R res = residual();
        last_res = res;
        if (res <= prm.tol) return;
        if (prm.accel && !accel_off && !(res <= res_prev)) {
          // the residual rose since the last check: restart the CG steps,
          // and stop them for the rest of the solve when the residual is
          // still far (> 1000 tol) from the tolerance
          std::fill(pend.begin(), pend.end(), R(0));
          std::fill(dir.begin(), dir.end(), R(0));
          rr_prev = 0;
          if (res > 1000 * prm.tol) accel_off = true;
        }
        res_prev = res;
      }
    }
    if (use_tol) prm.unconverged++;
  }

  // dxStepBody, default flags: ode/src/util.cpp:255-360
  void step_body(int bi, R h) {
    Body<R> &b = bodies[bi];
    for (int j = 0; j < 3; j++) b.pos[j] += h * b.lvel[j];
    R dq[4];
    quat_rate(dq, b.avel, b.q);
    for (int j = 0; j < 4; j++) b.q[j] += h * dq[j];
    normalize4(b.q);
    quat_to_rot(b.Rm, b.q);
    // dGeomMoved (collision_space.cpp:47-75): moved geoms go to the front
    for (int g : b.geoms) moved.push_back(g);
  }
  std::vector<int> moved;

  // dWorldQuickStep: ode.cpp:1807-1825; islands util.cpp:384-568
  void quickstep(R h) {
    moved.clear();
    step_res = 0;
    int nbod = (int)bodies.size();
    for (Body<R> &b : bodies) b.tag = 0;
    std::vector<char> jtag(joints.size(), 0);
    std::vector<int> stack;
    // the world body list is newest first
    for (int bb = nbod - 1; bb >= 0; bb--) {
      if (bodies[bb].tag) continue;
      bodies[bb].tag = 1;
      std::vector<int> ib, ij;
      ib.push_back(bb);
      stack.clear();
      int b = bb;
      while (true) {
        for (int ji : bodies[b].jlist) {
          if (jtag[ji]) continue;
          const Joint<R> &jt = joints[ji];
          bool enabled = bodies[jt.node0].invMass > 0 || (jt.node1 >= 0 && bodies[jt.node1].invMass > 0);
          if (enabled) {
            jtag[ji] = 1;
            ij.push_back(ji);
            int nbd = other_body(ji, b);
            if (nbd >= 0 && bodies[nbd].tag <= 0) {
              bodies[nbd].tag = 1;
              stack.push_back(nbd);
            }
          } else {
            jtag[ji] = 2;
          }
        }
        if (stack.empty()) break;
        b = stack.back();
        stack.pop_back();
        ib.push_back(b);
      }
      islands_b.push_back(ib);
      islands_j.push_back(ij);
    }
    if (prm.colored) {
      // the coloured mode solves the whole world at once (one colouring, one
      // sweep over all islands, world-wide stopping test and CG step), as
      // the MI355X coloured mode does; islands are independent, so a plain
      // sweep is the same either way
      std::vector<int> ab, aj;
      for (size_t k = 0; k < islands_b.size(); k++) {
        ab.insert(ab.end(), islands_b[k].begin(), islands_b[k].end());
        aj.insert(aj.end(), islands_j[k].begin(), islands_j[k].end());
      }
      last_res = 0;
      quickstep_island(ab, aj, h);
      // this step's contact impulses, the next step's starting point
      warm_prev.clear();
      // a solve that ended at the cap far from the tolerance (residual
      // > 100 tol) is no starting point: the next one starts from zero
      warm_ok = prm.warm != 0 && last_res <= 100 * prm.tol;
      if (warm_ok)
        for (size_t k = n_perm; k < joints.size(); k++) {
          const Joint<R> &jt = joints[k];
          warm_prev.push_back({jt.wkey, std::array<R, 3>{jt.lam[0], jt.lam[1], jt.lam[2]}});
        }
      std::sort(warm_prev.begin(), warm_prev.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    } else {
      for (size_t k = 0; k < islands_b.size(); k++) quickstep_island(islands_b[k], islands_j[k], h);
    }
    islands_b.clear();
    islands_j.clear();
    // new space list: moved geoms in reverse order of their move, then the
    // untouched ones in their old relative order
    std::vector<char> is_moved(geoms.size(), 0), in_first(geoms.size(), 0);
    for (int g : space) in_first[g] = 1;
    std::vector<int> ns;
    for (auto it = moved.rbegin(); it != moved.rend(); ++it) {
      if (in_first[*it] && !is_moved[*it]) { is_moved[*it] = 1; ns.push_back(*it); }
    }
    for (int g : space) if (!is_moved[g]) ns.push_back(g);
    space.swap(ns);
  }
  std::vector<std::vector<int>> islands_b, islands_j;
  std::vector<std::pair<long long, std::array<R, 3>>> warm_prev;   // sorted by key
  R last_res = 0;   // residual at the last check of the last coloured solve (0: none)
  // when set, the last coloured solve's colouring: (cid, b1, b2, colour,
  // round) per constraint, in the order the colours were assigned
  std::vector<int> *colour_log = nullptr;
  bool warm_ok = false;   // warm_prev and the joints' impulses hold a usable start
};

}  // namespace orc
