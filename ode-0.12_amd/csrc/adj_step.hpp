// adj_step.hpp -- reverse-mode derivative of one quickstep of one world with
// contact joints (spheres, boxes, planes), ball joints, hinges (motor rows
// included, not limit rows) and hinge2 joints (suspension, steering limit,
// motors, added torques), written once over the scalar type
// and a parallel context so that the GPU runs it with one workgroup per world
// (float) and the host runs the same code serially (double) for its checks.
//
// Forward step being differentiated (quickstep.cpp:598-1021, util.cpp:255-360,
// joints/contact.cpp:75-256): body preprocessing (world inertia, gyroscopic
// torque, gravity), contact rows J = [d, r1 x d, -d, -(r2 x d)] for d in
// {n, t1, t2} (t1, t2 = dPlaneSpace(n)), c = erp/h * max(depth - min_depth, 0)
// or the bounce term, b = c/h - J (v/h + M^-1 f), the LCP
//   A lambda - b  in the normal cone of [lo, hi],  A = J M^-1 J^T + cfm/h,
// friction bounds +-mu * lambda_normal, joint rows (joint.cpp:85-196 and
// 556-733, hinge.cpp:76-149, hinge2.cpp:127-180) with their own bounds, then
// v' = v + h M^-1 (J^T lambda + f),
// x' = x + h v', q' = normalize(q + h/2 (0, w') q).
//
// The LCP is differentiated implicitly at its (converged) solution on the
// active set: free rows satisfy (A lambda)_i = b_i, rows at a bound follow
// the bound (friction: lambda_f = sigma mu lambda_n).  The adjoint system
// K^T y = T^T lambda_bar (K = (A T)_FF) is solved with BiCGSTAB and a Jacobi
// preconditioner.  Collision geometry is differentiated by running the
// colliders on dual numbers (adj_colliders.hpp).  Rotational derivatives use
// a world-frame rotation vector theta (R -> exp([theta]x) R); the quaternion
// adjoint is q_bar = 2 (0, theta_bar) (x) q on top of the integrator's terms.
#pragma once
#include "adj_colliders.hpp"

namespace odemi {
namespace adj {

// contact record of the tape (the step's contacts in creation order)
template <class T>
struct TapeContact {
  T pos[3], normal[3], depth;   // as the collider returned them for (g1, g2)
  T lam[3];                      // converged impulses of the contact's rows
  int g1, g2, b1, b2;            // b1 >= 0 is the first body after attach
  int flags;                     // JF_REVERSE: the normal is negated for the rows
  int m;                         // rows (1 or 3)
};

// permanent-joint record of the tape: a ball joint, a hinge (its motor row
// included; not its limit row) or a hinge2 (suspension, steering limit and
// motors), with its frame data as the step used it, each row's bounds and
// cfm, and the converged impulses of its rows
template <class T>
struct TapeJoint {
  T lam[6];
  T lo[6], hi[6];       // row bounds
  T rcfm[6];            // row cfm / h
  T anc1[3], anc2[3];   // anchors in the body frames (anc2: world point without a second body)
  T ax1[3], ax2[3];     // axes in the body frames (ax2: world axis without a second body)
  T erp;                // the joint's Info2 erp
  T susp_erp;           // hinge2: erp of the suspension row
  T c0, s0;             // hinge2: sine and cosine terms of the rest angle
  T v1[3], v2[3];       // hinge2: steering-angle frame in body 1
  T lerp[2];            // limit row of limot k: the stop erp, c = -erp/h (angle - lstop)
  T lstop[2];
  T lc[2];              // motor row of limot k: c (the target velocity)
  T torque[2];          // hinge2 torques added this step about axis 1 and axis 2
  int type;             // JT_BALL, JT_HINGE or JT_HINGE2
  int b1, b2;           // node order after attach (b2 may be -1)
  int m;                // rows
  int lrow[2];          // row of limot 1 / 2, -1 without one
  int lkind[2];         // 1: limit row, 2: motor row
};

template <class T>
struct StepParams {
  T h, erp, cfm, min_depth, max_vel, gravity[3];
  T mu, bounce, bounce_vel;
  int surf_bounce;   // dContactBounce set
  int maxc;          // contacts per pair requested
  T act_tol;         // impulses at most this are treated as at their bound
  T lin_tol;         // BiCGSTAB relative residual
  int lin_maxit;
};

// one world's data for one step
template <class T>
struct StepData {
  int nb, nk, ng;
  // pre-step body state, stride 4: pos, q (w x y z), lvel, avel
  const T *x, *q, *v, *w;
  // body constants: mass, 1/mass, body inertia and its inverse (3 rows of 4)
  const T *mass, *invMass, *I, *invI;
  // geoms: type, body, params (4), static pose (pos stride 4, R 3 rows of 4)
  const int *gtype, *gbody;
  const T *gprm, *gpos, *gR;
  const TapeContact<T> *con;
  // permanent joints of the step (their rows follow the contacts' rows)
  int nj = 0;
  const TapeJoint<T> *jnt = nullptr;
};

// adjoints: in = of the post-step state, out = of the pre-step state
// (stride 4 each, pos / q / lvel / avel)
template <class T>
struct StepAdj {
  const T *gx_in, *gq_in, *gv_in, *gw_in;
  T *gx, *gq, *gv, *gw;
  T *gmu, *gbounce;   // accumulated (one scalar each)
  int *stats;         // [0] BiCGSTAB iterations, [1] collider count mismatches
};

// scratch sizes
constexpr int ADJ_BODY_FLOATS = 64;
constexpr int ADJ_ROW_FLOATS = 33;
// row slots: 3 per contact, then 6 per permanent joint
template <class T> OD_HD inline long long adj_scratch_elems(int nb, int nk, int nj = 0) {
  return (long long)ADJ_BODY_FLOATS * nb + (long long)ADJ_ROW_FLOATS * (3LL * nk + 6LL * nj) + 16;
}

// per-body scratch fields
enum {
  BR = 0,        // R 9
  BIW = 9,       // I_w 9
  BIIW = 18,     // I_w^-1 9
  BTAU = 27,     // gyroscopic + external torque 3
  BZ = 30,       // v/h + M^-1 f 6
  BJL = 36,      // J^T lambda 6
  BJY = 42,      // J^T y 6 (matvec) / J^T b_bar afterwards
  BU = 48,       // h M^-1 (v_bar', w_bar') 6
  BTH = 54,      // theta_bar 3
  BTAUB = 57,    // tau_bar 3
  BWB = 60       // w_bar accumulated from contacts 3 (v_bar in BX below)
};
// per-row scratch fields
enum {
  RJ = 0,        // J 12
  RJB = 12,      // J_bar 12
  RLB = 24,      // lambda_bar
  RY = 25,       // BiCGSTAB x (the adjoint y)
  RR = 26, RRH = 27, RP = 28, RV = 29, RS = 30, RT = 31,  // BiCGSTAB vectors
  RYB = 32       // the iterate with the smallest residual so far
};

template <class T> OD_HD inline V3<T> ld3(const T *p) { return v3(p[0], p[1], p[2]); }
template <class T> OD_HD inline void st3(T *p, const V3<T> &a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
template <class T> OD_HD inline M3<T> ldm(const T *p) {   // 9 floats, row-major
  M3<T> m;
  m.r0 = v3(p[0], p[1], p[2]);
  m.r1 = v3(p[3], p[4], p[5]);
  m.r2 = v3(p[6], p[7], p[8]);
  return m;
}
template <class T> OD_HD inline void stm(T *p, const M3<T> &m) {
  st3(p, m.r0); st3(p + 3, m.r1); st3(p + 6, m.r2);
}
template <class T> OD_HD inline M3<T> ldm4(const T *p) {  // 3 rows of 4
  M3<T> m;
  m.r0 = v3(p[0], p[1], p[2]);
  m.r1 = v3(p[4], p[5], p[6]);
  m.r2 = v3(p[8], p[9], p[10]);
  return m;
}

// Row state of contact k: free / bound flags and the sign of bound friction
template <class T>
struct RowAct {
  bool nfree;        // normal row free
  bool ffree[2];     // friction rows free
  T sig[2];          // sign of a friction row at its bound (0 if free or fixed)
};
template <class T> OD_HD inline RowAct<T> row_activity(const TapeContact<T> &c, const StepParams<T> &P) {
  RowAct<T> a;
  a.nfree = c.lam[0] > P.act_tol;
  a.ffree[0] = a.ffree[1] = false;
  a.sig[0] = a.sig[1] = T(0);
  if (c.m < 3 || !a.nfree) return a;
  T bound = P.mu * c.lam[0];
  for (int f = 0; f < 2; f++) {
    T l = c.lam[1 + f];
    if (std::fabs(l) >= bound - (P.act_tol + T(1e-4) * bound)) a.sig[f] = l >= T(0) ? T(1) : T(-1);
    else a.ffree[f] = true;
  }
  return a;
}

// body block (6) of a row times a body 6-vector
template <class T> OD_HD inline T jdot6(const T *J, const T *u) {
  return J[0] * u[0] + J[1] * u[1] + J[2] * u[2] + J[3] * u[3] + J[4] * u[4] + J[5] * u[5];
}

// M^-1 times a body 6-vector (linear: 1/m, angular: I_w^-1)
template <class T> OD_HD inline void minv6(const T *bs, T im, const T *u, T *o) {
  o[0] = im * u[0]; o[1] = im * u[1]; o[2] = im * u[2];
  M3<T> iI = ldm(bs + BIIW);
  st3(o + 3, iI.mul(v3(u[3], u[4], u[5])));
}

// dq = 1/2 (0, w) (x) q  (rotation.cpp:310-317) and its adjoint
template <class T> OD_HD inline void dq_of(const V3<T> &w, const T *q, T *dq) {
  dq[0] = T(0.5) * (-w.x * q[1] - w.y * q[2] - w.z * q[3]);
  dq[1] = T(0.5) * (w.x * q[0] + w.y * q[3] - w.z * q[2]);
  dq[2] = T(0.5) * (-w.x * q[3] + w.y * q[0] + w.z * q[1]);
  dq[3] = T(0.5) * (w.x * q[2] - w.y * q[1] + w.z * q[0]);
}
template <class T> OD_HD inline void dq_adj(const V3<T> &w, const T *q, const T *a, T *qb, V3<T> &wb) {
  qb[0] += T(0.5) * (w.x * a[1] + w.y * a[2] + w.z * a[3]);
  qb[1] += T(0.5) * (-w.x * a[0] + w.z * a[2] - w.y * a[3]);
  qb[2] += T(0.5) * (-w.y * a[0] - w.z * a[1] + w.x * a[3]);
  qb[3] += T(0.5) * (-w.z * a[0] + w.y * a[1] - w.x * a[2]);
  wb.x += T(0.5) * (-q[1] * a[0] + q[0] * a[1] - q[3] * a[2] + q[2] * a[3]);
  wb.y += T(0.5) * (-q[2] * a[0] + q[3] * a[1] + q[0] * a[2] - q[1] * a[3]);
  wb.z += T(0.5) * (-q[3] * a[0] - q[2] * a[1] + q[1] * a[2] + q[0] * a[3]);
}

// Rows of contact k (normal, t1, t2), J stored 12 per row; returns the
// row normal direction sign convention applied (reverse flag).
template <class T>
OD_HD inline void contact_rows(const StepData<T> &D, const TapeContact<T> &c, T *J3) {
  V3<T> n = ld3(c.normal);
  if (c.flags & JF_REVERSE) n = -n;
  V3<T> p = ld3(c.pos);
  V3<T> r1 = p - ld3(D.x + 4 * c.b1);
  V3<T> r2 = c.b2 >= 0 ? p - ld3(D.x + 4 * c.b2) : v3(T(0), T(0), T(0));
  V3<T> d[3];
  d[0] = n;
  if (c.m >= 3) plane_space(n, d[1], d[2]);
  for (int r = 0; r < c.m; r++) {
    T *J = J3 + ADJ_ROW_FLOATS * r;
    st3(J, d[r]);
    st3(J + 3, cross(r1, d[r]));
    if (c.b2 >= 0) {
      st3(J + 6, -d[r]);
      st3(J + 9, -cross(r2, d[r]));
    } else {
      for (int k = 6; k < 12; k++) J[k] = T(0);
    }
  }
}

// Rows of a permanent joint (ball.cpp:56-64 with the ball rows of
// joint.cpp:85-126; hinge.cpp:76-149; hinge2.cpp:127-180 with the aligned
// ball rows of joint.cpp:134-196; the limit/motor rows of joint.cpp:556-733):
// J (12 per row, rows `stride` apart) and c, over a scalar type S so that the
// same code gives the rows (S = T) and their derivatives (S = Dual).
template <class T, class S>
OD_HD inline void joint_rows(const TapeJoint<T> &j, const V3<S> &x1, const M3<S> &R1, const V3<S> &x2,
                             const M3<S> &R2, T hinv, S *J, int stride, S *c) {
  const bool two = j.b2 >= 0;
  const S one = S(T(1)), zero = S(T(0));
  for (int r = 0; r < j.m; r++) {
    for (int i = 0; i < 12; i++) J[stride * r + i] = zero;
    c[r] = zero;
  }
  const T k = j.erp * hinv;
  const V3<S> a1 = R1.mul(v3(S(j.anc1[0]), S(j.anc1[1]), S(j.anc1[2])));
  const V3<S> a2 = two ? R2.mul(v3(S(j.anc2[0]), S(j.anc2[1]), S(j.anc2[2]))) : v3(zero, zero, zero);
  // separation of the anchors: second anchor (or the world point) - first
  const V3<S> d = two ? ((a2 + x2) - a1) - x1 : (v3(S(j.anc2[0]), S(j.anc2[1]), S(j.anc2[2])) - a1) - x1;
  const V3<S> ax1 = R1.mul(v3(S(j.ax1[0]), S(j.ax1[1]), S(j.ax1[2])));
  const V3<S> wa2 = v3(S(j.ax2[0]), S(j.ax2[1]), S(j.ax2[2]));
  const V3<S> ax2 = two ? R2.mul(wa2) : wa2;
  // an angular row along u: J_a1 = u, J_a2 = -u
  auto ang_row = [&](int r, const V3<S> &u) {
    S *Jr = J + stride * r;
    Jr[3] = u.x; Jr[4] = u.y; Jr[5] = u.z;
    if (two) { Jr[9] = -u.x; Jr[10] = -u.y; Jr[11] = -u.z; }
  };
  int ax_rows = 0;   // first row of the limit / motor rows
  if (j.type == JT_HINGE2) {
    // three ball rows along ax1 and its plane space, the first with the
    // suspension erp
    V3<S> q1, q2;
    plane_space(ax1, q1, q2);
    const V3<S> dirs[3] = {ax1, q1, q2};
    for (int r = 0; r < 3; r++) {
      S *Jr = J + stride * r;
      const V3<S> u = dirs[r], c1 = cross(a1, u);
      Jr[0] = u.x; Jr[1] = u.y; Jr[2] = u.z;
      Jr[3] = c1.x; Jr[4] = c1.y; Jr[5] = c1.z;
      if (two) {
        const V3<S> c2 = cross(a2, u);
        Jr[6] = -u.x; Jr[7] = -u.y; Jr[8] = -u.z;
        Jr[9] = -c2.x; Jr[10] = -c2.y; Jr[11] = -c2.z;
      }
      c[r] = (r == 0 ? j.susp_erp * hinv : k) * dot(u, d);
    }
    // the row holding the angle between the axes: along ax1 x ax2 normalised
    const V3<S> axc = cross(ax1, ax2);
    const S sn = sqrt_(dot(axc, axc)), cs = dot(ax1, ax2);
    ang_row(3, (one / sn) * axc);
    c[3] = k * (j.c0 * sn - j.s0 * cs);
    ax_rows = 4;
  } else {
    S *J0 = J, *J1 = J + stride, *J2 = J + 2 * stride;
    // J_l1 = I, J_a1 = -[a1]x
    J0[0] = one; J1[1] = one; J2[2] = one;
    J0[4] = a1.z;  J0[5] = -a1.y;
    J1[3] = -a1.z; J1[5] = a1.x;
    J2[3] = a1.y;  J2[4] = -a1.x;
    if (two) {
      // J_l2 = -I, J_a2 = [a2]x
      J0[6] = -one; J1[7] = -one; J2[8] = -one;
      J0[10] = -a2.z; J0[11] = a2.y;
      J1[9] = a2.z;   J1[11] = -a2.x;
      J2[9] = -a2.y;  J2[10] = a2.x;
    }
    c[0] = k * d.x;
    c[1] = k * d.y;
    c[2] = k * d.z;
    if (j.type != JT_HINGE) return;
    // the two rows keeping the hinge axes aligned
    V3<S> p, q;
    plane_space(ax1, p, q);
    ang_row(3, p);
    ang_row(4, q);
    const V3<S> b = cross(ax1, ax2);
    c[3] = k * dot(b, p);
    c[4] = k * dot(b, q);
    ax_rows = 5;
  }
  (void)ax_rows;
  for (int l = 0; l < 2; l++) {
    const int r = j.lrow[l];
    if (r < 0) continue;
    ang_row(r, l == 0 ? ax1 : ax2);
    if (j.lkind[l] == 2) {
      c[r] = S(j.lc[l]);
    } else {
      // hinge2 steering angle (hinge2.cpp:34-43): -atan2(v2 . a, v1 . a)
      // with a = R1^T R2 axis2
      const V3<S> a = R1.tmul(R2.mul(wa2));
      const S ang = -atan2_(dot(v3(S(j.v2[0]), S(j.v2[1]), S(j.v2[2])), a),
                            dot(v3(S(j.v1[0]), S(j.v1[1]), S(j.v1[2])), a));
      c[r] = -(j.lerp[l] * hinv) * (ang - j.lstop[l]);
    }
  }
}

// row r of a joint is free (strictly inside its bounds)
template <class T> OD_HD inline bool joint_row_free(const TapeJoint<T> &j, int r, T tol) {
  return j.lam[r] > j.lo[r] + tol && j.lam[r] < j.hi[r] - tol;
}

// a body pose as dual numbers: the position seeded with the unit directions
// (kind 0) or the rotation R -> exp([theta]x) R seeded along theta (kind 1),
// or constant (kind -1)
template <class T>
OD_HD inline void pose_dual(const V3<T> &p, const M3<T> &R, int kind, V3<Dual<T, 3>> &pd, M3<Dual<T, 3>> &Rd) {
  typedef Dual<T, 3> S;
  pd = v3(S(p.x), S(p.y), S(p.z));
  Rd.r0 = v3(S(R.r0.x), S(R.r0.y), S(R.r0.z));
  Rd.r1 = v3(S(R.r1.x), S(R.r1.y), S(R.r1.z));
  Rd.r2 = v3(S(R.r2.x), S(R.r2.y), S(R.r2.z));
  if (kind == 0) {
    pd.x.d[0] = T(1); pd.y.d[1] = T(1); pd.z.d[2] = T(1);
  } else if (kind == 1) {
    // d R / d theta_j = [e_j]x R: column cc gets e_j x R_cc
    for (int cc = 0; cc < 3; cc++) {
      const V3<T> col = R.col(cc);
      S *e0 = cc == 0 ? &Rd.r0.x : (cc == 1 ? &Rd.r0.y : &Rd.r0.z);
      S *e1 = cc == 0 ? &Rd.r1.x : (cc == 1 ? &Rd.r1.y : &Rd.r1.z);
      S *e2 = cc == 0 ? &Rd.r2.x : (cc == 1 ? &Rd.r2.y : &Rd.r2.z);
      e0->d[0] = T(0);    e1->d[0] = -col.z; e2->d[0] = col.y;
      e0->d[1] = col.z;   e1->d[1] = T(0);   e2->d[1] = -col.x;
      e0->d[2] = -col.y;  e1->d[2] = col.x;  e2->d[2] = T(0);
    }
  }
}

// The adjoint of one step of one world.  Par provides tid(), nt(), sync(),
// sum(T) (all-reduce) and add(T*, T) (atomic accumulate).
template <class T, class Par>
OD_HD void adjoint_step(Par &par, const StepData<T> &D, const StepParams<T> &P, StepAdj<T> &G, T *scr) {
  const int nb = D.nb, nk = D.nk, nj = D.nj;
  const int tid = par.tid(), nt = par.nt();
  const T h = P.h, hinv = T(1) / P.h;
  // row slots: 3 per contact (slot 3k + r), then 6 per joint (3 nk + 6 j + r)
  const int ncs = 3 * nk, nslots = ncs + 6 * nj;
  T *bsc = scr;                                           // ADJ_BODY_FLOATS per body
  T *rsc = scr + (long long)ADJ_BODY_FLOATS * nb;         // ADJ_ROW_FLOATS per row slot
  T *wsc = rsc + (long long)ADJ_ROW_FLOATS * nslots;      // 16 world scalars
  // bodies and cfm / h of a row slot; false for a slot beyond its
  // constraint's rows
  auto slot_rows = [&](int s, int &b1, int &b2, T &cfm_h) -> bool {
    if (s < ncs) {
      const TapeContact<T> &c = D.con[s / 3];
      b1 = c.b1; b2 = c.b2; cfm_h = P.cfm * hinv;
      return s % 3 < c.m;
    }
    const TapeJoint<T> &jt = D.jnt[(s - ncs) / 6];
    const int r = (s - ncs) % 6;
    b1 = jt.b1; b2 = jt.b2; cfm_h = r < jt.m ? jt.rcfm[r] : T(0);
    return r < jt.m;
  };
  auto joint_slots = [&](int j) -> T * { return rsc + ADJ_ROW_FLOATS * (long long)(ncs + 6 * j); };

  // ---- 0. bodies: R, I_w, I_w^-1, torque, Z; clear accumulators
  for (int b = tid; b < nb; b += nt) {
    T *bs = bsc + ADJ_BODY_FLOATS * b;
    M3<T> R = quat_rot(D.q + 4 * b);
    M3<T> Iw = rotate_tensor(R, ldm4(D.I + 12 * b));
    M3<T> iIw = rotate_tensor(R, ldm4(D.invI + 12 * b));
    stm(bs + BR, R);
    stm(bs + BIW, Iw);
    stm(bs + BIIW, iIw);
    V3<T> w = ld3(D.w + 4 * b);
    V3<T> tau = -cross(w, Iw.mul(w));
    st3(bs + BTAU, tau);
    for (int k = BJL; k < ADJ_BODY_FLOATS; k++) bs[k] = T(0);
  }
  if (tid == 0)
    for (int k = 0; k < 16; k++) wsc[k] = T(0);
  par.sync();
  // hinge2 torques of the step (hinge2.cpp:393-410): t = t1 R1 axis1 + t2 R2
  // axis2 on body 1, -t on body 2
  for (int j = tid; j < nj; j += nt) {
    const TapeJoint<T> &jt = D.jnt[j];
    if (jt.type != JT_HINGE2 || jt.b2 < 0 || (jt.torque[0] == T(0) && jt.torque[1] == T(0))) continue;
    const V3<T> a1 = ldm(bsc + ADJ_BODY_FLOATS * jt.b1 + BR).mul(ld3(jt.ax1));
    const V3<T> a2 = ldm(bsc + ADJ_BODY_FLOATS * jt.b2 + BR).mul(ld3(jt.ax2));
    const V3<T> t = jt.torque[0] * a1 + jt.torque[1] * a2;
    for (int i = 0; i < 3; i++) {
      par.add(bsc + ADJ_BODY_FLOATS * jt.b1 + BTAU + i, comp(t, i));
      par.add(bsc + ADJ_BODY_FLOATS * jt.b2 + BTAU + i, -comp(t, i));
    }
  }
  par.sync();
  for (int b = tid; b < nb; b += nt) {
    T *bs = bsc + ADJ_BODY_FLOATS * b;
    const V3<T> v = ld3(D.v + 4 * b), w = ld3(D.w + 4 * b);
    const V3<T> g = v3(P.gravity[0], P.gravity[1], P.gravity[2]);
    st3(bs + BZ, hinv * v + g);
    st3(bs + BZ + 3, hinv * w + ldm(bs + BIIW).mul(ld3(bs + BTAU)));
  }
  par.sync();

  // ---- 1. contact rows, J^T lambda per body
  for (int k = tid; k < nk; k += nt) {
    const TapeContact<T> &c = D.con[k];
    T *J3 = rsc + ADJ_ROW_FLOATS * 3 * k;
    contact_rows(D, c, J3);
    for (int r = 0; r < c.m; r++) {
      const T *J = J3 + ADJ_ROW_FLOATS * r;
      T l = c.lam[r];
      T *b1 = bsc + ADJ_BODY_FLOATS * c.b1 + BJL;
      for (int i = 0; i < 6; i++) par.add(b1 + i, J[i] * l);
      if (c.b2 >= 0) {
        T *b2 = bsc + ADJ_BODY_FLOATS * c.b2 + BJL;
        for (int i = 0; i < 6; i++) par.add(b2 + i, J[6 + i] * l);
      }
    }
  }
  // joint rows (from the pre-step poses), J^T lambda per body
  for (int j = tid; j < nj; j += nt) {
    const TapeJoint<T> &jt = D.jnt[j];
    T *J6 = joint_slots(j);
    T c6[6];
    const V3<T> zero3 = v3(T(0), T(0), T(0));
    const M3<T> R2 = jt.b2 >= 0 ? ldm(bsc + ADJ_BODY_FLOATS * jt.b2 + BR) : M3<T>{zero3, zero3, zero3};
    joint_rows<T, T>(jt, ld3(D.x + 4 * jt.b1), ldm(bsc + ADJ_BODY_FLOATS * jt.b1 + BR),
                     jt.b2 >= 0 ? ld3(D.x + 4 * jt.b2) : zero3, R2, hinv, J6, ADJ_ROW_FLOATS, c6);
    for (int r = 0; r < jt.m; r++) {
      const T *J = J6 + ADJ_ROW_FLOATS * r;
      const T l = jt.lam[r];
      T *b1 = bsc + ADJ_BODY_FLOATS * jt.b1 + BJL;
      for (int i = 0; i < 6; i++) par.add(b1 + i, J[i] * l);
      if (jt.b2 >= 0) {
        T *b2 = bsc + ADJ_BODY_FLOATS * jt.b2 + BJL;
        for (int i = 0; i < 6; i++) par.add(b2 + i, J[6 + i] * l);
      }
    }
  }
  par.sync();

  // ---- 2. integrator and velocity update, backward
  for (int b = tid; b < nb; b += nt) {
    T *bs = bsc + ADJ_BODY_FLOATS * b;
    const T im = D.invMass[b];
    M3<T> iIw = ldm(bs + BIIW);
    V3<T> tau = ld3(bs + BTAU);
    V3<T> X = ld3(bs + BJL + 3) + tau;               // torque through I_w^-1
    V3<T> F = ld3(bs + BJL) + D.mass[b] * v3(P.gravity[0], P.gravity[1], P.gravity[2]);
    V3<T> v1 = ld3(D.v + 4 * b) + (h * im) * F;
    V3<T> w1 = ld3(D.w + 4 * b) + h * iIw.mul(X);
    (void)v1;
    const T *q = D.q + 4 * b;
    T dq[4], qt[4];
    dq_of(w1, q, dq);
    for (int i = 0; i < 4; i++) qt[i] = q[i] + h * dq[i];
    T ql = std::sqrt(qt[0] * qt[0] + qt[1] * qt[1] + qt[2] * qt[2] + qt[3] * qt[3]);
    T qn[4];
    for (int i = 0; i < 4; i++) qn[i] = qt[i] / ql;
    const T *gq1 = G.gq_in + 4 * b;
    T pr = qn[0] * gq1[0] + qn[1] * gq1[1] + qn[2] * gq1[2] + qn[3] * gq1[3];
    T qtb[4], qb[4], a[4];
    for (int i = 0; i < 4; i++) {
      qtb[i] = (gq1[i] - qn[i] * pr) / ql;
      qb[i] = qtb[i];
      a[i] = h * qtb[i];
    }
    V3<T> xb = ld3(G.gx_in + 4 * b);
    V3<T> vb1 = ld3(G.gv_in + 4 * b) + h * xb;
    V3<T> wb1 = ld3(G.gw_in + 4 * b);
    dq_adj(w1, q, a, qb, wb1);
    // v' = v + h/m F, w' = w + h I_w^-1 X
    V3<T> u_lin = (h * im) * vb1;
    V3<T> iwb = iIw.mul(wb1);
    V3<T> u_ang = h * iwb;
    st3(bs + BU, u_lin);
    st3(bs + BU + 3, u_ang);
    // I_w^-1 = R invI R^T depends on the rotation
    V3<T> th = h * (cross(iIw.mul(X), wb1) - cross(X, iwb));
    st3(bs + BTH, th);
    st3(bs + BTAUB, u_ang);     // tau_bar += h I_w^-1 w_bar'
    // outputs so far
    st3(G.gx + 4 * b, xb);
    st3(G.gv + 4 * b, vb1);
    st3(G.gw + 4 * b, wb1);
    for (int i = 0; i < 4; i++) G.gq[4 * b + i] = qb[i];
  }
  par.sync();

  // ---- 3. lambda_bar, J_bar (from J^T lambda), right-hand side of the
  //         adjoint system, preconditioner
  T rr_loc = T(0);
  for (int k = tid; k < nk; k += nt) {
    const TapeContact<T> &c = D.con[k];
    T *J3 = rsc + ADJ_ROW_FLOATS * 3 * k;
    const T *u1 = bsc + ADJ_BODY_FLOATS * c.b1 + BU;
    const T *u2 = c.b2 >= 0 ? bsc + ADJ_BODY_FLOATS * c.b2 + BU : nullptr;
    for (int r = 0; r < 3; r++) {
      T *R = J3 + ADJ_ROW_FLOATS * r;
      if (r >= c.m) {
        for (int i = 0; i < ADJ_ROW_FLOATS; i++) R[i] = T(0);
        continue;
      }
      T lb = jdot6(R, u1) + (u2 ? jdot6(R + 6, u2) : T(0));
      R[RLB] = lb;
      for (int i = 0; i < 6; i++) R[RJB + i] = c.lam[r] * u1[i];
      for (int i = 0; i < 6; i++) R[RJB + 6 + i] = u2 ? c.lam[r] * u2[i] : T(0);
    }
    RowAct<T> A = row_activity(c, P);
    // g = T^T lambda_bar on the free rows
    T g[3] = {T(0), T(0), T(0)};
    if (A.nfree) {
      g[0] = J3[RLB];
      for (int f = 0; f < 2; f++) {
        if (c.m < 3) break;
        if (A.ffree[f]) g[1 + f] = J3[ADJ_ROW_FLOATS * (1 + f) + RLB];
        else g[0] += A.sig[f] * P.mu * J3[ADJ_ROW_FLOATS * (1 + f) + RLB];
      }
    }
    for (int r = 0; r < 3; r++) {
      T *R = J3 + ADJ_ROW_FLOATS * r;
      R[RY] = T(0);
      R[RR] = g[r];
      R[RRH] = g[r];
      R[RP] = T(0);
      R[RV] = T(0);
      rr_loc += g[r] * g[r];
    }
  }
  // joint rows: g = lambda_bar on the free ones (bilateral rows always,
  // bounded limit and motor rows strictly inside their bounds)
  for (int j = tid; j < nj; j += nt) {
    const TapeJoint<T> &jt = D.jnt[j];
    T *J6 = joint_slots(j);
    const T *u1 = bsc + ADJ_BODY_FLOATS * jt.b1 + BU;
    const T *u2 = jt.b2 >= 0 ? bsc + ADJ_BODY_FLOATS * jt.b2 + BU : nullptr;
    for (int r = 0; r < 6; r++) {
      T *R = J6 + ADJ_ROW_FLOATS * r;
      if (r >= jt.m) {
        for (int i = 0; i < ADJ_ROW_FLOATS; i++) R[i] = T(0);
        continue;
      }
      const T lb = jdot6(R, u1) + (u2 ? jdot6(R + 6, u2) : T(0));
      R[RLB] = lb;
      for (int i = 0; i < 6; i++) R[RJB + i] = jt.lam[r] * u1[i];
      for (int i = 0; i < 6; i++) R[RJB + 6 + i] = u2 ? jt.lam[r] * u2[i] : T(0);
      const T g = joint_row_free(jt, r, P.act_tol) ? lb : T(0);
      R[RY] = T(0);
      R[RR] = g;
      R[RRH] = g;
      R[RP] = T(0);
      R[RV] = T(0);
      rr_loc += g * g;
    }
  }
  T gnorm2 = par.sum(rr_loc);

  // K^T applied to the row vector in field `src` (zero outside the free
  // set), result into field `dst`: z = A y (A = J M^-1 J^T + cfm/h), then
  // (K^T y)_j = z_j + sum over bound friction rows f of j: sigma mu z_f.
  // Also returns sum of dst*dot_with and dst*dst for the caller.
  auto apply_KT = [&](int src, int dst) {
    for (int b = tid; b < nb; b += nt)
      for (int i = 0; i < 6; i++) bsc[ADJ_BODY_FLOATS * b + BJY + i] = T(0);
    par.sync();
    for (int s = tid; s < nslots; s += nt) {
      int sb1, sb2;
      T scfm;
      if (!slot_rows(s, sb1, sb2, scfm)) continue;
      const T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
      const T y = R[src];
      if (y == T(0)) continue;
      T *b1 = bsc + ADJ_BODY_FLOATS * sb1 + BJY;
      for (int i = 0; i < 6; i++) par.add(b1 + i, R[i] * y);
      if (sb2 >= 0) {
        T *b2 = bsc + ADJ_BODY_FLOATS * sb2 + BJY;
        for (int i = 0; i < 6; i++) par.add(b2 + i, R[6 + i] * y);
      }
    }
    par.sync();
    for (int k = tid; k < nk; k += nt) {
      const TapeContact<T> &c = D.con[k];
      T *J3 = rsc + ADJ_ROW_FLOATS * 3 * k;
      T m1[6], m2[6];
      minv6(bsc + ADJ_BODY_FLOATS * c.b1, D.invMass[c.b1], bsc + ADJ_BODY_FLOATS * c.b1 + BJY, m1);
      if (c.b2 >= 0) minv6(bsc + ADJ_BODY_FLOATS * c.b2, D.invMass[c.b2], bsc + ADJ_BODY_FLOATS * c.b2 + BJY, m2);
      T z[3] = {T(0), T(0), T(0)};
      for (int r = 0; r < c.m; r++) {
        const T *R = J3 + ADJ_ROW_FLOATS * r;
        z[r] = jdot6(R, m1) + (c.b2 >= 0 ? jdot6(R + 6, m2) : T(0)) + P.cfm * hinv * R[src];
      }
      RowAct<T> A = row_activity(c, P);
      T o[3] = {T(0), T(0), T(0)};
      if (A.nfree) {
        o[0] = z[0];
        for (int f = 0; f < 2; f++) {
          if (c.m < 3) break;
          if (A.ffree[f]) o[1 + f] = z[1 + f];
          else o[0] += A.sig[f] * P.mu * z[1 + f];
        }
      }
      for (int r = 0; r < 3; r++) J3[ADJ_ROW_FLOATS * r + dst] = o[r];
    }
    for (int j = tid; j < nj; j += nt) {
      const TapeJoint<T> &jt = D.jnt[j];
      T *J6 = joint_slots(j);
      T m1[6], m2[6];
      minv6(bsc + ADJ_BODY_FLOATS * jt.b1, D.invMass[jt.b1], bsc + ADJ_BODY_FLOATS * jt.b1 + BJY, m1);
      if (jt.b2 >= 0) minv6(bsc + ADJ_BODY_FLOATS * jt.b2, D.invMass[jt.b2], bsc + ADJ_BODY_FLOATS * jt.b2 + BJY, m2);
      for (int r = 0; r < 6; r++) {
        T *R = J6 + ADJ_ROW_FLOATS * r;
        R[dst] = (r < jt.m && joint_row_free(jt, r, P.act_tol))
                     ? jdot6(R, m1) + (jt.b2 >= 0 ? jdot6(R + 6, m2) : T(0)) + jt.rcfm[r] * R[src]
                     : T(0);
      }
    }
    par.sync();
  };
  // Jacobi preconditioner: inverse diagonal of A on the row of slot s (0
  // beyond the constraint's rows)
  auto precond = [&](int s, T x) -> T {
    int sb1, sb2;
    T scfm;
    if (!slot_rows(s, sb1, sb2, scfm)) return T(0);
    const T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
    T m1[6], d = scfm;
    minv6(bsc + ADJ_BODY_FLOATS * sb1, D.invMass[sb1], R, m1);
    d += jdot6(R, m1);
    if (sb2 >= 0) {
      T m2[6];
      minv6(bsc + ADJ_BODY_FLOATS * sb2, D.invMass[sb2], R + 6, m2);
      d += jdot6(R + 6, m2);
    }
    return d > T(0) ? x / d : x;
  };

  // ---- 4. BiCGSTAB on K^T y = g (right Jacobi preconditioning); the
  //         preconditioned directions live in RT (y_hat) and RV/RS fields
  int iters = 0;
  if (gnorm2 > T(0)) {
    T rho0 = T(1), alpha = T(1), omega = T(1);
    const T stop2 = P.lin_tol * P.lin_tol * gnorm2;
    // safeguard: the iterate with the smallest residual is kept (RYB, from
    // y = 0 with residual |g|^2); a non-finite or exploding residual ends the
    // iteration and the kept iterate is used
    T best2 = gnorm2;
    for (int s = tid; s < nslots; s += nt) rsc[ADJ_ROW_FLOATS * (long long)s + RYB] = T(0);
    bool use_best = false;
    // fields: RY x, RR r, RRH r_hat, RP p, RV v, RS s, RT t; a hat vector is
    // kept in the J_bar-free slots RJB..: no -- use the dedicated scratch
    // below (wsc is scalars only), so the hat vectors reuse RS / RT in turn.
    for (iters = 0; iters < P.lin_maxit; iters++) {
      T loc = T(0);
      for (int s = tid; s < nslots; s += nt) {
        const T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
        loc += R[RRH] * R[RR];
      }
      T rho1 = par.sum(loc);
      if (rho1 == T(0)) break;
      T beta = (rho1 / rho0) * (alpha / omega);
      // p = r + beta (p - omega v); s := M^-1 p (hat)
      for (int s = tid; s < nslots; s += nt) {
        T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
        R[RP] = R[RR] + beta * (R[RP] - omega * R[RV]);
        R[RS] = precond(s, R[RP]);
      }
      par.sync();
      apply_KT(RS, RV);   // v = K^T M^-1 p
      loc = T(0);
      for (int s = tid; s < nslots; s += nt) {
        const T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
        loc += R[RRH] * R[RV];
      }
      T den = par.sum(loc);
      if (den == T(0)) break;
      alpha = rho1 / den;
      // x += alpha p_hat; s = r - alpha v
      loc = T(0);
      for (int s = tid; s < nslots; s += nt) {
        T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
        R[RY] += alpha * R[RS];
        R[RS] = R[RR] - alpha * R[RV];
        loc += R[RS] * R[RS];
      }
      T ss = par.sum(loc);
      if (ss <= stop2) {
        for (int s = tid; s < nslots; s += nt) rsc[ADJ_ROW_FLOATS * (long long)s + RR] = rsc[ADJ_ROW_FLOATS * (long long)s + RS];
        iters++;
        break;
      }
      // z_hat = M^-1 s into RP (p is rebuilt from r next iteration... keep p:
      // store z_hat in RT temporarily, t = K^T z_hat into RJB+11 is not free,
      // so t overwrites RT after z_hat has been copied to the x update below)
      for (int s = tid; s < nslots; s += nt) {
        T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
        R[RT] = precond(s, R[RS]);
      }
      par.sync();
      // t = K^T z_hat: computed into RJB+... not available; use RR as the
      // destination (r is recomputed from s and t below)
      apply_KT(RT, RR);   // RR := t
      T ts = T(0), tt = T(0);
      for (int s = tid; s < nslots; s += nt) {
        const T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
        ts += R[RR] * R[RS];
        tt += R[RR] * R[RR];
      }
      ts = par.sum(ts);
      tt = par.sum(tt);
      omega = tt > T(0) ? ts / tt : T(0);
      loc = T(0);
      for (int s = tid; s < nslots; s += nt) {
        T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
        R[RY] += omega * R[RT];
        R[RR] = R[RS] - omega * R[RR];   // r = s - omega t
        loc += R[RR] * R[RR];
      }
      T rr = par.sum(loc);
      rho0 = rho1;
      if (!(rr <= T(1e12) * gnorm2)) { use_best = true; iters++; break; }   // diverged or not finite
      if (rr < best2) {
        best2 = rr;
        for (int s = tid; s < nslots; s += nt) {
          T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
          R[RYB] = R[RY];
        }
      }
      if (rr <= stop2 || omega == T(0)) { iters++; break; }
      if (iters + 1 == P.lin_maxit && rr > best2) use_best = true;   // ended worse than the best iterate
    }
    if (use_best) {
      par.sync();
      for (int s = tid; s < nslots; s += nt) {
        T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
        R[RY] = R[RYB];
      }
    }
  }
  if (tid == 0 && G.stats) G.stats[0] += iters;
  par.sync();

  // ---- 5. with y: W_y = M^-1 J^T y, z = A y; b_bar = y; J_bar, theta_bar,
  //         mu_bar; then the rhs b = c/h - J Z and c backward
  apply_KT(RY, RS);   // leaves J^T y in BJY; RS unused afterwards
  for (int b = tid; b < nb; b += nt) {
    T *bs = bsc + ADJ_BODY_FLOATS * b;
    M3<T> iIw = ldm(bs + BIIW);
    V3<T> ay = ld3(bs + BJY + 3), al = ld3(bs + BJL + 3);
    V3<T> th = ld3(bs + BTH) - cross(iIw.mul(al), ay) + cross(al, iIw.mul(ay));
    st3(bs + BTH, th);
  }
  par.sync();
  T mu_loc = T(0), bounce_loc = T(0);
  for (int k = tid; k < nk; k += nt) {
    const TapeContact<T> &c = D.con[k];
    T *J3 = rsc + ADJ_ROW_FLOATS * 3 * k;
    T wl1[6], wl2[6], wy1[6], wy2[6];
    const T *b1s = bsc + ADJ_BODY_FLOATS * c.b1;
    minv6(b1s, D.invMass[c.b1], b1s + BJL, wl1);
    minv6(b1s, D.invMass[c.b1], b1s + BJY, wy1);
    const T *b2s = c.b2 >= 0 ? bsc + ADJ_BODY_FLOATS * c.b2 : nullptr;
    if (b2s) {
      minv6(b2s, D.invMass[c.b2], b2s + BJL, wl2);
      minv6(b2s, D.invMass[c.b2], b2s + BJY, wy2);
    }
    RowAct<T> A = row_activity(c, P);
    // z = A y on the bound friction rows (for mu_bar)
    for (int f = 0; f < 2; f++) {
      if (c.m < 3 || A.sig[f] == T(0)) continue;
      const T *R = J3 + ADJ_ROW_FLOATS * (1 + f);
      T z = jdot6(R, wy1) + (b2s ? jdot6(R + 6, wy2) : T(0));   // y_f = 0 on a bound row
      mu_loc += A.sig[f] * c.lam[0] * (R[RLB] - z);
    }
    // b_bar (free rows), J_bar from A and from b = c/h - J Z
    const T *Z1 = b1s + BZ;
    const T *Z2 = b2s ? b2s + BZ : nullptr;
    for (int r = 0; r < c.m; r++) {
      T *R = J3 + ADJ_ROW_FLOATS * r;
      T y = R[RY];
      T l = c.lam[r];
      for (int i = 0; i < 6; i++) R[RJB + i] += -(y * wl1[i] + l * wy1[i]) - y * Z1[i];
      if (b2s)
        for (int i = 0; i < 6; i++) R[RJB + 6 + i] += -(y * wl2[i] + l * wy2[i]) - y * Z2[i];
    }
  }
  for (int j = tid; j < nj; j += nt) {
    const TapeJoint<T> &jt = D.jnt[j];
    T *J6 = joint_slots(j);
    T wl1[6], wl2[6], wy1[6], wy2[6];
    const T *b1s = bsc + ADJ_BODY_FLOATS * jt.b1;
    minv6(b1s, D.invMass[jt.b1], b1s + BJL, wl1);
    minv6(b1s, D.invMass[jt.b1], b1s + BJY, wy1);
    const T *b2s = jt.b2 >= 0 ? bsc + ADJ_BODY_FLOATS * jt.b2 : nullptr;
    if (b2s) {
      minv6(b2s, D.invMass[jt.b2], b2s + BJL, wl2);
      minv6(b2s, D.invMass[jt.b2], b2s + BJY, wy2);
    }
    const T *Z1 = b1s + BZ;
    const T *Z2 = b2s ? b2s + BZ : nullptr;
    for (int r = 0; r < jt.m; r++) {
      T *R = J6 + ADJ_ROW_FLOATS * r;
      const T y = R[RY], l = jt.lam[r];
      for (int i = 0; i < 6; i++) R[RJB + i] += -(y * wl1[i] + l * wy1[i]) - y * Z1[i];
      if (b2s)
        for (int i = 0; i < 6; i++) R[RJB + 6 + i] += -(y * wl2[i] + l * wy2[i]) - y * Z2[i];
    }
  }
  par.sync();
  // Z_bar = -J^T b_bar per body (into BJY, reused), c_bar = b_bar / h
  for (int b = tid; b < nb; b += nt)
    for (int i = 0; i < 6; i++) bsc[ADJ_BODY_FLOATS * b + BJY + i] = T(0);
  par.sync();
  for (int s = tid; s < nslots; s += nt) {
    int sb1, sb2;
    T scfm;
    if (!slot_rows(s, sb1, sb2, scfm)) continue;
    const T *R = rsc + ADJ_ROW_FLOATS * (long long)s;
    const T y = R[RY];
    if (y == T(0)) continue;
    T *b1 = bsc + ADJ_BODY_FLOATS * sb1 + BJY;
    for (int i = 0; i < 6; i++) par.add(b1 + i, -R[i] * y);
    if (sb2 >= 0) {
      T *b2 = bsc + ADJ_BODY_FLOATS * sb2 + BJY;
      for (int i = 0; i < 6; i++) par.add(b2 + i, -R[6 + i] * y);
    }
  }
  par.sync();

  // ---- 6. c backward, J_bar -> geometry adjoints per contact (stored in
  //         the row scratch of row 0: RR.. reused as p_bar 3, n_bar 3, d_bar)
  for (int k = tid; k < nk; k += nt) {
    const TapeContact<T> &c = D.con[k];
    T *J3 = rsc + ADJ_ROW_FLOATS * 3 * k;
    V3<T> n = ld3(c.normal);
    const bool rev = (c.flags & JF_REVERSE) != 0;
    if (rev) n = -n;
    V3<T> p = ld3(c.pos);
    V3<T> x1 = ld3(D.x + 4 * c.b1);
    V3<T> x2 = c.b2 >= 0 ? ld3(D.x + 4 * c.b2) : v3(T(0), T(0), T(0));
    V3<T> r1 = p - x1, r2 = c.b2 >= 0 ? p - x2 : v3(T(0), T(0), T(0));
    T cb = J3[RY] * hinv;            // c_bar of the normal row
    T depth_bar = T(0);
    {
      // c = erp/h * max(depth - min_depth, 0), capped by max_vel; bounce
      T dd = c.depth - P.min_depth;
      T cerp = P.erp * hinv * (dd > T(0) ? dd : T(0));
      bool capped = false;
      if (cerp > P.max_vel) { cerp = P.max_vel; capped = true; }
      bool use_bounce = false;
      T out = T(0);
      if (P.surf_bounce) {
        const T *J = J3;
        out = J[0] * D.v[4 * c.b1] + J[1] * D.v[4 * c.b1 + 1] + J[2] * D.v[4 * c.b1 + 2] + J[3] * D.w[4 * c.b1] +
              J[4] * D.w[4 * c.b1 + 1] + J[5] * D.w[4 * c.b1 + 2];
        if (c.b2 >= 0)
          out += J[6] * D.v[4 * c.b2] + J[7] * D.v[4 * c.b2 + 1] + J[8] * D.v[4 * c.b2 + 2] + J[9] * D.w[4 * c.b2] +
                 J[10] * D.w[4 * c.b2 + 1] + J[11] * D.w[4 * c.b2 + 2];
        if (P.bounce_vel >= T(0) && -out > P.bounce_vel) {
          T nc = -P.bounce * out;
          if (nc > cerp) use_bounce = true;
        }
      }
      if (use_bounce) {
        bounce_loc += -out * cb;
        T ob = -P.bounce * cb;     // out_bar
        T *JB = J3 + RJB;
        const T *J = J3;
        for (int i = 0; i < 3; i++) {
          JB[i] += ob * D.v[4 * c.b1 + i];
          JB[3 + i] += ob * D.w[4 * c.b1 + i];
        }
        // velocity adjoints of the pre-step state
        for (int i = 0; i < 3; i++) {
          par.add(G.gv + 4 * c.b1 + i, ob * J[i]);
          par.add(G.gw + 4 * c.b1 + i, ob * J[3 + i]);
        }
        if (c.b2 >= 0) {
          for (int i = 0; i < 3; i++) {
            JB[6 + i] += ob * D.v[4 * c.b2 + i];
            JB[9 + i] += ob * D.w[4 * c.b2 + i];
            par.add(G.gv + 4 * c.b2 + i, ob * J[6 + i]);
            par.add(G.gw + 4 * c.b2 + i, ob * J[9 + i]);
          }
        }
      } else if (!capped && dd > T(0)) {
        depth_bar = P.erp * hinv * cb;
      }
    }
    // J_bar -> direction, lever arm adjoints
    V3<T> dbar[3], r1b = v3(T(0), T(0), T(0)), r2b = v3(T(0), T(0), T(0));
    V3<T> d[3];
    d[0] = n;
    if (c.m >= 3) plane_space(n, d[1], d[2]);
    for (int r = 0; r < c.m; r++) {
      const T *JB = J3 + ADJ_ROW_FLOATS * r + RJB;
      V3<T> l1 = ld3(JB), a1 = ld3(JB + 3), l2 = ld3(JB + 6), a2 = ld3(JB + 9);
      dbar[r] = l1 + cross(a1, r1);
      r1b = r1b + cross(d[r], a1);
      if (c.b2 >= 0) {
        dbar[r] = dbar[r] - l2 - cross(a2, r2);
        r2b = r2b - cross(d[r], a2);
      }
    }
    V3<T> nb_ = c.m >= 1 ? dbar[0] : v3(T(0), T(0), T(0));
    if (c.m >= 3) {
      // t1, t2 = dPlaneSpace(n): forward-mode over the three components of n
      Dual<T, 3> nd[3];
      nd[0] = Dual<T, 3>(n.x); nd[0].d[0] = T(1);
      nd[1] = Dual<T, 3>(n.y); nd[1].d[1] = T(1);
      nd[2] = Dual<T, 3>(n.z); nd[2].d[2] = T(1);
      V3<Dual<T, 3>> t1, t2;
      plane_space(v3(nd[0], nd[1], nd[2]), t1, t2);
      for (int j = 0; j < 3; j++) {
        T s = dbar[1].x * t1.x.d[j] + dbar[1].y * t1.y.d[j] + dbar[1].z * t1.z.d[j] + dbar[2].x * t2.x.d[j] +
              dbar[2].y * t2.y.d[j] + dbar[2].z * t2.z.d[j];
        if (j == 0) nb_.x += s;
        if (j == 1) nb_.y += s;
        if (j == 2) nb_.z += s;
      }
    }
    // r1 = p - x1, r2 = p - x2
    V3<T> pbar = r1b + r2b;
    for (int i = 0; i < 3; i++) par.add(G.gx + 4 * c.b1 + i, -comp(r1b, i));
    if (c.b2 >= 0)
      for (int i = 0; i < 3; i++) par.add(G.gx + 4 * c.b2 + i, -comp(r2b, i));
    if (rev) nb_ = -nb_;   // back to the collider's normal
    T *GA = J3 + ADJ_ROW_FLOATS * 2;   // row 2's J_bar slots are no longer needed
    st3(GA + RJB, pbar);
    st3(GA + RJB + 3, nb_);
    GA[RJB + 6] = depth_bar;
  }
  // joints: J and c = erp/h (...) depend on both poses; c_bar = b_bar / h.
  // Position and rotation adjoints of each body by dual numbers over the
  // joint's rows.
  for (int j = tid; j < nj; j += nt) {
    const TapeJoint<T> &jt = D.jnt[j];
    const T *J6 = joint_slots(j);
    const bool two = jt.b2 >= 0;
    const V3<T> zero3 = v3(T(0), T(0), T(0));
    const V3<T> x1 = ld3(D.x + 4 * jt.b1), x2 = two ? ld3(D.x + 4 * jt.b2) : zero3;
    const M3<T> R1 = ldm(bsc + ADJ_BODY_FLOATS * jt.b1 + BR);
    const M3<T> R2 = two ? ldm(bsc + ADJ_BODY_FLOATS * jt.b2 + BR) : M3<T>{zero3, zero3, zero3};
    for (int pass = 0; pass < (two ? 4 : 2); pass++) {
      const int side = pass >> 1, kind = pass & 1;   // body 1 / 2; position / rotation
      typedef Dual<T, 3> S;
      V3<S> px1, px2;
      M3<S> pR1, pR2;
      pose_dual(x1, R1, side == 0 ? kind : -1, px1, pR1);
      pose_dual(x2, R2, side == 1 ? kind : -1, px2, pR2);
      S Jd[6 * 12], cd[6];
      joint_rows<T, S>(jt, px1, pR1, px2, pR2, hinv, Jd, 12, cd);
      T acc[3] = {T(0), T(0), T(0)};
      for (int r = 0; r < jt.m; r++) {
        const T *R = J6 + ADJ_ROW_FLOATS * r;
        const T cbar = R[RY] * hinv;
        for (int dir = 0; dir < 3; dir++) {
          T a = cbar * cd[r].d[dir];
          for (int i = 0; i < 12; i++) a += R[RJB + i] * Jd[12 * r + i].d[dir];
          acc[dir] += a;
        }
      }
      const int body = side == 0 ? jt.b1 : jt.b2;
      T *dst = kind == 0 ? G.gx + 4 * body : bsc + ADJ_BODY_FLOATS * body + BTH;
      for (int dir = 0; dir < 3; dir++) par.add(dst + dir, acc[dir]);
    }
  }
  par.sync();

  // ---- 7. bodies: Z_bar, torque and gyroscopic backward
  for (int b = tid; b < nb; b += nt) {
    T *bs = bsc + ADJ_BODY_FLOATS * b;
    M3<T> iIw = ldm(bs + BIIW), Iw = ldm(bs + BIW);
    V3<T> zl = ld3(bs + BJY), za = ld3(bs + BJY + 3);
    V3<T> tau = ld3(bs + BTAU);
    V3<T> w = ld3(D.w + 4 * b);
    for (int i = 0; i < 3; i++) {
      par.add(G.gv + 4 * b + i, hinv * comp(zl, i));
      par.add(G.gw + 4 * b + i, hinv * comp(za, i));
    }
    V3<T> iza = iIw.mul(za);
    V3<T> taub = ld3(bs + BTAUB) + iza;
    V3<T> th = ld3(bs + BTH) + cross(iIw.mul(tau), za) - cross(tau, iza);
    // tau = -w x (I_w w)
    V3<T> L = Iw.mul(w);
    V3<T> a = cross(taub, w);
    V3<T> wb = -cross(L, taub) - Iw.mul(a);
    th = th - cross(L, a) + cross(w, Iw.mul(a));
    for (int i = 0; i < 3; i++) par.add(G.gw + 4 * b + i, comp(wb, i));
    st3(bs + BTH, th);
    st3(bs + BTAUB, taub);
  }
  par.sync();
  // hinge2 torques: tau_1 = t, tau_2 = -t with t = t1 R1 axis1 + t2 R2 axis2;
  // d(R a) = dtheta x (R a) gives theta_bar += t_k (R a_k) x (tau_bar_1 - tau_bar_2)
  for (int j = tid; j < nj; j += nt) {
    const TapeJoint<T> &jt = D.jnt[j];
    if (jt.type != JT_HINGE2 || jt.b2 < 0 || (jt.torque[0] == T(0) && jt.torque[1] == T(0))) continue;
    const T *b1s = bsc + ADJ_BODY_FLOATS * jt.b1, *b2s = bsc + ADJ_BODY_FLOATS * jt.b2;
    const V3<T> tb = ld3(b1s + BTAUB) - ld3(b2s + BTAUB);
    const V3<T> g1 = jt.torque[0] * cross(ldm(b1s + BR).mul(ld3(jt.ax1)), tb);
    const V3<T> g2 = jt.torque[1] * cross(ldm(b2s + BR).mul(ld3(jt.ax2)), tb);
    for (int i = 0; i < 3; i++) {
      par.add(bsc + ADJ_BODY_FLOATS * jt.b1 + BTH + i, comp(g1, i));
      par.add(bsc + ADJ_BODY_FLOATS * jt.b2 + BTH + i, comp(g2, i));
    }
  }
  par.sync();

  // ---- 8. collision geometry: dual-number colliders per pair (the
  //         contacts of one pair are consecutive in the tape)
  int mism = 0;
  for (int k = tid; k < nk; k += nt) {
    const TapeContact<T> &c = D.con[k];
    if (k > 0 && D.con[k - 1].g1 == c.g1 && D.con[k - 1].g2 == c.g2) continue;   // not the pair's first
    int cnt = 1;
    while (k + cnt < nk && D.con[k + cnt].g1 == c.g1 && D.con[k + cnt].g2 == c.g2) cnt++;
    const int gg[2] = {c.g1, c.g2};
    for (int side = 0; side < 2; side++) {
      int body = D.gbody[gg[side]];
      if (body < 0) continue;
      for (int kind = 0; kind < 2; kind++) {   // 0: position, 1: rotation
        typedef Dual<T, 3> S;
        GeomD<S> G2[2];
        for (int s2 = 0; s2 < 2; s2++) {
          int g = gg[s2];
          GeomD<S> &o = G2[s2];
          o.type = D.gtype[g];
          for (int i = 0; i < 4; i++) o.prm[i] = (float)D.gprm[4 * g + i];
          int bb = D.gbody[g];
          V3<T> pp;
          M3<T> RR;
          if (bb >= 0) {
            pp = ld3(D.x + 4 * bb);
            RR = quat_rot(D.q + 4 * bb);
          } else {
            pp = ld3(D.gpos + 4 * g);
            RR = ldm4(D.gR + 12 * g);
          }
          o.p = v3(S(pp.x), S(pp.y), S(pp.z));
          o.R.r0 = v3(S(RR.r0.x), S(RR.r0.y), S(RR.r0.z));
          o.R.r1 = v3(S(RR.r1.x), S(RR.r1.y), S(RR.r1.z));
          o.R.r2 = v3(S(RR.r2.x), S(RR.r2.y), S(RR.r2.z));
          if (s2 == side) {
            if (kind == 0) {
              o.p.x.d[0] = T(1); o.p.y.d[1] = T(1); o.p.z.d[2] = T(1);
            } else {
              // d R / d theta_j = [e_j]x R: column c gets e_j x R_c
              for (int cc = 0; cc < 3; cc++) {
                V3<T> col = RR.col(cc);
                // e_x x col = (0, -col.z, col.y); e_y x col = (col.z, 0, -col.x); e_z x col = (-col.y, col.x, 0)
                S *e0 = cc == 0 ? &o.R.r0.x : (cc == 1 ? &o.R.r0.y : &o.R.r0.z);
                S *e1 = cc == 0 ? &o.R.r1.x : (cc == 1 ? &o.R.r1.y : &o.R.r1.z);
                S *e2 = cc == 0 ? &o.R.r2.x : (cc == 1 ? &o.R.r2.y : &o.R.r2.z);
                e0->d[0] = T(0);    e1->d[0] = -col.z; e2->d[0] = col.y;
                e0->d[1] = col.z;   e1->d[1] = T(0);   e2->d[1] = -col.x;
                e0->d[2] = -col.y;  e1->d[2] = col.x;  e2->d[2] = T(0);
              }
            }
          }
        }
        CP<S> out[MAX_SLOTS];
        int n = d_collide(G2[0], G2[1], P.maxc > MAX_SLOTS ? MAX_SLOTS : P.maxc, out);
        if (n != cnt) {
          if (side == 0 && kind == 0) mism++;
          continue;
        }
        T acc[3] = {T(0), T(0), T(0)};
        for (int j = 0; j < n; j++) {
          const T *GA = rsc + ADJ_ROW_FLOATS * (3 * (k + j) + 2) + RJB;
          for (int dir = 0; dir < 3; dir++)
            acc[dir] += GA[0] * out[j].pos.x.d[dir] + GA[1] * out[j].pos.y.d[dir] + GA[2] * out[j].pos.z.d[dir] +
                        GA[3] * out[j].n.x.d[dir] + GA[4] * out[j].n.y.d[dir] + GA[5] * out[j].n.z.d[dir] +
                        GA[6] * out[j].depth.d[dir];
        }
        T *dst = kind == 0 ? G.gx + 4 * body : bsc + ADJ_BODY_FLOATS * body + BTH;
        for (int dir = 0; dir < 3; dir++) par.add(dst + dir, acc[dir]);
      }
    }
  }
  T mu_sum = par.sum(mu_loc);
  T bounce_sum = par.sum(bounce_loc);
  T mism_sum = par.sum(T(mism));
  if (tid == 0) {
    *G.gmu += mu_sum;
    *G.gbounce += bounce_sum;
    if (G.stats) G.stats[1] += (int)mism_sum;
  }
  par.sync();

  // ---- 9. rotation adjoint -> quaternion: q_bar += 2 (0, theta_bar) (x) q
  for (int b = tid; b < nb; b += nt) {
    const T *bs = bsc + ADJ_BODY_FLOATS * b;
    V3<T> th = ld3(bs + BTH);
    const T *q = D.q + 4 * b;
    T *gq = G.gq + 4 * b;
    // (0, a) (x) q = (-a.q_v, q0 a + a x q_v)
    V3<T> qv = v3(q[1], q[2], q[3]);
    V3<T> cv = cross(th, qv);
    gq[0] += T(2) * (-dot(th, qv));
    gq[1] += T(2) * (q[0] * th.x + cv.x);
    gq[2] += T(2) * (q[0] * th.y + cv.y);
    gq[3] += T(2) * (q[0] * th.z + cv.z);
  }
  par.sync();
}

}  // namespace adj
}  // namespace odemi
