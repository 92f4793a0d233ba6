// joints_dev.hpp -- constraint row construction on the device: contact rows
// and the ball / hinge / hinge2 joints (getInfo1 / getInfo2 semantics of the
// reference, cited per function).  Rows are written unscaled; the solver
// applies Ad = w / (J inv(M) J^T + cfm) itself.
#pragma once
#include "devmath.hpp"
#include "layout.hpp"

namespace odemi {

struct BodyPose {
  float3 p;
  Mat3 R;
  float4 q;
  float3 v, w;
};

__device__ inline BodyPose load_pose(const DevBatch &B, int gb) {
  BodyPose s;
  s.p = xyz(B.pos[gb]);
  s.R = load_mat(B.R + 3 * (size_t)gb);
  s.q = B.q[gb];
  s.v = xyz(B.lvel[gb]);
  s.w = xyz(B.avel[gb]);
  return s;
}

// a row builder working on 12-float rows with stride 36 floats (Row layout)
struct RowOut {
  Row *r;   // first row of the constraint
  int m;
  __device__ float *J(int i) { return r[i].J; }
  __device__ void set_l1(int i, float3 v) { r[i].J[0] = v.x; r[i].J[1] = v.y; r[i].J[2] = v.z; }
  __device__ void set_a1(int i, float3 v) { r[i].J[3] = v.x; r[i].J[4] = v.y; r[i].J[5] = v.z; }
  __device__ void set_l2(int i, float3 v) { r[i].J[6] = v.x; r[i].J[7] = v.y; r[i].J[8] = v.z; }
  __device__ void set_a2(int i, float3 v) { r[i].J[9] = v.x; r[i].J[10] = v.y; r[i].J[11] = v.z; }
};

__device__ inline void row_clear(Row &r, float cfm) {
  for (int k = 0; k < 12; k++) r.J[k] = 0.f;
  r.rhs = 0.f;  // holds c until the rhs pass
  r.cfm = cfm;
  r.lo = -__int_as_float(0x7f800000);
  r.hi = __int_as_float(0x7f800000);
  r.findex = -1;
  r.lambda = 0.f;
}

// hinge angle from the relative rotation (joint.cpp:381-450)
__device__ inline float hinge_angle(float4 q1, const float4 *q2, float4 qrel, float3 axis) {
  float4 qr;
  if (q2) qr = qmul2(qmul1(q1, *q2), qrel);
  else qr = qmul3(q1, qrel);
  float cost2 = qr.x;
  float sint2 = sqrtf(qr.y * qr.y + qr.z * qr.z + qr.w * qr.w);
  float3 v = mk3(qr.y, qr.z, qr.w);
  float th = (dot(v, axis) >= 0.f) ? 2.f * atan2f(sint2, cost2) : 2.f * atan2f(sint2, -cost2);
  if (th > 3.14159265358979323846f) th -= 2.f * 3.14159265358979323846f;
  return -th;
}

// limit test (joint.cpp:538-553)
__device__ inline int limot_test(Limot &lm, float angle) {
  if (angle <= lm.lostop) { lm.limit = 1; lm.limit_err = angle - lm.lostop; return 1; }
  if (angle >= lm.histop) { lm.limit = 2; lm.limit_err = angle - lm.histop; return 1; }
  lm.limit = 0;
  return 0;
}

// powered / limited rotational axis row (joint.cpp:556-733).  A motor that
// pushes against its limit adds a torque to the bodies; it is returned in
// `extra` (body 1 gets -extra, body 2 +extra) so that it is applied in a
// deterministic order.
__device__ inline int limot_row(Limot &lm, Row &r, float3 ax1, bool has2, const BodyPose &s1, const BodyPose *s2,
                                float fps, float3 &extra) {
  int powered = lm.fmax > 0.f;
  if (!(powered || lm.limit)) return 0;
  r.J[3] = ax1.x; r.J[4] = ax1.y; r.J[5] = ax1.z;
  if (has2) { r.J[9] = -ax1.x; r.J[10] = -ax1.y; r.J[11] = -ax1.z; }
  if (lm.limit && lm.lostop == lm.histop) powered = 0;
  if (powered) {
    r.cfm = lm.normal_cfm;
    if (!lm.limit) {
      r.rhs = lm.vel;
      r.lo = -lm.fmax;
      r.hi = lm.fmax;
    } else {
      float fm = lm.fmax;
      if (lm.vel > 0.f || (lm.vel == 0.f && lm.limit == 2)) fm = -fm;
      if ((lm.limit == 1 && lm.vel > 0.f) || (lm.limit == 2 && lm.vel < 0.f)) fm *= lm.fudge;
      extra = extra + fm * ax1;
    }
  }
  if (lm.limit) {
    float k = fps * lm.stop_erp;
    r.rhs = -k * lm.limit_err;
    r.cfm = lm.stop_cfm;
    const float inf = __int_as_float(0x7f800000);
    if (lm.lostop == lm.histop) {
      r.lo = -inf;
      r.hi = inf;
    } else {
      if (lm.limit == 1) { r.lo = 0.f; r.hi = inf; }
      else { r.lo = -inf; r.hi = 0.f; }
      if (lm.bounce > 0.f) {
        float v = dot(s1.w, ax1);
        if (has2) v -= dot(s2->w, ax1);
        if (lm.limit == 1) {
          if (v < 0.f) { float nc = -lm.bounce * v; if (nc > r.rhs) r.rhs = nc; }
        } else {
          if (v > 0.f) { float nc = -lm.bounce * v; if (nc < r.rhs) r.rhs = nc; }
        }
      }
    }
  }
  return 1;
}

// hinge2 axis data (hinge2.cpp:107-115)
__device__ inline void hinge2_axes(const JointDev &j, const BodyPose &s1, const BodyPose &s2, float3 &ax1, float3 &ax2,
                                   float3 &axc, float &sn, float &cs) {
  ax1 = s1.R.mul(mk3(j.axis1[0], j.axis1[1], j.axis1[2]));
  ax2 = s2.R.mul(mk3(j.axis2[0], j.axis2[1], j.axis2[2]));
  axc = cross(ax1, ax2);
  sn = sqrtf(dot(axc, axc));
  cs = dot(ax1, ax2);
}

// hinge2 steering angle (hinge2.cpp:34-43)
__device__ inline float hinge2_angle(const JointDev &j, const BodyPose &s1, const BodyPose &s2) {
  float3 a1 = s2.R.mul(mk3(j.axis2[0], j.axis2[1], j.axis2[2]));
  float3 a2 = s1.R.tmul(a1);
  float x = dot(mk3(j.v1[0], j.v1[1], j.v1[2]), a2);
  float y = dot(mk3(j.v2[0], j.v2[1], j.v2[2]), a2);
  return -atan2f(y, x);
}

// getInfo1 (ball.cpp:48-52, hinge.cpp:53-74, hinge2.cpp:80-100); updates
// the limit state stored in the joint
__device__ inline int joint_info1(JointDev &j, const BodyPose &s1, const BodyPose *s2) {
  const float PI = 3.14159265358979323846f;
  if (j.type == JT_BALL) return 3;
  if (j.type == JT_HINGE) {
    int m = (j.lm1.fmax > 0.f) ? 6 : 5;
    if ((j.lm1.lostop >= -PI || j.lm1.histop <= PI) && j.lm1.lostop <= j.lm1.histop) {
      float ang = hinge_angle(s1.q, s2 ? &s2->q : nullptr, make_float4(j.qrel[0], j.qrel[1], j.qrel[2], j.qrel[3]),
                              mk3(j.axis1[0], j.axis1[1], j.axis1[2]));
      if (limot_test(j.lm1, ang)) m = 6;
    }
    return m;
  }
  if (j.type == JT_HINGE2) {
    int m = 4;
    j.lm1.limit = 0;
    if ((j.lm1.lostop >= -PI || j.lm1.histop <= PI) && j.lm1.lostop <= j.lm1.histop)
      limot_test(j.lm1, hinge2_angle(j, s1, *s2));
    if (j.lm1.limit || j.lm1.fmax > 0.f) m++;
    j.lm2.limit = 0;
    if (j.lm2.fmax > 0.f) m++;
    return m;
  }
  return 0;
}

// three ball-and-socket rows (joint.cpp:85-126); erp is the Info2 erp
__device__ inline void rows_ball(Row *r, float3 anc1, float3 anc2, const BodyPose &s1, const BodyPose *s2, float fps,
                                 float erp) {
  float3 a1 = s1.R.mul(anc1);
  r[0].J[0] = 1.f; r[1].J[1] = 1.f; r[2].J[2] = 1.f;
  // -[a1]x
  r[0].J[4] = a1.z;  r[0].J[5] = -a1.y;
  r[1].J[3] = -a1.z; r[1].J[5] = a1.x;
  r[2].J[3] = a1.y;  r[2].J[4] = -a1.x;
  float k = fps * erp;
  if (s2) {
    float3 a2 = s2->R.mul(anc2);
    r[0].J[6] = -1.f; r[1].J[7] = -1.f; r[2].J[8] = -1.f;
    r[0].J[10] = -a2.z; r[0].J[11] = a2.y;
    r[1].J[9] = a2.z;   r[1].J[11] = -a2.x;
    r[2].J[9] = -a2.y;  r[2].J[10] = a2.x;
    float3 c = ((a2 + s2->p) - a1) - s1.p;
    r[0].rhs = k * c.x; r[1].rhs = k * c.y; r[2].rhs = k * c.z;
  } else {
    float3 c = (anc2 - a1) - s1.p;
    r[0].rhs = k * c.x; r[1].rhs = k * c.y; r[2].rhs = k * c.z;
  }
}

// ball rows aligned to a unit axis (joint.cpp:134-196)
__device__ inline void rows_ball2(Row *r, float3 anc1, float3 anc2, float3 axis, float erp1, const BodyPose &s1,
                                  const BodyPose *s2, float fps, float erp) {
  float3 q1, q2;
  plane_space(axis, q1, q2);
  float3 dirs[3] = {axis, q1, q2};
  float3 a1 = s1.R.mul(anc1);
  float3 a2 = mk3(0.f, 0.f, 0.f);
  if (s2) a2 = s2->R.mul(anc2);
  for (int i = 0; i < 3; i++) {
    float3 d = dirs[i];
    r[i].J[0] = d.x; r[i].J[1] = d.y; r[i].J[2] = d.z;
    float3 c1 = cross(a1, d);
    r[i].J[3] = c1.x; r[i].J[4] = c1.y; r[i].J[5] = c1.z;
    if (s2) {
      r[i].J[6] = -d.x; r[i].J[7] = -d.y; r[i].J[8] = -d.z;
      float3 c2 = cross(a2, d);
      r[i].J[9] = -c2.x; r[i].J[10] = -c2.y; r[i].J[11] = -c2.z;
    }
  }
  float k1 = fps * erp1, k = fps * erp;
  float3 w1 = a1 + s1.p;
  float3 d = s2 ? (a2 + s2->p) - w1 : anc2 - w1;
  r[0].rhs = k1 * dot(axis, d);
  r[1].rhs = k * dot(q1, d);
  r[2].rhs = k * dot(q2, d);
}

// getInfo2 of a permanent joint into its m rows (ball.cpp:56-64,
// hinge.cpp:76-149, hinge2.cpp:127-180).  `erp` is the running Info2 erp,
// which a ball joint overwrites for the joints processed after it.
__device__ inline void joint_info2(JointDev &j, Row *r, const BodyPose &s1, const BodyPose *s2, float fps, float &erp,
                                   float3 &extra1, float3 &extra2) {
  float3 anc1 = mk3(j.anchor1[0], j.anchor1[1], j.anchor1[2]);
  float3 anc2 = mk3(j.anchor2[0], j.anchor2[1], j.anchor2[2]);
  if (j.type == JT_BALL) {
    erp = j.erp;
    r[0].cfm = r[1].cfm = r[2].cfm = j.cfm;
    rows_ball(r, anc1, anc2, s1, s2, fps, erp);
  } else if (j.type == JT_HINGE) {
    rows_ball(r, anc1, anc2, s1, s2, fps, erp);
    float3 ax1 = s1.R.mul(mk3(j.axis1[0], j.axis1[1], j.axis1[2]));
    float3 p, q;
    plane_space(ax1, p, q);
    r[3].J[3] = p.x; r[3].J[4] = p.y; r[3].J[5] = p.z;
    r[4].J[3] = q.x; r[4].J[4] = q.y; r[4].J[5] = q.z;
    if (s2) {
      r[3].J[9] = -p.x; r[3].J[10] = -p.y; r[3].J[11] = -p.z;
      r[4].J[9] = -q.x; r[4].J[10] = -q.y; r[4].J[11] = -q.z;
    }
    float3 ax2 = s2 ? s2->R.mul(mk3(j.axis2[0], j.axis2[1], j.axis2[2])) : mk3(j.axis2[0], j.axis2[1], j.axis2[2]);
    float3 b = cross(ax1, ax2);
    float k = fps * erp;
    r[3].rhs = k * dot(b, p);
    r[4].rhs = k * dot(b, q);
    float3 ex = mk3(0.f, 0.f, 0.f);
    limot_row(j.lm1, r[5], ax1, s2 != nullptr, s1, s2, fps, ex);
    extra1 = extra1 - ex;
    extra2 = extra2 + ex;
  } else if (j.type == JT_HINGE2) {
    float3 ax1, ax2, qv;
    float sn, cs;
    hinge2_axes(j, s1, *s2, ax1, ax2, qv, sn, cs);
    qv = safe_normalize(qv);
    rows_ball2(r, anc1, anc2, ax1, j.susp_erp, s1, s2, fps, erp);
    r[3].J[3] = qv.x; r[3].J[4] = qv.y; r[3].J[5] = qv.z;
    r[3].J[9] = -qv.x; r[3].J[10] = -qv.y; r[3].J[11] = -qv.z;
    float k = fps * erp;
    r[3].rhs = k * (j.c0 * sn - j.s0 * cs);
    float3 ex = mk3(0.f, 0.f, 0.f);
    int row = 4 + limot_row(j.lm1, r[4], ax1, true, s1, s2, fps, ex);
    limot_row(j.lm2, r[row], ax2, true, s1, s2, fps, ex);
    extra1 = extra1 - ex;
    extra2 = extra2 + ex;
    r[0].cfm = j.susp_cfm;
  }
}

// contact rows (joints/contact.cpp:47-71 getInfo1, 75-256 getInfo2)
struct Surf {
  int mode;
  float mu, mu2, bounce, bounce_vel, soft_erp, soft_cfm, motion1, motion2, motionN, slip1, slip2;
  float3 fdir1;
};

__device__ inline int contact_info1(Surf &s) {
  int m = 1;
  if (s.mu < 0.f) s.mu = 0.f;
  if (s.mode & SM_MU2) {
    if (s.mu > 0.f || s.mu2 > 0.f) m++;
    if (s.mu2 < 0.f) s.mu2 = 0.f;
    if (s.mu2 > 0.f) m++;
  } else if (s.mu > 0.f) {
    m += 2;
  }
  return m;
}

__device__ inline void contact_info2(Row *r, int m, const Surf &s, float3 cpos, float3 normal_in, float depth_in,
                                     bool reverse, const BodyPose &s1, const BodyPose *s2, float fps, float erp,
                                     float min_depth, float max_vel) {
  float3 n = reverse ? -normal_in : normal_in;
  float3 c1 = cpos - s1.p;
  float3 c2 = mk3(0.f, 0.f, 0.f);
  r[0].J[0] = n.x; r[0].J[1] = n.y; r[0].J[2] = n.z;
  float3 a = cross(c1, n);
  r[0].J[3] = a.x; r[0].J[4] = a.y; r[0].J[5] = a.z;
  if (s2) {
    c2 = cpos - s2->p;
    r[0].J[6] = -n.x; r[0].J[7] = -n.y; r[0].J[8] = -n.z;
    float3 b = -cross(c2, n);
    r[0].J[9] = b.x; r[0].J[10] = b.y; r[0].J[11] = b.z;
  }
  float e = (s.mode & SM_SOFT_ERP) ? s.soft_erp : erp;
  float k = fps * e;
  float depth = depth_in - min_depth;
  if (depth < 0.f) depth = 0.f;
  if (s.mode & SM_SOFT_CFM) r[0].cfm = s.soft_cfm;
  float motionN = (s.mode & SM_MOTIONN) ? s.motionN : 0.f;
  float c = k * depth + motionN;
  if (c > max_vel) c = max_vel;
  if (s.mode & SM_BOUNCE) {
    float out = dot(n, s1.v) + dot(a, s1.w);
    if (s2) out += dot(-n, s2->v) + dot(mk3(r[0].J[9], r[0].J[10], r[0].J[11]), s2->w);
    out -= motionN;
    if (s.bounce_vel >= 0.f && (-out) > s.bounce_vel) {
      float nc = -s.bounce * out + motionN;
      if (nc > c) c = nc;
    }
  }
  r[0].rhs = c;
  r[0].lo = 0.f;
  r[0].hi = __int_as_float(0x7f800000);
  float3 t1, t2;
  if (m >= 2) {
    if (s.mode & SM_FDIR1) {
      t1 = s.fdir1;
      t2 = cross(n, t1);
    } else {
      plane_space(n, t1, t2);
    }
    r[1].J[0] = t1.x; r[1].J[1] = t1.y; r[1].J[2] = t1.z;
    float3 a1 = cross(c1, t1);
    r[1].J[3] = a1.x; r[1].J[4] = a1.y; r[1].J[5] = a1.z;
    if (s2) {
      r[1].J[6] = -t1.x; r[1].J[7] = -t1.y; r[1].J[8] = -t1.z;
      float3 b1 = -cross(c2, t1);
      r[1].J[9] = b1.x; r[1].J[10] = b1.y; r[1].J[11] = b1.z;
    }
    if (s.mode & SM_MOTION1) r[1].rhs = s.motion1;
    r[1].lo = -s.mu;
    r[1].hi = s.mu;
    if (s.mode & SM_APPROX1_1) r[1].findex = 0;
    if (s.mode & SM_SLIP1) r[1].cfm = s.slip1;
  }
  if (m >= 3) {
    r[2].J[0] = t2.x; r[2].J[1] = t2.y; r[2].J[2] = t2.z;
    float3 a2 = cross(c1, t2);
    r[2].J[3] = a2.x; r[2].J[4] = a2.y; r[2].J[5] = a2.z;
    if (s2) {
      r[2].J[6] = -t2.x; r[2].J[7] = -t2.y; r[2].J[8] = -t2.z;
      float3 b2 = -cross(c2, t2);
      r[2].J[9] = b2.x; r[2].J[10] = b2.y; r[2].J[11] = b2.z;
    }
    if (s.mode & SM_MOTION2) r[2].rhs = s.motion2;
    if (s.mode & SM_MU2) { r[2].lo = -s.mu2; r[2].hi = s.mu2; }
    else { r[2].lo = -s.mu; r[2].hi = s.mu; }
    if (s.mode & SM_APPROX1_2) r[2].findex = 0;
    if (s.mode & SM_SLIP2) r[2].cfm = s.slip2;
  }
}

}  // namespace odemi
