// host_common.hpp -- host-side helpers shared by the batch API and the ODE
// API layer: error reporting, device buffers, and the setup-time math of the
// reference's setters (mass inversion, joint anchors and axes).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "layout.hpp"

namespace odemi {

void report_error(const char *fmt, ...);

#define OD_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess) {                                                          \
      ::odemi::report_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                     \
      return 0;                                                                      \
    }                                                                                \
  } while (0)

// a device allocation owned by a host object
struct DevBuf {
  void *p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  bool alloc(size_t bytes) {
    if (bytes <= n && p) return true;
    release();
    if (bytes == 0) bytes = 16;
    if (hipMalloc(&p, bytes) != hipSuccess) { p = nullptr; return false; }
    n = bytes;
    return true;
  }
  template <class T> T *as() const { return (T *)p; }
};

// ---------------------------------------------------------------- setup math
// (host copies of the reference's setters; float, 3x4 row-major rotations)

inline void h_quat_to_rot(const float *q, float *R) {  // rotation.cpp:236-255
  float xx = 2 * q[1] * q[1], yy = 2 * q[2] * q[2], zz = 2 * q[3] * q[3];
  R[0] = 1 - yy - zz; R[1] = 2 * (q[1] * q[2] - q[0] * q[3]); R[2] = 2 * (q[1] * q[3] + q[0] * q[2]); R[3] = 0;
  R[4] = 2 * (q[1] * q[2] + q[0] * q[3]); R[5] = 1 - xx - zz; R[6] = 2 * (q[2] * q[3] - q[0] * q[1]); R[7] = 0;
  R[8] = 2 * (q[1] * q[3] - q[0] * q[2]); R[9] = 2 * (q[2] * q[3] + q[0] * q[1]); R[10] = 1 - xx - yy; R[11] = 0;
}
inline void h_normalize4(float *q) {  // odemath.cpp:119-137
  float l = q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3];
  if (l > 0) {
    l = 1.f / std::sqrt(l);
    for (int k = 0; k < 4; k++) q[k] *= l;
  } else {
    q[0] = 1; q[1] = q[2] = q[3] = 0;
  }
}
inline int h_normalize3(float *a) {  // odemath.cpp:42-91
  float ax = std::fabs(a[0]), ay = std::fabs(a[1]), az = std::fabs(a[2]);
  float m = std::fmax(ax, std::fmax(ay, az));
  if (!(m > 0)) { a[0] = 1; a[1] = 0; a[2] = 0; return 0; }
  a[0] /= m; a[1] /= m; a[2] /= m;
  float l = 1.f / std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  a[0] *= l; a[1] *= l; a[2] *= l;
  return 1;
}
inline void h_mv(float *r, const float *M, const float *v) {
  float x = M[0] * v[0] + M[1] * v[1] + M[2] * v[2];
  float y = M[4] * v[0] + M[5] * v[1] + M[6] * v[2];
  float z = M[8] * v[0] + M[9] * v[1] + M[10] * v[2];
  r[0] = x; r[1] = y; r[2] = z;
}
inline void h_mtv(float *r, const float *M, const float *v) {
  float x = M[0] * v[0] + M[4] * v[1] + M[8] * v[2];
  float y = M[1] * v[0] + M[5] * v[1] + M[9] * v[2];
  float z = M[2] * v[0] + M[6] * v[1] + M[10] * v[2];
  r[0] = x; r[1] = y; r[2] = z;
}
inline void h_cross(float *r, const float *a, const float *b) {
  float x = a[1] * b[2] - a[2] * b[1], y = a[2] * b[0] - a[0] * b[2], z = a[0] * b[1] - a[1] * b[0];
  r[0] = x; r[1] = y; r[2] = z;
}
inline float h_dot(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void h_qmul1(float *a, const float *b, const float *c) {
  a[0] = b[0] * c[0] + b[1] * c[1] + b[2] * c[2] + b[3] * c[3];
  a[1] = b[0] * c[1] - b[1] * c[0] - b[2] * c[3] + b[3] * c[2];
  a[2] = b[0] * c[2] - b[2] * c[0] - b[3] * c[1] + b[1] * c[3];
  a[3] = b[0] * c[3] - b[3] * c[0] - b[1] * c[2] + b[2] * c[1];
}

// inverse of a symmetric positive definite 3x3 by Cholesky (matrix.cpp:124-222)
inline int h_invert_pd3(const float *A, float *Ainv) {
  float L[12];
  std::memcpy(L, A, sizeof(L));
  float rcp[3];
  for (int i = 0; i < 3; i++) {
    float *row = L + 4 * i;
    for (int j = 0; j < i; j++) {
      float s = row[j];
      for (int k = 0; k < j; k++) s -= row[k] * L[4 * j + k];
      row[j] = s * rcp[j];
    }
    float s = row[i];
    for (int k = 0; k < i; k++) s -= row[k] * row[k];
    if (s <= 0) return 0;
    float r = std::sqrt(s);
    row[i] = r;
    rcp[i] = 1.f / r;
  }
  std::memset(Ainv, 0, 12 * sizeof(float));
  for (int c = 0; c < 3; c++) {
    float b[3] = {0, 0, 0}, y[3];
    b[c] = 1;
    for (int i = 0; i < 3; i++) {
      float s = 0;
      for (int k = 0; k < i; k++) s += L[4 * i + k] * y[k];
      y[i] = (b[i] - s) / L[4 * i + i];
    }
    for (int i = 2; i >= 0; i--) {
      float s = 0;
      for (int k = i + 1; k < 3; k++) s += L[4 * k + i] * b[k];
      b[i] = (y[i] - s) / L[4 * i + i];
    }
    for (int i = 0; i < 3; i++) Ainv[4 * i + c] = b[i];
  }
  return 1;
}

// pose of a body as seen by joint setup
struct HPose {
  float pos[4], q[4], R[12];
};

// joint.cpp:260-288
inline void h_set_anchors(JointDev &j, const HPose *b1, const HPose *b2, float x, float y, float z) {
  if (b1) {
    float q[4] = {x - b1->pos[0], y - b1->pos[1], z - b1->pos[2], 0};
    h_mtv(j.anchor1, b1->R, q);
    if (b2) {
      float q2[4] = {x - b2->pos[0], y - b2->pos[1], z - b2->pos[2], 0};
      h_mtv(j.anchor2, b2->R, q2);
    } else {
      j.anchor2[0] = x; j.anchor2[1] = y; j.anchor2[2] = z;
    }
  }
  j.anchor1[3] = 0;
  j.anchor2[3] = 0;
}
// joint.cpp:293-324
inline void h_set_axes(JointDev &j, const HPose *b1, const HPose *b2, float x, float y, float z, bool a1, bool a2) {
  if (!b1) return;
  float q[4] = {x, y, z, 0};
  h_normalize3(q);
  if (a1) { h_mtv(j.axis1, b1->R, q); j.axis1[3] = 0; }
  if (a2) {
    if (b2) h_mtv(j.axis2, b2->R, q);
    else { j.axis2[0] = x; j.axis2[1] = y; j.axis2[2] = z; }
    j.axis2[3] = 0;
  }
}
// hinge.cpp:377-395
inline void h_hinge_qrel(JointDev &j, const HPose *b1, const HPose *b2) {
  if (!b1) return;
  if (b2) h_qmul1(j.qrel, b1->q, b2->q);
  else { j.qrel[0] = b1->q[0]; j.qrel[1] = -b1->q[1]; j.qrel[2] = -b1->q[2]; j.qrel[3] = -b1->q[3]; }
}
// hinge2.cpp:107-115
inline void h_hinge2_axis_info(const JointDev &j, const HPose *b1, const HPose *b2, float &sn, float &cs) {
  float ax1[3], ax2[3], c[3];
  h_mv(ax1, b1->R, j.axis1);
  h_mv(ax2, b2->R, j.axis2);
  h_cross(c, ax1, ax2);
  sn = std::sqrt(h_dot(c, c));
  cs = h_dot(ax1, ax2);
}
// hinge2.cpp:186-211
inline void h_hinge2_v1v2(JointDev &j, const HPose *b1, const HPose *b2) {
  if (!b1 || !b2) return;
  float ax1[4], ax2[4], v[3];
  h_mv(ax1, b1->R, j.axis1);
  h_mv(ax2, b2->R, j.axis2);
  if ((ax1[0] == 0 && ax1[1] == 0 && ax1[2] == 0) || (ax2[0] == 0 && ax2[1] == 0 && ax2[2] == 0) ||
      (ax1[0] == ax2[0] && ax1[1] == ax2[1] && ax1[2] == ax2[2]))
    return;
  float k = h_dot(ax1, ax2);
  for (int i = 0; i < 3; i++) ax2[i] -= k * ax1[i];
  ax2[3] = 0;
  h_normalize3(ax2);
  h_cross(v, ax1, ax2);
  h_mtv(j.v1, b1->R, ax2);
  h_mtv(j.v2, b1->R, v);
}

inline void h_limot_init(Limot &l, float erp, float cfm) {  // joint.cpp:456-469
  l.vel = 0; l.fmax = 0; l.lostop = -INFINITY; l.histop = INFINITY; l.fudge = 1;
  l.normal_cfm = cfm; l.stop_erp = erp; l.stop_cfm = cfm; l.bounce = 0; l.limit = 0; l.limit_err = 0; l.pad = 0;
}
// dxJointLimitMotor::set (joint.cpp:472-509); param within the group
inline void h_limot_set(Limot &l, int p, float v) {
  switch (p) {
    case 0: l.lostop = v; break;
    case 1: l.histop = v; break;
    case 2: l.vel = v; break;
    case 3: if (v >= 0) l.fmax = v; break;
    case 4: if (v >= 0 && v <= 1) l.fudge = v; break;
    case 5: l.bounce = v; break;
    case 6: l.normal_cfm = v; break;
    case 7: l.stop_erp = v; break;
    case 8: l.stop_cfm = v; break;
  }
}
inline float h_limot_get(const Limot &l, int p) {
  switch (p) {
    case 0: return l.lostop; case 1: return l.histop; case 2: return l.vel; case 3: return l.fmax;
    case 4: return l.fudge; case 5: return l.bounce; case 6: return l.normal_cfm; case 7: return l.stop_erp;
    case 8: return l.stop_cfm;
  }
  return 0;
}

inline void h_joint_defaults(JointDev &j, int type, float erp, float cfm) {
  std::memset(&j, 0, sizeof(j));
  j.type = type;
  j.b1 = -1;
  j.b2 = -1;
  j.axis1[0] = 1;
  if (type == JT_HINGE2) j.axis2[1] = 1; else j.axis2[0] = 1;
  j.v1[0] = 1;
  j.v2[1] = 1;
  j.erp = erp;
  j.cfm = cfm;
  j.susp_erp = erp;
  j.susp_cfm = cfm;
  j.fb = -1;
  h_limot_init(j.lm1, erp, cfm);
  h_limot_init(j.lm2, erp, cfm);
}

// capacities of a world's scratch buffers
inline void world_capacities(int nb, int ng, int nj, double scale, int &pcap, int &kcap, int &rcap) {
  long long p = (long long)(scale * (16.0 * ng + 64));
  pcap = (int)od_pow2(p);
  kcap = (int)(scale * (6.0 * ng + 64));
  rcap = 3 * kcap + 6 * nj + 8;
}

// kernel launch entry points (step_kernels.hip)
struct GeomView;
struct CPoint;
// world kernel modes (step_kernels.hip)
enum {
  KM_BATCH_COLLIDE = 1,   // AABB + broadphase + narrowphase of every world
  KM_BATCH_STEP = 2,      // quickstep on the contacts of KM_BATCH_COLLIDE
  KM_FUSED = 3,           // both in one kernel
  KM_API_COLLIDE = 5,     // one world, pairs and contacts copied out (dSpaceCollide)
  KM_API_STEP = 10,       // one world, contact joints from the host (dWorldQuickStep)
  KM_BATCH_STEP_TOL = 18, // KM_BATCH_STEP honouring the solver tolerance
  KM_FUSED_TOL = 19,      // KM_FUSED honouring the solver tolerance
  KM_FUSED_PRE = 51,      // KM_FUSED_TOL up to the coloured solver's set-up
  KM_POST = 82,           // the rest of the step after the sweeps
  KM_SOR = 1000           // the coloured solver's sweeps (between KM_FUSED_PRE and KM_POST)
};
hipError_t launch_world_kernel(int mode, const DevBatch &B, int first_world, int nworlds_launch, int grid, int wg,
                               int api_maxc, hipStream_t st, int queue = 0);
hipError_t launch_init_rot(const DevBatch &B, int nbodies, hipStream_t st);
// the persistent queue's order of B.world_order by expected cost (rank: the
// static order; est: the per-world cost estimates it keeps); at most 8192
// worlds
hipError_t launch_order_worlds(const DevBatch &B, const int *rank, float *est, hipStream_t st);
int world_kernel_static_lds();
hipError_t allow_world_kernel_lds(int bytes);
int world_kernel_occupancy(int mode, int wg, int lds_bytes);

}  // namespace odemi
