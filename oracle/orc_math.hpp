// orc_math.hpp -- small fixed-size linear algebra used by the CPU restatement.
//
// TEST INFRASTRUCTURE (oracle).  Only tests/, the smoke check and the CPU
// baseline leg of bench.py may load code built from this directory.
//
// Conventions follow the reference's storage layout: a 3-vector occupies 4
// reals (include/ode/common.h:139-141), a rotation is 3 rows of 4 reals,
// row-major (common.h:142).  Sums are evaluated left to right in the order
// the reference's inline helpers use (include/ode/odemath.h:175-300), so the
// double build of this restatement tracks the reference double build to a
// few ulps.
#pragma once
#include <cmath>
#include <cstring>

namespace orc {

template <class R> inline R rsqrt(R x) { return R(1) / std::sqrt(x); }
template <class R> inline R recip(R x) { return R(1) / x; }

// dot of a (stride sa) and b (stride sb): odemath.h:175-178
template <class R> inline R dot3(const R *a, const R *b, int sa = 1, int sb = 1) {
  return a[0] * b[0] + a[sa] * b[sb] + a[2 * sa] * b[2 * sb];
}
// cross product: odemath.h:196-206
template <class R> inline void cross3(R *r, const R *a, const R *b, int sr = 1, int sa = 1, int sb = 1) {
  R x = a[sa] * b[2 * sb] - a[2 * sa] * b[sb];
  R y = a[2 * sa] * b[0] - a[0] * b[2 * sb];
  R z = a[0] * b[sb] - a[sa] * b[0];
  r[0] = x;
  r[sr] = y;
  r[2 * sr] = z;
}
// r = M v  (odemath.h:280-288, dMultiply0_331)
template <class R> inline void mv(R *r, const R *M, const R *v) {
  R x = dot3(M, v), y = dot3(M + 4, v), z = dot3(M + 8, v);
  r[0] = x; r[1] = y; r[2] = z;
}
// r = M^T v  (dMultiply1_331)
template <class R> inline void mtv(R *r, const R *M, const R *v) {
  R x = dot3(M, v, 4, 1), y = dot3(M + 1, v, 4, 1), z = dot3(M + 2, v, 4, 1);
  r[0] = x; r[1] = y; r[2] = z;
}
// r = A B  (dMultiply0_333)
template <class R> inline void mm(R *r, const R *A, const R *B) {
  for (int i = 0; i < 3; i++) {
    R t0 = dot3(B + 0, A + 4 * i, 4, 1), t1 = dot3(B + 1, A + 4 * i, 4, 1), t2 = dot3(B + 2, A + 4 * i, 4, 1);
    r[4 * i] = t0; r[4 * i + 1] = t1; r[4 * i + 2] = t2;
  }
}
// r = A B^T  (dMultiply2_333)
template <class R> inline void mmt(R *r, const R *A, const R *B) {
  for (int i = 0; i < 3; i++) {
    R t0 = dot3(B, A + 4 * i), t1 = dot3(B + 4, A + 4 * i), t2 = dot3(B + 8, A + 4 * i);
    r[4 * i] = t0; r[4 * i + 1] = t1; r[4 * i + 2] = t2;
  }
}
template <class R> inline void zero(R *a, int n) { for (int i = 0; i < n; i++) a[i] = 0; }
template <class R> inline void identity3(R *M) {
  zero(M, 12);
  M[0] = M[5] = M[10] = 1;
}

// dSafeNormalize3 (ode/src/odemath.cpp:42-91): scale by the largest
// magnitude first, then by 1/length.
template <class R> inline int normalize3(R *a) {
  R ax = std::fabs(a[0]), ay = std::fabs(a[1]), az = std::fabs(a[2]);
  int k;
  if (ay > ax) k = (az > ay) ? 2 : 1;
  else if (az > ax) k = 2;
  else {
    if (ax <= 0) { a[0] = 1; a[1] = 0; a[2] = 0; return 0; }
    k = 0;
  }
  R m = (k == 0) ? ax : (k == 1 ? ay : az);
  a[0] /= m; a[1] /= m; a[2] /= m;
  R l = rsqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  a[0] *= l; a[1] *= l; a[2] *= l;
  return 1;
}
// dSafeNormalize4 (odemath.cpp:119-137)
template <class R> inline void normalize4(R *a) {
  R l = dot3(a, a) + a[3] * a[3];
  if (l > 0) {
    l = rsqrt(l);
    a[0] *= l; a[1] *= l; a[2] *= l; a[3] *= l;
  } else {
    a[0] = 1; a[1] = 0; a[2] = 0; a[3] = 0;
  }
}
// dPlaneSpace (odemath.cpp:151-178): two unit vectors p, q spanning the
// plane normal to unit n, with q = n x p.
template <class R> inline void plane_space(const R *n, R *p, R *q) {
  const R half_sqrt2 = R(0.7071067811865475244008443621048490);
  if (std::fabs(n[2]) > half_sqrt2) {
    R a = n[1] * n[1] + n[2] * n[2];
    R k = rsqrt(a);
    p[0] = 0; p[1] = -n[2] * k; p[2] = n[1] * k;
    q[0] = a * k; q[1] = -n[0] * p[2]; q[2] = n[0] * p[1];
  } else {
    R a = n[0] * n[0] + n[1] * n[1];
    R k = rsqrt(a);
    p[0] = -n[1] * k; p[1] = n[0] * k; p[2] = 0;
    q[0] = -n[2] * p[1]; q[1] = n[2] * p[0]; q[2] = a * k;
  }
}
// rotation from unit quaternion (w,x,y,z): ode/src/rotation.cpp:236-255
template <class R> inline void quat_to_rot(R *M, const R *q) {
  R xx = 2 * q[1] * q[1], yy = 2 * q[2] * q[2], zz = 2 * q[3] * q[3];
  M[0] = 1 - yy - zz;            M[1] = 2 * (q[1] * q[2] - q[0] * q[3]); M[2] = 2 * (q[1] * q[3] + q[0] * q[2]); M[3] = 0;
  M[4] = 2 * (q[1] * q[2] + q[0] * q[3]); M[5] = 1 - xx - zz;            M[6] = 2 * (q[2] * q[3] - q[0] * q[1]); M[7] = 0;
  M[8] = 2 * (q[1] * q[3] - q[0] * q[2]); M[9] = 2 * (q[2] * q[3] + q[0] * q[1]); M[10] = 1 - xx - yy;           M[11] = 0;
}
// quaternion derivative for angular velocity w: rotation.cpp:310-317
template <class R> inline void quat_rate(R *dq, const R *w, const R *q) {
  dq[0] = R(0.5) * (-w[0] * q[1] - w[1] * q[2] - w[2] * q[3]);
  dq[1] = R(0.5) * (w[0] * q[0] + w[1] * q[3] - w[2] * q[2]);
  dq[2] = R(0.5) * (-w[0] * q[3] + w[1] * q[0] + w[2] * q[1]);
  dq[3] = R(0.5) * (w[0] * q[2] - w[1] * q[1] + w[2] * q[0]);
}
// quaternion products: rotation.cpp:191-233 (0: b*c, 1: conj(b)*c,
// 2: b*conj(c), 3: conj(b)*conj(c))
template <class R> inline void qmul0(R *a, const R *b, const R *c) {
  a[0] = b[0] * c[0] - b[1] * c[1] - b[2] * c[2] - b[3] * c[3];
  a[1] = b[0] * c[1] + b[1] * c[0] + b[2] * c[3] - b[3] * c[2];
  a[2] = b[0] * c[2] + b[2] * c[0] + b[3] * c[1] - b[1] * c[3];
  a[3] = b[0] * c[3] + b[3] * c[0] + b[1] * c[2] - b[2] * c[1];
}
template <class R> inline void qmul1(R *a, const R *b, const R *c) {
  a[0] = b[0] * c[0] + b[1] * c[1] + b[2] * c[2] + b[3] * c[3];
  a[1] = b[0] * c[1] - b[1] * c[0] - b[2] * c[3] + b[3] * c[2];
  a[2] = b[0] * c[2] - b[2] * c[0] - b[3] * c[1] + b[1] * c[3];
  a[3] = b[0] * c[3] - b[3] * c[0] - b[1] * c[2] + b[2] * c[1];
}
template <class R> inline void qmul2(R *a, const R *b, const R *c) {
  a[0] = b[0] * c[0] + b[1] * c[1] + b[2] * c[2] + b[3] * c[3];
  a[1] = -b[0] * c[1] + b[1] * c[0] - b[2] * c[3] + b[3] * c[2];
  a[2] = -b[0] * c[2] + b[2] * c[0] - b[3] * c[1] + b[1] * c[3];
  a[3] = -b[0] * c[3] + b[3] * c[0] - b[1] * c[2] + b[2] * c[1];
}
template <class R> inline void qmul3(R *a, const R *b, const R *c) {
  a[0] = b[0] * c[0] - b[1] * c[1] - b[2] * c[2] - b[3] * c[3];
  a[1] = -b[0] * c[1] - b[1] * c[0] + b[2] * c[3] - b[3] * c[2];
  a[2] = -b[0] * c[2] - b[2] * c[0] + b[3] * c[1] - b[1] * c[3];
  a[3] = -b[0] * c[3] - b[3] * c[0] + b[1] * c[2] - b[2] * c[1];
}

// Inverse of a symmetric positive definite 3x3 (padded rows of 4) through a
// Cholesky factorisation, as dInvertPDMatrix does (ode/src/matrix.cpp:124-222).
// Returns 0 if the matrix is not positive definite.
template <class R> inline int invert_pd3(const R *A, R *Ainv) {
  R L[12];
  std::memcpy(L, A, sizeof(L));
  R rcp[3];
  for (int i = 0; i < 3; i++) {
    R *row = L + 4 * i;
    for (int j = 0; j < i; j++) {
      R s = row[j];
      for (int k = 0; k < j; k++) s -= row[k] * L[4 * j + k];
      row[j] = s * rcp[j];
    }
    R s = row[i];
    for (int k = 0; k < i; k++) s -= row[k] * row[k];
    if (s <= 0) return 0;
    R r = std::sqrt(s);
    row[i] = r;
    rcp[i] = recip(r);
  }
  zero(Ainv, 12);
  for (int c = 0; c < 3; c++) {
    R b[3] = {0, 0, 0}, y[3];
    b[c] = 1;
    for (int i = 0; i < 3; i++) {
      R s = 0;
      for (int k = 0; k < i; k++) s += L[4 * i + k] * y[k];
      y[i] = (b[i] - s) / L[4 * i + i];
    }
    for (int i = 2; i >= 0; i--) {
      R s = 0;
      for (int k = i + 1; k < 3; k++) s += L[4 * k + i] * b[k];
      b[i] = (y[i] - s) / L[4 * i + i];
    }
    for (int i = 0; i < 3; i++) Ainv[4 * i + c] = b[i];
  }
  return 1;
}

}  // namespace orc
