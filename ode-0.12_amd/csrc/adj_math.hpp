// adj_math.hpp -- scalar-generic small vector algebra and forward-mode dual
// numbers for the adjoint of the step.  The same code compiles for the GPU
// (hipcc, float) and for a host build in double (the adjoint's own CPU
// checks); OD_HD marks functions usable on both.
#pragma once
#include <cmath>
#include <type_traits>

#include "layout.hpp"

namespace odemi {
namespace adj {

// ------------------------------------------------------------------ dual numbers

// value + N directional derivatives
template <class T, int N>
struct Dual {
  T v;
  T d[N];
  OD_HD Dual() : v(0) {
    for (int i = 0; i < N; i++) d[i] = 0;
  }
  OD_HD Dual(T x) : v(x) {
    for (int i = 0; i < N; i++) d[i] = 0;
  }
};

template <class T> OD_HD inline T val(T x) { return x; }
template <class T, int N> OD_HD inline T val(const Dual<T, N> &x) { return x.v; }

#define OD_DUAL template <class T, int N> OD_HD inline
OD_DUAL Dual<T, N> operator+(const Dual<T, N> &a, const Dual<T, N> &b) {
  Dual<T, N> r;
  r.v = a.v + b.v;
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] + b.d[i];
  return r;
}
OD_DUAL Dual<T, N> operator-(const Dual<T, N> &a, const Dual<T, N> &b) {
  Dual<T, N> r;
  r.v = a.v - b.v;
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] - b.d[i];
  return r;
}
OD_DUAL Dual<T, N> operator-(const Dual<T, N> &a) {
  Dual<T, N> r;
  r.v = -a.v;
  for (int i = 0; i < N; i++) r.d[i] = -a.d[i];
  return r;
}
OD_DUAL Dual<T, N> operator*(const Dual<T, N> &a, const Dual<T, N> &b) {
  Dual<T, N> r;
  r.v = a.v * b.v;
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
  return r;
}
OD_DUAL Dual<T, N> operator/(const Dual<T, N> &a, const Dual<T, N> &b) {
  Dual<T, N> r;
  r.v = a.v / b.v;
  T ib = T(1) / b.v;
  for (int i = 0; i < N; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
  return r;
}
// mixed with plain arithmetic scalars (float constants in double builds too)
#define OD_MIXED template <class T, int N, class U, class = typename std::enable_if<std::is_arithmetic<U>::value>::type> OD_HD inline
OD_MIXED Dual<T, N> operator+(const Dual<T, N> &a, U b) { return a + Dual<T, N>(T(b)); }
OD_MIXED Dual<T, N> operator+(U a, const Dual<T, N> &b) { return Dual<T, N>(T(a)) + b; }
OD_MIXED Dual<T, N> operator-(const Dual<T, N> &a, U b) { return a - Dual<T, N>(T(b)); }
OD_MIXED Dual<T, N> operator-(U a, const Dual<T, N> &b) { return Dual<T, N>(T(a)) - b; }
OD_MIXED Dual<T, N> operator*(const Dual<T, N> &a, U b) {
  Dual<T, N> r;
  r.v = a.v * T(b);
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * T(b);
  return r;
}
OD_MIXED Dual<T, N> operator*(U a, const Dual<T, N> &b) { return b * a; }
OD_MIXED Dual<T, N> operator/(const Dual<T, N> &a, U b) { return a / Dual<T, N>(T(b)); }
OD_MIXED Dual<T, N> operator/(U a, const Dual<T, N> &b) { return Dual<T, N>(T(a)) / b; }
#undef OD_MIXED
OD_DUAL Dual<T, N> sqrt_(const Dual<T, N> &a) {
  Dual<T, N> r;
  r.v = std::sqrt(a.v);
  T k = r.v > T(0) ? T(0.5) / r.v : T(0);
  for (int i = 0; i < N; i++) r.d[i] = a.d[i] * k;
  return r;
}
OD_DUAL Dual<T, N> fabs_(const Dual<T, N> &a) { return a.v < T(0) ? -a : a; }
OD_DUAL Dual<T, N> atan2_(const Dual<T, N> &y, const Dual<T, N> &x) {
  Dual<T, N> r;
  r.v = std::atan2(y.v, x.v);
  const T den = x.v * x.v + y.v * y.v;
  const T k = den > T(0) ? T(1) / den : T(0);
  for (int i = 0; i < N; i++) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) * k;
  return r;
}
#undef OD_DUAL
template <class T> OD_HD inline T sqrt_(T a) { return std::sqrt(a); }
template <class T> OD_HD inline T fabs_(T a) { return std::fabs(a); }
template <class T> OD_HD inline T atan2_(T y, T x) { return std::atan2(y, x); }

// ------------------------------------------------------------------ vectors

template <class S>
struct V3 {
  S x, y, z;
};
template <class S> OD_HD inline V3<S> v3(S x, S y, S z) { return V3<S>{x, y, z}; }
template <class S> OD_HD inline V3<S> operator+(const V3<S> &a, const V3<S> &b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class S> OD_HD inline V3<S> operator-(const V3<S> &a, const V3<S> &b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class S> OD_HD inline V3<S> operator-(const V3<S> &a) { return v3(-a.x, -a.y, -a.z); }
template <class S, class K> OD_HD inline V3<S> operator*(const K &s, const V3<S> &a) { return v3(S(s * a.x), S(s * a.y), S(s * a.z)); }
template <class S> OD_HD inline S dot(const V3<S> &a, const V3<S> &b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class S> OD_HD inline V3<S> cross(const V3<S> &a, const V3<S> &b) {
  return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class S> OD_HD inline S comp(const V3<S> &a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
template <class S> OD_HD inline V3<S> sel(int k, const V3<S> &a, const V3<S> &b, const V3<S> &c) {
  return k == 0 ? a : (k == 1 ? b : c);
}
// value part of a vector
template <class S> OD_HD inline auto vval(const V3<S> &a) -> V3<decltype(val(a.x))> {
  return v3(val(a.x), val(a.y), val(a.z));
}
// promote a plain vector to scalar type S
template <class S, class T> OD_HD inline V3<S> lift(const V3<T> &a) { return v3(S(a.x), S(a.y), S(a.z)); }

template <class S>
struct M3 {
  V3<S> r0, r1, r2;   // rows
  OD_HD V3<S> col(int k) const { return v3(comp(r0, k), comp(r1, k), comp(r2, k)); }
  OD_HD V3<S> mul(const V3<S> &v) const { return v3(dot(r0, v), dot(r1, v), dot(r2, v)); }
  OD_HD V3<S> tmul(const V3<S> &v) const { return v.x * r0 + v.y * r1 + v.z * r2; }
};
template <class S> OD_HD inline M3<S> matmul(const M3<S> &A, const M3<S> &B) {
  M3<S> r;
  r.r0 = A.r0.x * B.r0 + A.r0.y * B.r1 + A.r0.z * B.r2;
  r.r1 = A.r1.x * B.r0 + A.r1.y * B.r1 + A.r1.z * B.r2;
  r.r2 = A.r2.x * B.r0 + A.r2.y * B.r1 + A.r2.z * B.r2;
  return r;
}
template <class S> OD_HD inline M3<S> matmul_bt(const M3<S> &A, const M3<S> &B) {
  M3<S> r;
  r.r0 = v3(dot(A.r0, B.r0), dot(A.r0, B.r1), dot(A.r0, B.r2));
  r.r1 = v3(dot(A.r1, B.r0), dot(A.r1, B.r1), dot(A.r1, B.r2));
  r.r2 = v3(dot(A.r2, B.r0), dot(A.r2, B.r1), dot(A.r2, B.r2));
  return r;
}
// R M R^T
template <class S> OD_HD inline M3<S> rotate_tensor(const M3<S> &R, const M3<S> &M) { return matmul(R, matmul_bt(M, R)); }

// rotation of a unit quaternion q = (w, x, y, z) (rotation.cpp:236-255)
template <class S> OD_HD inline M3<S> quat_rot(const S *q) {
  S xx = S(2) * q[1] * q[1], yy = S(2) * q[2] * q[2], zz = S(2) * q[3] * q[3];
  M3<S> m;
  m.r0 = v3(S(1) - yy - zz, S(2) * (q[1] * q[2] - q[0] * q[3]), S(2) * (q[1] * q[3] + q[0] * q[2]));
  m.r1 = v3(S(2) * (q[1] * q[2] + q[0] * q[3]), S(1) - xx - zz, S(2) * (q[2] * q[3] - q[0] * q[1]));
  m.r2 = v3(S(2) * (q[1] * q[3] - q[0] * q[2]), S(2) * (q[2] * q[3] + q[0] * q[1]), S(1) - xx - yy);
  return m;
}

// dPlaneSpace (odemath.cpp:151-178): p, q unit and normal to unit n
template <class S> OD_HD inline void plane_space(const V3<S> &n, V3<S> &p, V3<S> &q) {
  using T = decltype(val(n.x));
  if (std::fabs(val(n.z)) > T(0.70710678118654752)) {
    S a = n.y * n.y + n.z * n.z;
    S k = S(1) / sqrt_(a);
    p = v3(S(0), -n.z * k, n.y * k);
    q = v3(a * k, -n.x * p.z, n.x * p.y);
  } else {
    S a = n.x * n.x + n.y * n.y;
    S k = S(1) / sqrt_(a);
    p = v3(-n.y * k, n.x * k, S(0));
    q = v3(-n.z * p.y, n.z * p.x, a * k);
  }
}

}  // namespace adj
}  // namespace odemi
