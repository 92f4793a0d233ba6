// devmath.hpp -- small vector/quaternion helpers for the HIP kernels.
// Rotations are stored as three float4 rows (row-major 3x3, w unused), the
// layout the reference uses for dMatrix3 (include/ode/common.h:142).
#pragma once
#include <hip/hip_runtime.h>

namespace odemi {

__device__ __forceinline__ float3 mk3(float x, float y, float z) { return make_float3(x, y, z); }
__device__ __forceinline__ float3 xyz(const float4 &a) { return make_float3(a.x, a.y, a.z); }
__device__ __forceinline__ float4 f4(const float3 &a, float w = 0.f) { return make_float4(a.x, a.y, a.z, w); }
__device__ __forceinline__ float3 operator+(float3 a, float3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ float3 operator-(float3 a, float3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float3 operator-(float3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ float3 operator*(float s, float3 a) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float3 operator*(float3 a, float s) { return mk3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ float dot(float3 a, float3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float3 cross(float3 a, float3 b) {
  return mk3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float comp(float3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ float3 sel3(int k, float3 a, float3 b, float3 c) { return k == 0 ? a : (k == 1 ? b : c); }

struct Mat3 {
  float3 r0, r1, r2;  // rows
  __device__ float3 col(int k) const { return mk3(comp(r0, k), comp(r1, k), comp(r2, k)); }
  __device__ float3 mul(float3 v) const { return mk3(dot(r0, v), dot(r1, v), dot(r2, v)); }
  // transpose times v
  __device__ float3 tmul(float3 v) const { return v.x * r0 + v.y * r1 + v.z * r2; }
};

__device__ __forceinline__ Mat3 load_mat(const float4 *R) {
  Mat3 m;
  m.r0 = xyz(R[0]);
  m.r1 = xyz(R[1]);
  m.r2 = xyz(R[2]);
  return m;
}
__device__ __forceinline__ void store_mat(float4 *R, const Mat3 &m) {
  R[0] = f4(m.r0);
  R[1] = f4(m.r1);
  R[2] = f4(m.r2);
}
// A * B
__device__ __forceinline__ Mat3 matmul(const Mat3 &A, const Mat3 &B) {
  Mat3 r;
  r.r0 = A.r0.x * B.r0 + A.r0.y * B.r1 + A.r0.z * B.r2;
  r.r1 = A.r1.x * B.r0 + A.r1.y * B.r1 + A.r1.z * B.r2;
  r.r2 = A.r2.x * B.r0 + A.r2.y * B.r1 + A.r2.z * B.r2;
  return r;
}
// A * B^T
__device__ __forceinline__ Mat3 matmul_bt(const Mat3 &A, const Mat3 &B) {
  Mat3 r;
  r.r0 = mk3(dot(A.r0, B.r0), dot(A.r0, B.r1), dot(A.r0, B.r2));
  r.r1 = mk3(dot(A.r1, B.r0), dot(A.r1, B.r1), dot(A.r1, B.r2));
  r.r2 = mk3(dot(A.r2, B.r0), dot(A.r2, B.r1), dot(A.r2, B.r2));
  return r;
}
// R M R^T (world-frame inertia from the body frame)
__device__ __forceinline__ Mat3 rotate_tensor(const Mat3 &R, const Mat3 &M) { return matmul(R, matmul_bt(M, R)); }

// rotation of a unit quaternion (w,x,y,z) (ode/src/rotation.cpp:236-255)
__device__ __forceinline__ Mat3 quat_to_mat(float4 q) {
  float xx = 2.f * q.y * q.y, yy = 2.f * q.z * q.z, zz = 2.f * q.w * q.w;
  Mat3 m;
  m.r0 = mk3(1.f - yy - zz, 2.f * (q.y * q.z - q.x * q.w), 2.f * (q.y * q.w + q.x * q.z));
  m.r1 = mk3(2.f * (q.y * q.z + q.x * q.w), 1.f - xx - zz, 2.f * (q.z * q.w - q.x * q.y));
  m.r2 = mk3(2.f * (q.y * q.w - q.x * q.z), 2.f * (q.z * q.w + q.x * q.y), 1.f - xx - yy);
  return m;
}
// Note: quaternions are float4 with x = w (scalar), y,z,w = vector part.

// unit vector; zero vector -> (1,0,0) (ode/src/odemath.cpp:42-91)
__device__ __forceinline__ float3 safe_normalize(float3 a) {
  float ax = fabsf(a.x), ay = fabsf(a.y), az = fabsf(a.z);
  float m = fmaxf(ax, fmaxf(ay, az));
  if (!(m > 0.f)) return mk3(1.f, 0.f, 0.f);
  a = (1.f / m) * a;
  return rsqrtf(dot(a, a)) * a;
}
// the same in the reference's own operations (divide by the largest
// magnitude, then multiply by the reciprocal square root), for colliders whose
// contacts follow the reference's rounding closely
__device__ __forceinline__ float3 safe_normalize_div(float3 a) {
  float ax = fabsf(a.x), ay = fabsf(a.y), az = fabsf(a.z);
  float m = (ay > ax) ? ((az > ay) ? az : ay) : ((az > ax) ? az : ax);
  if (!(m > 0.f)) return mk3(1.f, 0.f, 0.f);
  a = mk3(a.x / m, a.y / m, a.z / m);
  const float l = 1.f / sqrtf(a.x * a.x + a.y * a.y + a.z * a.z);
  return mk3(a.x * l, a.y * l, a.z * l);
}
__device__ __forceinline__ float4 quat_normalize(float4 q) {
  float l = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
  if (l > 0.f) {
    float s = 1.f / sqrtf(l);
    return make_float4(q.x * s, q.y * s, q.z * s, q.w * s);
  }
  return make_float4(1.f, 0.f, 0.f, 0.f);
}
// two unit vectors normal to unit n, q = n x p (ode/src/odemath.cpp:151-178)
__device__ __forceinline__ void plane_space(float3 n, float3 &p, float3 &q) {
  if (fabsf(n.z) > 0.70710678118654752f) {
    float a = n.y * n.y + n.z * n.z;
    float k = 1.f / sqrtf(a);
    p = mk3(0.f, -n.z * k, n.y * k);
    q = mk3(a * k, -n.x * p.z, n.x * p.y);
  } else {
    float a = n.x * n.x + n.y * n.y;
    float k = 1.f / sqrtf(a);
    p = mk3(-n.y * k, n.x * k, 0.f);
    q = mk3(-n.z * p.y, n.z * p.x, a * k);
  }
}
// quaternion products (ode/src/rotation.cpp:191-233): 0 b*c, 1 b~*c, 2 b*c~, 3 b~*c~
__device__ __forceinline__ float4 qmul0(float4 b, float4 c) {
  return make_float4(b.x * c.x - b.y * c.y - b.z * c.z - b.w * c.w, b.x * c.y + b.y * c.x + b.z * c.w - b.w * c.z,
                     b.x * c.z + b.z * c.x + b.w * c.y - b.y * c.w, b.x * c.w + b.w * c.x + b.y * c.z - b.z * c.y);
}
__device__ __forceinline__ float4 qmul1(float4 b, float4 c) {
  return make_float4(b.x * c.x + b.y * c.y + b.z * c.z + b.w * c.w, b.x * c.y - b.y * c.x - b.z * c.w + b.w * c.z,
                     b.x * c.z - b.z * c.x - b.w * c.y + b.y * c.w, b.x * c.w - b.w * c.x - b.y * c.z + b.z * c.y);
}
__device__ __forceinline__ float4 qmul2(float4 b, float4 c) {
  return make_float4(b.x * c.x + b.y * c.y + b.z * c.z + b.w * c.w, -b.x * c.y + b.y * c.x - b.z * c.w + b.w * c.z,
                     -b.x * c.z + b.z * c.x - b.w * c.y + b.y * c.w, -b.x * c.w + b.w * c.x - b.y * c.z + b.z * c.y);
}
__device__ __forceinline__ float4 qmul3(float4 b, float4 c) {
  return make_float4(b.x * c.x - b.y * c.y - b.z * c.z - b.w * c.w, -b.x * c.y - b.y * c.x + b.z * c.w - b.w * c.z,
                     -b.x * c.z - b.z * c.x + b.w * c.y - b.y * c.w, -b.x * c.w - b.w * c.x + b.y * c.z - b.z * c.y);
}

// ------------------------------------------------------------------ RNG
// LCG x -> 1664525 x + 1013904223 (mod 2^32) and the xor-folded bounded
// draw used by the row shuffle (ode/src/misc.cpp:33-37, 66-117).
__device__ __forceinline__ unsigned lcg_next(unsigned &s) {
  s = 1664525u * s + 1013904223u;
  return s;
}
__device__ __forceinline__ int lcg_below(unsigned &s, int n) {
  unsigned un = (unsigned)n;
  unsigned r = lcg_next(s);
  if (un <= 0x10u) {
    r ^= r >> 16; r ^= r >> 8; r ^= r >> 4;
    if (un <= 0x2u) { r ^= r >> 2; r ^= r >> 1; }
    else if (un <= 0x4u) { r ^= r >> 2; }
  } else if (un <= 0x100u) {
    r ^= r >> 16; r ^= r >> 8;
  } else if (un <= 0x10000u) {
    r ^= r >> 16;
  }
  return (int)(r % un);
}
// state after k draws: compose the affine map by squaring (O(log k))
__device__ __forceinline__ unsigned lcg_skip(unsigned s, unsigned long long k) {
  unsigned a = 1664525u, c = 1013904223u;   // current power of the map
  unsigned A = 1u, C = 0u;                   // accumulated map
  while (k) {
    if (k & 1ull) { A = a * A; C = a * C + c; }
    c = a * c + c;
    a = a * a;
    k >>= 1;
  }
  return A * s + C;
}

}  // namespace odemi
