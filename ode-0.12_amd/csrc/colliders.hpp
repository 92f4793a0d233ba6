// colliders.hpp -- narrowphase for the path's geom pairs (sphere, box, plane,
// capsule, ray, cylinder), one pair per lane.  Each collider writes up to `maxc` contacts (position, normal into
// body 1, depth) and returns their count; the contact normal convention and
// the contact selection are those of the reference colliders cited below.
#pragma once
#include "devmath.hpp"
#include "layout.hpp"

namespace odemi {

struct GeomView {
  int type;
  float3 p;      // world position (unused for planes)
  Mat3 R;        // world rotation
  float4 prm;    // radius | lx,ly,lz | plane a,b,c,d (unit normal) | capsule or cylinder radius, length | ray length
};

struct CPoint {
  float3 pos;
  float3 n;
  float depth;
};

// world AABB (minx,maxx,miny,maxy,minz,maxz): box.cpp:60-77, sphere.cpp:59-67,
// plane.cpp:80-107 (an axis-aligned plane bounds a half space, others are
// unbounded)
__device__ inline void geom_aabb(const GeomView &g, float *a) {
  if (g.type == GEOM_SPHERE) {
    float r = g.prm.x;
    a[0] = g.p.x - r; a[1] = g.p.x + r;
    a[2] = g.p.y - r; a[3] = g.p.y + r;
    a[4] = g.p.z - r; a[5] = g.p.z + r;
  } else if (g.type == GEOM_BOX) {
    float3 s = mk3(g.prm.x, g.prm.y, g.prm.z);
    float ex = 0.5f * (fabsf(g.R.r0.x * s.x) + fabsf(g.R.r0.y * s.y) + fabsf(g.R.r0.z * s.z));
    float ey = 0.5f * (fabsf(g.R.r1.x * s.x) + fabsf(g.R.r1.y * s.y) + fabsf(g.R.r1.z * s.z));
    float ez = 0.5f * (fabsf(g.R.r2.x * s.x) + fabsf(g.R.r2.y * s.y) + fabsf(g.R.r2.z * s.z));
    a[0] = g.p.x - ex; a[1] = g.p.x + ex;
    a[2] = g.p.y - ey; a[3] = g.p.y + ey;
    a[4] = g.p.z - ez; a[5] = g.p.z + ez;
  } else if (g.type == GEOM_RAY) {
    // ray.cpp:56-88: the box of the start point and the end point (the
    // direction is column 2 of R)
    const float3 e = mk3(g.p.x + g.R.r0.z * g.prm.x, g.p.y + g.R.r1.z * g.prm.x, g.p.z + g.R.r2.z * g.prm.x);
    if (g.p.x < e.x) { a[0] = g.p.x; a[1] = e.x; } else { a[0] = e.x; a[1] = g.p.x; }
    if (g.p.y < e.y) { a[2] = g.p.y; a[3] = e.y; } else { a[2] = e.y; a[3] = g.p.y; }
    if (g.p.z < e.z) { a[4] = g.p.z; a[5] = e.z; } else { a[4] = e.z; a[5] = g.p.z; }
  } else if (g.type == GEOM_CYLINDER) {
    // cylinder.cpp:58-74: the radius along columns 0 and 1 of R, half the
    // length along column 2
    const float r = g.prm.x, l = g.prm.y;
    float ex = fabsf(g.R.r0.x * r) + fabsf(g.R.r0.y * r) + 0.5f * fabsf(g.R.r0.z * l);
    float ey = fabsf(g.R.r1.x * r) + fabsf(g.R.r1.y * r) + 0.5f * fabsf(g.R.r1.z * l);
    float ez = fabsf(g.R.r2.x * r) + fabsf(g.R.r2.y * r) + 0.5f * fabsf(g.R.r2.z * l);
    a[0] = g.p.x - ex; a[1] = g.p.x + ex;
    a[2] = g.p.y - ey; a[3] = g.p.y + ey;
    a[4] = g.p.z - ez; a[5] = g.p.z + ez;
  } else if (g.type == GEOM_CAPSULE) {
    // capsule.cpp:60-74: the axis is column 2 of R
    float ex = fabsf(g.R.r0.z * g.prm.y) * 0.5f + g.prm.x;
    float ey = fabsf(g.R.r1.z * g.prm.y) * 0.5f + g.prm.x;
    float ez = fabsf(g.R.r2.z * g.prm.y) * 0.5f + g.prm.x;
    a[0] = g.p.x - ex; a[1] = g.p.x + ex;
    a[2] = g.p.y - ey; a[3] = g.p.y + ey;
    a[4] = g.p.z - ez; a[5] = g.p.z + ez;
  } else {
    const float inf = __int_as_float(0x7f800000);
    float4 p = g.prm;
    for (int k = 0; k < 6; k++) a[k] = (k & 1) ? inf : -inf;
    int ax = -1;
    if (p.y == 0.f && p.z == 0.f) ax = 0;
    else if (p.x == 0.f && p.z == 0.f) ax = 1;
    else if (p.x == 0.f && p.y == 0.f) ax = 2;
    if (ax >= 0) {
      float c = ax == 0 ? p.x : (ax == 1 ? p.y : p.z);
      a[2 * ax] = (c > 0.f) ? -inf : -p.w;
      a[2 * ax + 1] = (c > 0.f) ? p.w : inf;
    }
  }
}

// sphere-sphere: collision_util.cpp:38-67
__device__ inline int coll_sphere_sphere(float3 p1, float r1, float3 p2, float r2, CPoint *c) {
  float3 d = p1 - p2;
  float dist = sqrtf(dot(d, d));
  if (dist > r1 + r2) return 0;
  if (dist <= 0.f) {
    c->pos = p1;
    c->n = mk3(1.f, 0.f, 0.f);
    c->depth = r1 + r2;
  } else {
    c->n = (1.f / dist) * d;
    c->pos = p1 + (0.5f * (r2 - r1 - dist)) * c->n;
    c->depth = r1 + r2 - dist;
  }
  return 1;
}

// sphere-box: sphere.cpp:131-219
__device__ inline int coll_sphere_box(float3 ps, float rad, float3 pb, const Mat3 &R, float3 side, CPoint *c) {
  float3 p = ps - pb;
  float3 l = 0.5f * side;
  float t[3];
  bool border = false;
  for (int k = 0; k < 3; k++) {
    float3 ax = R.col(k);
    float v = dot(p, ax);
    float lk = comp(l, k);
    if (v < -lk) { v = -lk; border = true; }
    if (v > lk) { v = lk; border = true; }
    t[k] = v;
  }
  if (!border) {
    float best = l.x - fabsf(t[0]);
    int bi = 0;
    float f1 = l.y - fabsf(t[1]);
    if (f1 < best) { best = f1; bi = 1; }
    float f2 = l.z - fabsf(t[2]);
    if (f2 < best) { best = f2; bi = 2; }
    c->pos = ps;
    float sgn = (t[bi] > 0.f) ? 1.f : -1.f;
    c->n = sgn * R.col(bi);
    c->depth = best + rad;
    return 1;
  }
  float3 q = R.mul(mk3(t[0], t[1], t[2]));
  float3 r = p - q;
  float depth = rad - sqrtf(dot(r, r));
  if (depth < 0.f) return 0;
  c->pos = q + pb;
  c->n = safe_normalize(r);
  c->depth = depth;
  return 1;
}

// sphere-plane: sphere.cpp:222-251
__device__ inline int coll_sphere_plane(float3 ps, float rad, float4 pl, CPoint *c) {
  float3 n = mk3(pl.x, pl.y, pl.z);
  float depth = pl.w - dot(ps, n) + rad;
  if (depth < 0.f) return 0;
  c->n = n;
  c->pos = ps - rad * n;
  c->depth = depth;
  return 1;
}

// box-plane: box.cpp:745-878 -- deepest vertex, then the vertices along the
// two box edges with the smallest projections on the normal, then (for 4
// requested contacts) the fourth vertex of that face.
__device__ inline int coll_box_plane(float3 pb, const Mat3 &R, float3 side, float4 pl, int maxc, CPoint *c) {
  float3 n = mk3(pl.x, pl.y, pl.z);
  float3 ax0 = R.col(0), ax1 = R.col(1), ax2 = R.col(2);
  float A[3] = {side.x * dot(n, ax0), side.y * dot(n, ax1), side.z * dot(n, ax2)};
  float B[3] = {fabsf(A[0]), fabsf(A[1]), fabsf(A[2])};
  float depth = pl.w + 0.5f * (B[0] + B[1] + B[2]) - dot(n, pb);
  if (depth < 0.f) return 0;
  if (maxc > 4) maxc = 4;
  float3 p = pb;
  p = (A[0] > 0.f) ? p - (0.5f * side.x) * ax0 : p + (0.5f * side.x) * ax0;
  p = (A[1] > 0.f) ? p - (0.5f * side.y) * ax1 : p + (0.5f * side.y) * ax1;
  p = (A[2] > 0.f) ? p - (0.5f * side.z) * ax2 : p + (0.5f * side.z) * ax2;
  c[0].pos = p;
  c[0].depth = depth;
  int ret = 1;
  int first, second;
  if (B[0] < B[1]) {
    if (B[2] < B[0]) { first = 2; second = 0; }
    else { first = 0; second = (B[1] < B[2]) ? 1 : 2; }
  } else {
    if (B[2] < B[1]) { first = 2; second = (B[0] < B[1]) ? 0 : 1; }
    else { first = 1; second = (B[0] < B[2]) ? 0 : 2; }
  }
  if (maxc > 1) {
    for (int s = 0; s < 2; s++) {
      if (s == 1 && maxc <= 2) break;
      int ax = s == 0 ? first : second;
      float Bk = B[ax];
      if (depth - Bk < 0.f) break;
      float3 axv = sel3(ax, ax0, ax1, ax2);
      float sd = comp(side, ax);
      c[1 + s].pos = (A[ax] > 0.f) ? p + sd * axv : p - sd * axv;
      c[1 + s].depth = depth - Bk;
      ret++;
    }
  }
  if (maxc == 4 && ret == 3) {
    float d4 = c[1].depth + c[2].depth - depth;
    if (d4 > 0.f) {
      c[3].pos = c[1].pos + c[2].pos - p;
      c[3].depth = d4;
      ret++;
    }
  }
  for (int j = 0; j < ret; j++) c[j].n = n;
  return ret;
}

// closest approach of two lines (collision_util.cpp:70-92)
__device__ inline void line_closest(float3 pa, float3 ua, float3 pb, float3 ub, float &al, float &be) {
  float3 p = pb - pa;
  float uaub = dot(ua, ub);
  float q1 = dot(ua, p);
  float q2 = -dot(ub, p);
  float d = 1.f - uaub * uaub;
  if (d <= 0.0001f) { al = 0.f; be = 0.f; return; }
  d = 1.f / d;
  al = (q1 + uaub * q2) * d;
  be = (uaub * q1 + q2) * d;
}

// clip a quadrilateral against the rectangle |x|<h0, |y|<h1 with four
// half-plane chops; stops once a chop has produced 8 points (box.cpp:187-238)
__device__ inline int clip_rect_quad(float h0, float h1, const float *quad, float *ret) {
  float bufA[16], bufB[16];
  for (int i = 0; i < 8; i++) bufA[i] = quad[i];
  float *q = bufA, *r = bufB;
  int nq = 4, nr = 0;
  for (int chop = 0; chop < 4; chop++) {
    int dir = chop >> 1;
    float sign = (chop & 1) ? 1.f : -1.f;
    float h = dir ? h1 : h0;
    nr = 0;
    bool full = false;
    for (int i = 0; i < nq && !full; i++) {
      float ax = q[2 * i], ay = q[2 * i + 1];
      int j = (i + 1 < nq) ? i + 1 : 0;
      float bx = q[2 * j], by = q[2 * j + 1];
      float ad = dir ? ay : ax, bd = dir ? by : bx;
      bool ain = sign * ad < h, bin = sign * bd < h;
      if (ain) {
        r[2 * nr] = ax; r[2 * nr + 1] = ay; nr++;
        if (nr & 8) { full = true; break; }
      }
      if (ain != bin) {
        float ao = dir ? ax : ay, bo = dir ? bx : by;
        float o = ao + (bo - ao) / (bd - ad) * (sign * h - ad);
        float d = sign * h;
        r[2 * nr] = dir ? o : d;
        r[2 * nr + 1] = dir ? d : o;
        nr++;
        if (nr & 8) { full = true; break; }
      }
    }
    float *t = q; q = r; r = t;
    nq = nr;
    if (full) break;
  }
  for (int i = 0; i < 2 * nq; i++) ret[i] = q[i];
  return nq;
}

// choose m of n (<=8) points spread around the centroid, starting at i0
// (box.cpp:249-311)
__device__ inline void cull_points(int n, const float *p, int m, int i0, int *iret) {
  float cx, cy;
  if (n == 1) { cx = p[0]; cy = p[1]; }
  else if (n == 2) { cx = 0.5f * (p[0] + p[2]); cy = 0.5f * (p[1] + p[3]); }
  else {
    float a = 0.f, q;
    cx = 0.f; cy = 0.f;
    for (int i = 0; i < n - 1; i++) {
      q = p[2 * i] * p[2 * i + 3] - p[2 * i + 2] * p[2 * i + 1];
      a += q;
      cx += q * (p[2 * i] + p[2 * i + 2]);
      cy += q * (p[2 * i + 1] + p[2 * i + 3]);
    }
    q = p[2 * n - 2] * p[1] - p[0] * p[2 * n - 1];
    a = 1.f / (3.f * (a + q));
    cx = a * (cx + q * (p[2 * n - 2] + p[0]));
    cy = a * (cy + q * (p[2 * n - 1] + p[1]));
  }
  float ang[8];
  for (int i = 0; i < n; i++) ang[i] = atan2f(p[2 * i + 1] - cy, p[2 * i] - cx);
  unsigned avail = (1u << n) - 1u;
  avail &= ~(1u << i0);
  iret[0] = i0;
  const float PI = 3.14159265358979323846f;
  for (int j = 1; j < m; j++) {
    float a = (float)j * (2.f * PI / (float)m) + ang[i0];
    if (a > PI) a -= 2.f * PI;
    float best = 1e9f;
    int bi = i0;
    for (int i = 0; i < n; i++) {
      if (!(avail & (1u << i))) continue;
      float diff = fabsf(ang[i] - a);
      if (diff > PI) diff = 2.f * PI - diff;
      if (diff < best) { best = diff; bi = i; }
    }
    iret[j] = bi;
    avail &= ~(1u << bi);
  }
}

// box-box: separating axis test over the 15 axes with the 1.05 preference for
// face axes, then one edge-edge contact or face clipping with culling to
// maxc points.  box.cpp:331-712, normal negated as in 716-742.
__device__ __noinline__ int coll_box_box(float3 p1, const Mat3 &R1, float3 side1, float3 p2, const Mat3 &R2, float3 side2,
                                   int maxc, CPoint *out) {
  const float fudge = 1.05f;
  float3 u0 = R1.col(0), u1 = R1.col(1), u2 = R1.col(2);
  float3 v0 = R2.col(0), v1 = R2.col(1), v2 = R2.col(2);
  float3 p = p2 - p1;
  float pp0 = dot(u0, p), pp1 = dot(u1, p), pp2 = dot(u2, p);
  float3 A = 0.5f * side1, B = 0.5f * side2;
  float R11 = dot(u0, v0), R12 = dot(u0, v1), R13 = dot(u0, v2);
  float R21 = dot(u1, v0), R22 = dot(u1, v1), R23 = dot(u1, v2);
  float R31 = dot(u2, v0), R32 = dot(u2, v1), R33 = dot(u2, v2);
  float Q11 = fabsf(R11), Q12 = fabsf(R12), Q13 = fabsf(R13);
  float Q21 = fabsf(R21), Q22 = fabsf(R22), Q23 = fabsf(R23);
  float Q31 = fabsf(R31), Q32 = fabsf(R32), Q33 = fabsf(R33);
  float s = -__int_as_float(0x7f800000);
  bool invert = false;
  int code = 0;
  float3 nrm = mk3(0.f, 0.f, 0.f);   // face axis (world) or edge axis (box-1 frame)
  // face axes of box 1 and box 2
  auto face = [&](float e1, float e2, float3 ax, int cc) -> bool {
    float s2 = fabsf(e1) - e2;
    if (s2 > 0.f) return false;
    if (s2 > s) { s = s2; nrm = ax; invert = e1 < 0.f; code = cc; }
    return true;
  };
  auto edge = [&](float e1, float e2, float n1, float n2, float n3, int cc) -> bool {
    float s2 = fabsf(e1) - e2;
    if (s2 > 0.f) return false;
    float l = sqrtf(n1 * n1 + n2 * n2 + n3 * n3);
    if (l > 0.f) {
      s2 /= l;
      if (s2 * fudge > s) { s = s2; nrm = mk3(n1 / l, n2 / l, n3 / l); invert = e1 < 0.f; code = cc; }
    }
    return true;
  };
  if (!face(pp0, A.x + B.x * Q11 + B.y * Q12 + B.z * Q13, u0, 1)) return 0;
  if (!face(pp1, A.y + B.x * Q21 + B.y * Q22 + B.z * Q23, u1, 2)) return 0;
  if (!face(pp2, A.z + B.x * Q31 + B.y * Q32 + B.z * Q33, u2, 3)) return 0;
  if (!face(dot(v0, p), A.x * Q11 + A.y * Q21 + A.z * Q31 + B.x, v0, 4)) return 0;
  if (!face(dot(v1, p), A.x * Q12 + A.y * Q22 + A.z * Q32 + B.y, v1, 5)) return 0;
  if (!face(dot(v2, p), A.x * Q13 + A.y * Q23 + A.z * Q33 + B.z, v2, 6)) return 0;
  if (!edge(pp2 * R21 - pp1 * R31, A.y * Q31 + A.z * Q21 + B.y * Q13 + B.z * Q12, 0.f, -R31, R21, 7)) return 0;
  if (!edge(pp2 * R22 - pp1 * R32, A.y * Q32 + A.z * Q22 + B.x * Q13 + B.z * Q11, 0.f, -R32, R22, 8)) return 0;
  if (!edge(pp2 * R23 - pp1 * R33, A.y * Q33 + A.z * Q23 + B.x * Q12 + B.y * Q11, 0.f, -R33, R23, 9)) return 0;
  if (!edge(pp0 * R31 - pp2 * R11, A.x * Q31 + A.z * Q11 + B.y * Q23 + B.z * Q22, R31, 0.f, -R11, 10)) return 0;
  if (!edge(pp0 * R32 - pp2 * R12, A.x * Q32 + A.z * Q12 + B.x * Q23 + B.z * Q21, R32, 0.f, -R12, 11)) return 0;
  if (!edge(pp0 * R33 - pp2 * R13, A.x * Q33 + A.z * Q13 + B.x * Q22 + B.y * Q21, R33, 0.f, -R13, 12)) return 0;
  if (!edge(pp1 * R11 - pp0 * R21, A.x * Q21 + A.y * Q11 + B.y * Q33 + B.z * Q32, -R21, R11, 0.f, 13)) return 0;
  if (!edge(pp1 * R12 - pp0 * R22, A.x * Q22 + A.y * Q12 + B.x * Q33 + B.z * Q31, -R22, R12, 0.f, 14)) return 0;
  if (!edge(pp1 * R13 - pp0 * R23, A.x * Q23 + A.y * Q13 + B.x * Q32 + B.y * Q31, -R23, R13, 0.f, 15)) return 0;
  if (!code) return 0;
  float3 normal = (code <= 6) ? nrm : R1.mul(nrm);
  if (invert) normal = -normal;
  float depth = -s;
  int n_out;
  if (code > 6) {
    // edge-edge: halfway between the closest points of the two edges
    float3 pa = p1;
    pa = pa + ((dot(normal, u0) > 0.f) ? 1.f : -1.f) * A.x * u0;
    pa = pa + ((dot(normal, u1) > 0.f) ? 1.f : -1.f) * A.y * u1;
    pa = pa + ((dot(normal, u2) > 0.f) ? 1.f : -1.f) * A.z * u2;
    float3 pb = p2;
    pb = pb + ((dot(normal, v0) > 0.f) ? -1.f : 1.f) * B.x * v0;
    pb = pb + ((dot(normal, v1) > 0.f) ? -1.f : 1.f) * B.y * v1;
    pb = pb + ((dot(normal, v2) > 0.f) ? -1.f : 1.f) * B.z * v2;
    float3 ua = sel3((code - 7) / 3, u0, u1, u2);
    float3 ub = sel3((code - 7) % 3, v0, v1, v2);
    float al, be;
    line_closest(pa, ua, pb, ub, al, be);
    pa = pa + al * ua;
    pb = pb + be * ub;
    out[0].pos = 0.5f * (pa + pb);
    out[0].depth = depth;
    n_out = 1;
  } else {
    // face-something: clip the incident face of box b against the reference
    // face of box a
    bool aIs1 = code <= 3;
    float3 ra0 = aIs1 ? u0 : v0, ra1 = aIs1 ? u1 : v1, ra2 = aIs1 ? u2 : v2;
    float3 rb0 = aIs1 ? v0 : u0, rb1 = aIs1 ? v1 : u1, rb2 = aIs1 ? v2 : u2;
    float3 pa = aIs1 ? p1 : p2, pb = aIs1 ? p2 : p1;
    float3 Sa = aIs1 ? A : B, Sb = aIs1 ? B : A;
    float3 n2 = aIs1 ? normal : -normal;
    float nr0 = dot(rb0, n2), nr1 = dot(rb1, n2), nr2 = dot(rb2, n2);
    float an0 = fabsf(nr0), an1 = fabsf(nr1), an2 = fabsf(nr2);
    int lanr, a1, a2;
    if (an1 > an0) {
      if (an1 > an2) { a1 = 0; lanr = 1; a2 = 2; }
      else { a1 = 0; a2 = 1; lanr = 2; }
    } else {
      if (an0 > an2) { lanr = 0; a1 = 1; a2 = 2; }
      else { a1 = 0; a2 = 1; lanr = 2; }
    }
    float nrl = lanr == 0 ? nr0 : (lanr == 1 ? nr1 : nr2);
    float3 rbl = sel3(lanr, rb0, rb1, rb2);
    float sbl = comp(Sb, lanr);
    float3 center = (nrl < 0.f) ? (pb - pa) + sbl * rbl : (pb - pa) - sbl * rbl;
    int codeN = aIs1 ? code - 1 : code - 4;
    int code1 = (codeN == 0) ? 1 : 0;
    int code2 = (codeN == 2) ? 1 : 2;
    float3 rc1 = sel3(code1, ra0, ra1, ra2), rc2 = sel3(code2, ra0, ra1, ra2);
    float3 rba1 = sel3(a1, rb0, rb1, rb2), rba2 = sel3(a2, rb0, rb1, rb2);
    float c1 = dot(center, rc1), c2 = dot(center, rc2);
    float m11 = dot(rc1, rba1), m12 = dot(rc1, rba2), m21 = dot(rc2, rba1), m22 = dot(rc2, rba2);
    float quad[8];
    {
      float sa1 = comp(Sb, a1), sa2 = comp(Sb, a2);
      float k1 = m11 * sa1, k2 = m21 * sa1, k3 = m12 * sa2, k4 = m22 * sa2;
      quad[0] = c1 - k1 - k3; quad[1] = c2 - k2 - k4;
      quad[2] = c1 - k1 + k3; quad[3] = c2 - k2 + k4;
      quad[4] = c1 + k1 + k3; quad[5] = c2 + k2 + k4;
      quad[6] = c1 + k1 - k3; quad[7] = c2 + k2 - k4;
    }
    float ret[16];
    int n = clip_rect_quad(comp(Sa, code1), comp(Sa, code2), quad, ret);
    if (n < 1) return 0;
    float det1 = 1.f / (m11 * m22 - m12 * m21);
    m11 *= det1; m12 *= det1; m21 *= det1; m22 *= det1;
    float3 point[8];
    float dep[8];
    float san = comp(Sa, codeN);
    int cnum = 0;
    for (int j = 0; j < n; j++) {
      float k1 = m22 * (ret[2 * j] - c1) - m12 * (ret[2 * j + 1] - c2);
      float k2 = -m21 * (ret[2 * j] - c1) + m11 * (ret[2 * j + 1] - c2);
      float3 pt = center + k1 * rba1 + k2 * rba2;
      float d = san - dot(n2, pt);
      if (d >= 0.f) {
        point[cnum] = pt;
        dep[cnum] = d;
        ret[2 * cnum] = ret[2 * j];
        ret[2 * cnum + 1] = ret[2 * j + 1];
        cnum++;
      }
    }
    if (cnum < 1) return 0;
    int mc = maxc;
    if (mc > cnum) mc = cnum;
    if (mc < 1) mc = 1;
    if (cnum <= mc) {
      for (int j = 0; j < cnum; j++) {
        out[j].pos = point[j] + pa;
        out[j].depth = dep[j];
      }
      n_out = cnum;
    } else {
      int i1 = 0;
      float md = dep[0];
      for (int i = 1; i < cnum; i++)
        if (dep[i] > md) { md = dep[i]; i1 = i; }
      int iret[8];
      cull_points(cnum, ret, mc, i1, iret);
      for (int j = 0; j < mc; j++) {
        out[j].pos = point[iret[j]] + pa;
        out[j].depth = dep[iret[j]];
      }
      n_out = mc;
    }
  }
  for (int j = 0; j < n_out; j++) out[j].n = -normal;
  return n_out;
}

// dCollide dispatch (collision_kernel.cpp:292-339): sphere/box come first in
// the collider argument order; a reversed call negates the normals.
// ------------------------------------------------------------ capsules
// A capsule (prm: radius, length) lies along column 2 of its rotation.

// closest points of segments a1-a2 and b1-b2 (collision_util.cpp:109-240):
// the vertex-vertex, edge-vertex and edge-edge features in the reference's
// order of tests
__device__ inline void closest_segment_points(float3 a1, float3 a2, float3 b1, float3 b2, float3 &cp1,
                                              float3 &cp2) {
  const float3 a1a2 = a2 - a1, b1b2 = b2 - b1, a1b1 = b1 - a1;
  const float da1 = dot(a1a2, a1b1), db1 = dot(b1b2, a1b1);
  if (da1 <= 0.f && db1 >= 0.f) { cp1 = a1; cp2 = b1; return; }
  const float3 a1b2 = b2 - a1;
  const float da2 = dot(a1a2, a1b2), db2 = dot(b1b2, a1b2);
  if (da2 <= 0.f && db2 <= 0.f) { cp1 = a1; cp2 = b2; return; }
  const float3 a2b1 = b1 - a2;
  const float da3 = dot(a1a2, a2b1), db3 = dot(b1b2, a2b1);
  if (da3 >= 0.f && db3 >= 0.f) { cp1 = a2; cp2 = b1; return; }
  const float3 a2b2 = b2 - a2;
  const float da4 = dot(a1a2, a2b2), db4 = dot(b1b2, a2b2);
  if (da4 >= 0.f && db4 <= 0.f) { cp1 = a2; cp2 = b2; return; }
  const float la = dot(a1a2, a1a2);
  if (da1 >= 0.f && da3 <= 0.f) {
    const float k = da1 / la;
    if (dot(b1b2, a1b1 - k * a1a2) >= 0.f) { cp1 = a1 + k * a1a2; cp2 = b1; return; }
  }
  if (da2 >= 0.f && da4 <= 0.f) {
    const float k = da2 / la;
    if (dot(b1b2, a1b2 - k * a1a2) <= 0.f) { cp1 = a1 + k * a1a2; cp2 = b2; return; }
  }
  const float lb = dot(b1b2, b1b2);
  if (db1 <= 0.f && db2 >= 0.f) {
    const float k = -db1 / lb;
    if (dot(a1a2, -a1b1 - k * b1b2) >= 0.f) { cp1 = a1; cp2 = b1 + k * b1b2; return; }
  }
  if (db3 <= 0.f && db4 >= 0.f) {
    const float k = -db3 / lb;
    if (dot(a1a2, -a2b1 - k * b1b2) <= 0.f) { cp1 = a2; cp2 = b1 + k * b1b2; return; }
  }
  const float k = dot(a1a2, b1b2);
  float det = la * lb - k * k;
  if (det <= 0.f) { cp1 = a1; cp2 = b1; return; }
  det = 1.f / det;
  const float alpha = (lb * da1 - k * db1) * det;
  const float beta = (k * da1 - la * db1) * det;
  cp1 = a1 + alpha * a1a2;
  cp2 = b1 + beta * b1b2;
}

// the point of segment p1-p2 closest to a box and the box point closest to
// it (collision_util.cpp:247-375): the derivative of the squared distance
// along the segment is piecewise linear in box coordinates; walk its pieces
// until it turns non-negative
__device__ inline void closest_line_box_points(float3 p1, float3 p2, float3 c, const Mat3 &R, float3 side,
                                               float3 &lret, float3 &bret) {
  const float3 d = p2 - p1;
  float s[3], v[3], sg[3], v2[3], h[3], ta[3];
  int rg[3];
  {
    const float3 s3 = R.tmul(p1 - c), v3 = R.tmul(d);
    s[0] = s3.x; s[1] = s3.y; s[2] = s3.z;
    v[0] = v3.x; v[1] = v3.y; v[2] = v3.z;
  }
  for (int i = 0; i < 3; i++) {
    if (v[i] < 0.f) { s[i] = -s[i]; v[i] = -v[i]; sg[i] = -1.f; }
    else sg[i] = 1.f;
    v2[i] = v[i] * v[i];
  }
  h[0] = 0.5f * side.x; h[1] = 0.5f * side.y; h[2] = 0.5f * side.z;
  for (int i = 0; i < 3; i++) {
    if (v[i] > 1e-19f) {
      if (s[i] < -h[i]) { rg[i] = -1; ta[i] = (-h[i] - s[i]) / v[i]; }
      else { rg[i] = (s[i] > h[i]) ? 1 : 0; ta[i] = (h[i] - s[i]) / v[i]; }
    } else {
      rg[i] = 0;
      ta[i] = 2.f;
    }
  }
  float t = 0.f, dd2dt = 0.f;
  for (int i = 0; i < 3; i++) dd2dt -= (rg[i] ? v2[i] : 0.f) * ta[i];
  if (!(dd2dt >= 0.f)) {
    // each pass crosses at least one clip plane (at most six in all)
    for (int pass = 0; pass < 8; pass++) {
      float next_t = 1.f;
      for (int i = 0; i < 3; i++)
        if (ta[i] > t && ta[i] < 1.f && ta[i] < next_t) next_t = ta[i];
      float nd = 0.f;
      for (int i = 0; i < 3; i++) nd += (rg[i] ? v2[i] : 0.f) * (next_t - ta[i]);
      if (nd >= 0.f) {
        const float m = (nd - dd2dt) / (next_t - t);
        t -= dd2dt / m;
        break;
      }
      for (int i = 0; i < 3; i++)
        if (ta[i] == next_t) { ta[i] = (h[i] - s[i]) / v[i]; rg[i]++; }
      t = next_t;
      dd2dt = nd;
      if (!(t < 1.f)) { t = 1.f; break; }
    }
  }
  lret = p1 + t * d;
  float b[3];
  for (int i = 0; i < 3; i++) {
    b[i] = sg[i] * (s[i] + t * v[i]);
    if (b[i] < -h[i]) b[i] = -h[i];
    else if (b[i] > h[i]) b[i] = h[i];
  }
  bret = R.mul(mk3(b[0], b[1], b[2])) + c;
}

__device__ inline float3 capsule_axis(const Mat3 &R) { return mk3(R.r0.z, R.r1.z, R.r2.z); }

// capsule-sphere: capsule.cpp:130-162
__device__ inline int coll_capsule_sphere(const GeomView &cap, float3 ps, float rs, CPoint *c) {
  const float3 ax = capsule_axis(cap.R);
  float alpha = ax.x * (ps.x - cap.p.x) + ax.y * (ps.y - cap.p.y) + ax.z * (ps.z - cap.p.z);
  const float lz2 = cap.prm.y * 0.5f;
  if (alpha > lz2) alpha = lz2;
  if (alpha < -lz2) alpha = -lz2;
  return coll_sphere_sphere(cap.p + alpha * ax, cap.prm.x, ps, rs, c);
}

// capsule-box: capsule.cpp:178-231
__device__ inline int coll_capsule_box(const GeomView &cap, const GeomView &box, CPoint *c) {
  const float3 ax = capsule_axis(cap.R);
  const float clen = cap.prm.y * 0.5f;
  const float3 p1 = cap.p + clen * ax, p2 = cap.p - clen * ax;
  float3 pl, pb;
  closest_line_box_points(p1, p2, box.p, box.R, mk3(box.prm.x, box.prm.y, box.prm.z), pl, pb);
  const float3 d = pl - pb;
  const float rad = cap.prm.x;
  if (sqrtf(d.x * d.x + d.y * d.y + d.z * d.z) < 1e-9f) {
    // the axis point lies in the box: the normal points from the box
    // centre (dCollideSpheresZeroDist, capsule.cpp:165-176)
    c->n = safe_normalize(pb - box.p);
    c->depth = rad + 0.f;
    c->pos = pl + (0.5f * (0.f - rad)) * c->n;
    return 1;
  }
  return coll_sphere_sphere(pl, rad, pb, 0.f, c);
}

// capsule-capsule: capsule.cpp:234-349 -- two contacts along nearly parallel
// axes when both are made, else one at the middle of the overlap, else the
// closest points of the axis segments
__device__ inline int coll_capsule_capsule(const GeomView &g1, const GeomView &g2, int maxc, CPoint *c) {
  const float tolerance = 1e-5f;
  const float lz1 = g1.prm.y * 0.5f, lz2 = g2.prm.y * 0.5f;
  const float3 axis1 = capsule_axis(g1.R);
  float3 axis2 = capsule_axis(g2.R);
  const float r1 = g1.prm.x, r2 = g2.prm.x;
  const float a1a2 = dot(axis1, axis2);
  const float det = 1.0f - a1a2 * a1a2;
  if (det < tolerance) {
    if (a1a2 < 0.f) axis2 = -axis2;
    const float k = dot(axis1, g1.p - g2.p);
    const float a1lo = -lz1, a1hi = lz1, a2lo = -lz2 - k, a2hi = lz2 - k;
    const float lo = (a1lo > a2lo) ? a1lo : a2lo;
    const float hi = (a1hi < a2hi) ? a1hi : a2hi;
    if (lo <= hi) {
      if (maxc >= 2 && lo < hi) {
        if (coll_sphere_sphere(g1.p + lo * axis1, r1, g2.p + (lo + k) * axis2, r2, c) &&
            coll_sphere_sphere(g1.p + hi * axis1, r1, g2.p + (hi + k) * axis2, r2, c + 1))
          return 2;
      }
      const float alpha1 = (lo + hi) * 0.5f;
      const float alpha2 = alpha1 + k;
      return coll_sphere_sphere(g1.p + alpha1 * axis1, r1, g2.p + alpha2 * axis2, r2, c);
    }
  }
  float3 s1, s2;
  closest_segment_points(g1.p + lz1 * axis1, g1.p - lz1 * axis1, g2.p + lz2 * axis2, g2.p - lz2 * axis2, s1, s2);
  return coll_sphere_sphere(s1, r1, s2, r2, c);
}

// capsule-plane: capsule.cpp:352-409 -- the deeper end cap, then the other
__device__ inline int coll_capsule_plane(const GeomView &cap, float4 pl, int maxc, CPoint *c) {
  const float3 n = mk3(pl.x, pl.y, pl.z);
  const float3 ax = capsule_axis(cap.R);
  const float sign = (dot(n, ax) > 0.f) ? -1.0f : 1.0f;
  const float rad = cap.prm.x;
  float3 p = cap.p + (cap.prm.y * 0.5f * sign) * ax;
  float depth = pl.w - dot(p, n) + rad;
  if (depth < 0.f) return 0;
  c[0].n = n;
  c[0].pos = p - rad * n;
  c[0].depth = depth;
  if (maxc < 2) return 1;
  p = cap.p - (cap.prm.y * 0.5f * sign) * ax;
  depth = pl.w - dot(p, n) + rad;
  if (depth < 0.f) return 1;
  c[1].n = n;
  c[1].pos = p - rad * n;
  c[1].depth = depth;
  return 2;
}

// ------------------------------------------------------------ rays
// A ray (prm: length) starts at its position along column 2 of its rotation;
// the contact depth is the distance along the ray.

// ray.cpp:196-232: where the ray enters a sphere (mode 1 from outside: where
// it leaves)
__device__ inline int ray_sphere_hit(const GeomView &ray, float3 ps, float rad, int mode, CPoint *c) {
  const float3 dir = capsule_axis(ray.R);
  const float3 q = ray.p - ps;
  const float B = q.x * dir.x + q.y * dir.y + q.z * dir.z;
  const float C = dot(q, q) - rad * rad;
  float k = B * B - C;
  if (k < 0.f) return 0;
  k = sqrtf(k);
  float alpha;
  if (mode && C >= 0.f) {
    alpha = -B + k;
    if (alpha < 0.f) return 0;
  } else {
    alpha = -B - k;
    if (alpha < 0.f) {
      alpha = -B + k;
      if (alpha < 0.f) return 0;
    }
  }
  if (alpha > ray.prm.x) return 0;
  c->pos = ray.p + alpha * dir;
  const float nsign = (C < 0.f || mode) ? -1.0f : 1.0f;
  c->n = safe_normalize(nsign * (c->pos - ps));
  c->depth = alpha;
  return 1;
}

// ray-box: ray.cpp:253-349 -- slab intersection in box coordinates
__device__ inline int coll_ray_box(const GeomView &ray, const GeomView &box, CPoint *c) {
  const float3 dir = capsule_axis(ray.R);
  float s[3], v[3], sg[3], h[3];
  {
    const float3 s3 = box.R.tmul(ray.p - box.p), v3 = box.R.tmul(dir);
    s[0] = s3.x; s[1] = s3.y; s[2] = s3.z;
    v[0] = v3.x; v[1] = v3.y; v[2] = v3.z;
  }
  for (int i = 0; i < 3; i++) {
    if (v[i] < 0.f) { s[i] = -s[i]; v[i] = -v[i]; sg[i] = 1.f; }
    else sg[i] = -1.f;
  }
  h[0] = 0.5f * box.prm.x; h[1] = 0.5f * box.prm.y; h[2] = 0.5f * box.prm.z;
  if ((s[0] < -h[0] && v[0] <= 0.f) || s[0] > h[0] || (s[1] < -h[1] && v[1] <= 0.f) || s[1] > h[1] ||
      (s[2] < -h[2] && v[2] <= 0.f) || s[2] > h[2] || (v[0] == 0.f && v[1] == 0.f && v[2] == 0.f))
    return 0;
  const float inf = __int_as_float(0x7f800000);
  float lo = -inf, hi = inf;
  int nlo = 0, nhi = 0;
  for (int i = 0; i < 3; i++) {
    if (v[i] != 0.f) {
      float k = (-h[i] - s[i]) / v[i];
      if (k > lo) { lo = k; nlo = i; }
      k = (h[i] - s[i]) / v[i];
      if (k < hi) { hi = k; nhi = i; }
    }
  }
  if (lo > hi) return 0;
  const float alpha = lo >= 0.f ? lo : hi;
  const int n = lo >= 0.f ? nlo : nhi;
  if (alpha < 0.f || alpha > ray.prm.x) return 0;
  c->pos = ray.p + alpha * dir;
  c->n = sg[n] * box.R.col(n);
  c->depth = alpha;
  return 1;
}

// ray-capsule: ray.cpp:352-464 -- the infinite cylinder, then the end caps
__device__ inline int coll_ray_capsule(const GeomView &ray, const GeomView &cap, CPoint *c) {
  const float3 dir = capsule_axis(ray.R), ax = capsule_axis(cap.R);
  const float lz2 = cap.prm.y * 0.5f, rad = cap.prm.x;
  const float3 cs = ray.p - cap.p;
  float k = ax.x * cs.x + ax.y * cs.y + ax.z * cs.z;
  float3 q = k * ax - cs;
  const float C = dot(q, q) - rad * rad;
  int inside = 0;
  if (C < 0.f) {
    if (k < -lz2) k = -lz2;
    else if (k > lz2) k = lz2;
    const float3 r = cap.p + k * ax;
    const float3 d = ray.p - r;
    if (d.x * d.x + d.y * d.y + d.z * d.z < rad * rad) inside = 1;
  }
  if (!inside && C < 0.f) {
    k = (k < 0.f) ? -lz2 : lz2;
  } else {
    const float uv = ax.x * dir.x + ax.y * dir.y + ax.z * dir.z;
    const float3 r = uv * ax - dir;
    float A = dot(r, r);
    const float B = 2.f * dot(q, r);
    k = B * B - 4.f * A * C;
    if (k < 0.f) {
      if (!inside) return 0;
      k = (uv < 0.f) ? -lz2 : lz2;
    } else {
      k = sqrtf(k);
      A = 1.f / (2.f * A);
      float alpha = (-B - k) * A;
      if (alpha < 0.f) {
        alpha = (-B + k) * A;
        if (alpha < 0.f) return 0;
      }
      if (alpha > ray.prm.x) return 0;
      c->pos = ray.p + alpha * dir;
      q = c->pos - cap.p;
      k = q.x * ax.x + q.y * ax.y + q.z * ax.z;
      const float nsign = inside ? -1.0f : 1.0f;
      if (k >= -lz2 && k <= lz2) {
        c->n = safe_normalize(nsign * (c->pos - (cap.p + k * ax)));
        c->depth = alpha;
        return 1;
      }
      k = (k < 0.f) ? -lz2 : lz2;
    }
  }
  return ray_sphere_hit(ray, cap.p + k * ax, rad, inside, c);
}

// ray-plane: ray.cpp:467-497
__device__ inline int coll_ray_plane(const GeomView &ray, float4 pl, CPoint *c) {
  const float3 n = mk3(pl.x, pl.y, pl.z), dir = capsule_axis(ray.R);
  float alpha = pl.w - dot(n, ray.p);
  const float nsign = (alpha > 0.f) ? -1.0f : 1.0f;
  const float k = n.x * dir.x + n.y * dir.y + n.z * dir.z;
  if (k == 0.f) return 0;
  alpha /= k;
  if (alpha < 0.f || alpha > ray.prm.x) return 0;
  c->pos = ray.p + alpha * dir;
  c->n = nsign * n;
  c->depth = alpha;
  return 1;
}

// ------------------------------------------------------------ cylinders
// A cylinder (prm: radius, length) has its axis along column 2 of R.  The
// reference without libccd has cylinder-sphere, cylinder-plane,
// cylinder-box and ray-cylinder colliders; cylinder-cylinder and
// capsule-cylinder pairs give no contacts.  The parallel-test tolerance is
// the single-precision build's (collision_cylinder_sphere.cpp:66-71).
constexpr float CYL_TOL = 0.0001f;

// cylinder-sphere: collision_cylinder_sphere.cpp:57-278 -- s along the axis
// from the first disc, t from the axis; the side, a disc or a disc's rim
__device__ inline int coll_cylinder_sphere(const GeomView &cyl, float3 ps, float r2, CPoint *c) {
  const float radius = cyl.prm.x, length = cyl.prm.y;
  const float3 ax = capsule_axis(cyl.R);
  const float h = length * 0.5f;
  const float3 top = mk3(ax.x * h + cyl.p.x, ax.y * h + cyl.p.y, ax.z * h + cyl.p.z);
  const float3 bot = mk3(ax.x * -h + cyl.p.x, ax.y * -h + cyl.p.y, ax.z * -h + cyl.p.z);
  const float s = (ps.x - bot.x) * ax.x - (bot.y - ps.y) * ax.y - (bot.z - ps.z) * ax.z;
  if (s < -r2 || s > length + r2) return 0;
  float3 C = mk3(s * ax.x + bot.x - ps.x, s * ax.y + bot.y - ps.y, s * ax.z + bot.z - ps.z);
  const float t = sqrtf(C.x * C.x + C.y * C.y + C.z * C.z);
  if (t > radius + r2) return 0;
  if (t > radius && (s < 0.f || s > length)) {
    const float3 disc = (s <= 0.f) ? bot : top;
    const float e = (s <= 0.f) ? s : s - length;
    c->depth = r2 - sqrtf(e * e + (t - radius) * (t - radius));
    if (c->depth < 0.f) return 0;
    c->pos = mk3(C.x / t * -radius + disc.x, C.y / t * -radius + disc.y, C.z / t * -radius + disc.z);
    const float den = r2 - c->depth;
    c->n = mk3((c->pos.x - ps.x) / den, (c->pos.y - ps.y) / den, (c->pos.z - ps.z) / den);
    return 1;
  }
  if ((radius - t) <= s && (radius - t) <= (length - s)) {
    c->depth = (r2 + radius) - t;
    if (c->depth < 0.f) return 0;
    if (t > r2 + CYL_TOL) {
      C = mk3(C.x / t, C.y / t, C.z / t);
      c->pos = mk3(C.x * r2 + ps.x, C.y * r2 + ps.y, C.z * r2 + ps.z);
      c->n = C;
    } else {
      c->pos = C + ps;
      c->n = mk3(C.x / t, C.y / t, C.z / t);
    }
    return 1;
  }
  if (s <= length * 0.5f) {
    c->depth = s + r2;
    if (c->depth < 0.f) return 0;
    c->pos = mk3(r2 * ax.x + ps.x, r2 * ax.y + ps.y, r2 * ax.z + ps.z);
    c->n = ax;
  } else {
    c->depth = r2 + length - s;
    if (c->depth < 0.f) return 0;
    c->pos = mk3(r2 * -ax.x + ps.x, r2 * -ax.y + ps.y, r2 * -ax.z + ps.z);
    c->n = -ax;
  }
  return 1;
}

// cylinder-plane: collision_cylinder_plane.cpp:40-267 -- a disc parallel to
// the plane gives up to four rim points, otherwise the deepest rim point of
// each disc
__device__ inline int coll_cylinder_plane(const GeomView &cyl, float4 pl, int maxc, CPoint *c) {
  const float radius = cyl.prm.x, length = cyl.prm.y;
  const float3 ax = capsule_axis(cyl.R), pn = mk3(pl.x, pl.y, pl.z);
  const float h = length * 0.5f;
  const float3 top = mk3(ax.x * h + cyl.p.x, ax.y * h + cyl.p.y, ax.z * h + cyl.p.z);
  const float3 bot = mk3(ax.x * -h + cyl.p.x, ax.y * -h + cyl.p.y, ax.z * -h + cyl.p.z);
  int n = 0;
  // a candidate point; `plain`: depth from the dot product (else the
  // expanded form the reference writes for the second disc)
  auto emit = [&](float3 q, bool strict, bool plain) -> bool {
    const float d = plain ? pl.w - dot(pn, q) : pl.w - pl.x * q.x - pl.y * q.y - pl.z * q.z;
    if (strict ? d > 0.f : d >= 0.f) {
      c[n].pos = q;
      c[n].n = pn;
      c[n].depth = d;
      n++;
      return n >= maxc;
    }
    return false;
  };
  float s = ax.x * pl.x + ax.y * pl.y + ax.z * pl.z;
  if (s < 0.f) s += 1.0f;
  else s -= 1.0f;
  if (s < CYL_TOL && s > -CYL_TOL) {
    const float d1 = pl.w - dot(pn, bot), d2 = pl.w - dot(pn, top);
    float3 P;
    if (d1 >= d2) {
      if (!(d1 >= 0.f)) return 0;
      P = bot;
    } else {
      if (!(d2 >= 0.f)) return 0;
      P = top;
    }
    float3 V1 = (ax.x < CYL_TOL && ax.x > -CYL_TOL) ? mk3(ax.x + 1.0f, ax.y, ax.z) : mk3(ax.x, ax.y + 1.0f, ax.z);
    float3 V2 = cross(V1, ax);
    float l = sqrtf(V2.x * V2.x + V2.y * V2.y + V2.z * V2.z);
    l = radius / l;
    V2 = mk3(V2.x * l, V2.y * l, V2.z * l);
    V1 = cross(V2, ax);
    if (emit(P + V1, true, true)) return n;
    if (emit(P - V1, true, true)) return n;
    if (emit(P + V2, true, true)) return n;
    emit(P - V2, true, true);
    return n;
  }
  const float t = dot(pn, ax);
  float3 C = mk3(ax.x * t - pl.x, ax.y * t - pl.y, ax.z * t - pl.z);
  float l = sqrtf(C.x * C.x + C.y * C.y + C.z * C.z);
  l = radius / l;
  C = mk3(C.x * l, C.y * l, C.z * l);
  if (emit(C + bot, false, true)) return n;
  emit(C + top, false, false);
  return n;
}

// cylinder-box: collision_cylinder_box.cpp -- separating axes over the box
// faces, the cylinder axis, their cross products, the box vertices and the
// twelve box edges against both rims; then the cylinder's side segment
// clipped by the box, or the box face clipped to an 8-gon of the disc
struct CylBoxDev {
  float3 pc, pb, ax, half, diff, nrm;
  Mat3 Rc, Rb;
  float radius, size, best_depth, best_rb, best_rc;
  int best_axis;
  float3 vert[8];

  __device__ CylBoxDev(const GeomView &cyl, const GeomView &box) {
    pc = cyl.p; pb = box.p; Rc = cyl.R; Rb = box.R;
    radius = cyl.prm.x; size = cyl.prm.y;
    ax = capsule_axis(Rc);
    half = mk3(box.prm.x * 0.5f, box.prm.y * 0.5f, box.prm.z * 0.5f);
    // the vertex numbering of collision_cylinder_box.cpp:130-180
    const float sx[8] = {-1, 1, -1, 1, 1, 1, -1, -1}, sy[8] = {1, 1, -1, -1, 1, -1, -1, 1},
                sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
#pragma unroll
    for (int v = 0; v < 8; v++) vert[v] = Rb.mul(mk3(sx[v] * half.x, sy[v] * half.y, sz[v] * half.z)) + pb;
    diff = pc - pb;
    nrm = mk3(0.f, 0.f, 0.f);
    best_depth = __int_as_float(0x7f800000);
    best_rb = best_rc = 0.f;
    best_axis = 0;
  }

  // collision_cylinder_box.cpp:206-272
  __device__ bool test_axis(float3 n, int id) {
    const float l = sqrtf(n.x * n.x + n.y * n.y + n.z * n.z);
    if (l < 1e-5f) return true;
    n = safe_normalize_div(n);
    const float d1 = dot(ax, n);
    float rc;
    if (d1 > 1.0f || d1 < -1.0f) rc = size * 0.5f;
    else rc = fabsf(d1 * (size * 0.5f)) + radius * sqrtf(1.0f - d1 * d1);
    float rb = fabsf(dot(Rb.col(0), n)) * half.x;
    rb += fabsf(dot(Rb.col(1), n)) * half.y;
    rb += fabsf(dot(Rb.col(2), n)) * half.z;
    const float fd = dot(diff, n);
    float depth = rc + rb;
    if (fabsf(fd) > depth) return false;
    depth -= fabsf(fd);
    if (depth < best_depth) {
      best_depth = depth;
      nrm = fd > 0.f ? -n : n;
      best_axis = id;
      best_rb = rb;
      best_rc = rc;
    }
    return true;
  }

  // collision_cylinder_box.cpp:275-320
  __device__ bool test_edge_rim(float3 centre, float3 v0, float3 v1, int id) {
    const float3 dir = safe_normalize_div(v1 - v0);
    const float d2 = dot(dir, ax);
    if (fabsf(d2) < 1e-5f) return true;
    const float d1 = dot(centre - v0, ax);
    const float k = d1 / d2;
    const float3 p = mk3(v0.x + dir.x * k, v0.y + dir.y * k, v0.z + dir.z * k);
    const float3 tan = cross(centre - p, ax);
    return test_axis(cross(tan, dir), id);
  }

  // collision_cylinder_box.cpp:323-560
  __device__ bool separated() {
    for (int k = 0; k < 3; k++)
      if (!test_axis(Rb.col(k), 1 + k)) return true;
    if (!test_axis(ax, 4)) return true;
    for (int k = 0; k < 3; k++) {
      const float3 a = cross(ax, Rb.col(k));
      if (dot(a, a) > 1e-6f && !test_axis(a, 5 + k)) return true;
    }
    for (int v = 0; v < 8; v++) {
      const float3 a = cross(ax, cross(ax, vert[v] - pc));
      if (dot(a, a) > 1e-6f && !test_axis(a, 8 + v)) return true;
    }
    const int e0[12] = {1, 1, 2, 2, 4, 4, 0, 5, 5, 2, 4, 6}, e1[12] = {0, 3, 3, 0, 1, 7, 7, 3, 6, 6, 5, 7};
    const float h = size * 0.5f;
    for (int side = 0; side < 2; side++) {
      const float3 cc = side == 0 ? mk3(pc.x + ax.x * h, pc.y + ax.y * h, pc.z + ax.z * h)
                                  : mk3(pc.x - ax.x * h, pc.y - ax.y * h, pc.z - ax.z * h);
      for (int e = 0; e < 12; e++)
        if (!test_edge_rim(cc, vert[e0[e]], vert[e1[e]], 16 + 12 * side + e)) return true;
    }
    return false;
  }

  // collision_util.cpp:471-509
  __device__ static bool clip_edge(float3 &a, float3 &b, float4 pl) {
    const float d0 = pl.x * a.x + pl.y * a.y + pl.z * a.z + pl.w;
    const float d1 = pl.x * b.x + pl.y * b.y + pl.z * b.z + pl.w;
    if (d0 < 0.f && d1 < 0.f) return false;
    if (d0 > 0.f && d1 > 0.f) return true;
    if ((d0 > 0.f && d1 < 0.f) || (d0 < 0.f && d1 > 0.f)) {
      const float3 x = mk3(a.x - (a.x - b.x) * d0 / (d0 - d1), a.y - (a.y - b.y) * d0 / (d0 - d1),
                           a.z - (a.z - b.z) * d0 / (d0 - d1));
      if (d0 < 0.f) a = x;
      else b = x;
    }
    return true;
  }

  // collision_util.cpp:512-555 (at most 16 points: 4 corners + one per plane)
  __device__ static void clip_poly(const float3 *in, int nin, float3 *out, int &nout, float4 pl) {
    nout = 0;
    for (int i1 = 0, i0 = nin - 1; i1 < nin; i0 = i1, i1++) {
      const float d0 = pl.x * in[i0].x + pl.y * in[i0].y + pl.z * in[i0].z + pl.w;
      const float d1 = pl.x * in[i1].x + pl.y * in[i1].y + pl.z * in[i1].z + pl.w;
      if (d0 >= 0.f) out[nout++] = in[i0];
      if ((d0 > 0.f && d1 < 0.f) || (d0 < 0.f && d1 > 0.f))
        out[nout++] = mk3(in[i0].x - (in[i0].x - in[i1].x) * d0 / (d0 - d1),
                          in[i0].y - (in[i0].y - in[i1].y) * d0 / (d0 - d1),
                          in[i0].z - (in[i0].z - in[i1].z) * d0 / (d0 - d1));
    }
  }

  // collision_cylinder_box.cpp:563-680
  __device__ int clip_side(int maxc, CPoint *c) {
    const float t = dot(ax, nrm);
    const float3 vn = safe_normalize_div(mk3(nrm.x - ax.x * t, nrm.y - ax.y * t, nrm.z - ax.z * t));
    const float h = size * 0.5f;
    const float3 q = mk3(pc.x + vn.x * radius, pc.y + vn.y * radius, pc.z + vn.z * radius);
    float3 a = mk3(q.x + ax.x * h, q.y + ax.y * h, q.z + ax.z * h) - pb;
    float3 b = mk3(q.x - ax.x * h, q.y - ax.y * h, q.z - ax.z * h) - pb;
    for (int sg = 0; sg < 2; sg++)
      for (int k = 0; k < 3; k++) {
        float3 col = Rb.col(k);
        if (sg) col = -col;
        if (!clip_edge(a, b, make_float4(col.x, col.y, col.z, comp(half, k)))) return 0;
      }
    float dep0 = best_rb + dot(a, nrm), dep1 = best_rb + dot(b, nrm);
    if (dep0 < 0.f) dep0 = 0.f;
    if (dep1 < 0.f) dep1 = 0.f;
    int n = 0;
    c[n].pos = a + pb; c[n].n = -nrm; c[n].depth = dep0; n++;
    if (n != maxc) { c[n].pos = b + pb; c[n].n = -nrm; c[n].depth = dep1; n++; }
    return n;
  }

  // collision_util.h:176-210, the inverse as written there (the division
  // binds to the second product of the diagonal terms)
  __device__ static Mat3 inv3(const Mat3 &M) {
    const float m0 = M.r0.x, m1 = M.r0.y, m2 = M.r0.z, m4 = M.r1.x, m5 = M.r1.y, m6 = M.r1.z, m8 = M.r2.x,
                m9 = M.r2.y, m10 = M.r2.z;
    const float det = m0 * (m5 * m10 - m9 * m6) - m1 * (m4 * m10 - m8 * m6) + m2 * (m4 * m9 - m8 * m5);
    Mat3 d;
    if (fabsf(det) < 0.0005f) {
      d.r0 = mk3(1.f, 0.f, 0.f); d.r1 = mk3(0.f, 1.f, 0.f); d.r2 = mk3(0.f, 0.f, 1.f);
      return d;
    }
    d.r0 = mk3(m5 * m10 - m6 * m9 / det, -(m1 * m10 - m9 * m2) / det, m1 * m6 - m5 * m2 / det);
    d.r1 = mk3(-(m4 * m10 - m6 * m8) / det, m0 * m10 - m8 * m2 / det, -(m0 * m6 - m4 * m2) / det);
    d.r2 = mk3(m4 * m9 - m8 * m5 / det, -(m0 * m9 - m8 * m1) / det, m0 * m5 - m1 * m4 / det);
    return d;
  }

  // collision_cylinder_box.cpp:683-1000
  __device__ int clip_face(int maxc, CPoint *c) {
    const float h = size * 0.5f;
    float3 cpos;
    float cz;
    if (dot(ax, nrm) > 0.f) { cpos = mk3(pc.x + ax.x * h, pc.y + ax.y * h, pc.z + ax.z * h); cz = -1.0f; }
    else { cpos = mk3(pc.x - ax.x * h, pc.y - ax.y * h, pc.z - ax.z * h); cz = 1.0f; }
    const float3 nr = inv3(Rb).mul(nrm);
    const float a0 = fabsf(nr.x), a1 = fabsf(nr.y), a2 = fabsf(nr.z);
    int b0, b1, b2;
    if (a1 > a0) {
      if (a0 > a2) { b0 = 1; b1 = 0; b2 = 2; }
      else if (a1 > a2) { b0 = 1; b1 = 2; b2 = 0; }
      else { b0 = 2; b1 = 1; b2 = 0; }
    } else {
      if (a1 > a2) { b0 = 0; b1 = 1; b2 = 2; }
      else if (a0 > a2) { b0 = 0; b1 = 2; b2 = 1; }
      else { b0 = 2; b1 = 0; b2 = 1; }
    }
    const float3 c0 = Rb.col(b0);
    const float hb0 = comp(half, b0);
    const float3 ctr = comp(nr, b0) > 0.f ? mk3(pb.x - hb0 * c0.x, pb.y - hb0 * c0.y, pb.z - hb0 * c0.z)
                                          : mk3(pb.x + hb0 * c0.x, pb.y + hb0 * c0.y, pb.z + hb0 * c0.z);
    const float3 u1 = comp(half, b1) * Rb.col(b1), u2 = comp(half, b2) * Rb.col(b2);
    float3 A[16], B[16];
    float3 pts[4];
    pts[0] = (ctr + u1) - u2;
    pts[1] = (ctr - u1) - u2;
    pts[2] = (ctr - u1) + u2;
    pts[3] = (ctr + u1) + u2;
    const Mat3 ci = inv3(Rc);
    for (int k = 0; k < 4; k++) pts[k] = ci.mul(pts[k] - cpos);
    int n1 = 0, n2 = 0;
    clip_poly(pts, 4, A, n1, make_float4(0.f, 0.f, cz, 0.f));
    float ang = (float)(3.14159265358979323846 / 8);
    const float inc = ang * 2.0f;
    for (int seg = 0; seg < 8; seg++) {
      const float4 pl = make_float4(-cosf(ang), -sinf(ang), 0.f, radius);
      ang += inc;
      if (seg % 2 == 0) clip_poly(A, n1, B, n2, pl);
      else clip_poly(B, n2, A, n1, pl);
    }
    int n = 0;
    for (int k = 0; k < n1; k++) {
      const float3 p = Rc.mul(A[k]) + cpos;
      const float d = best_rc - dot(p - pc, nrm);
      if (d > 0.f) {
        c[n].pos = p; c[n].n = -nrm; c[n].depth = d;
        if (++n == maxc) break;
      }
    }
    return n;
  }
};

__device__ __attribute__((noinline)) int coll_cylinder_box(const GeomView &cyl, const GeomView &box, int maxc,
                                                          CPoint *c) {
  CylBoxDev S(cyl, box);
  if (S.separated() || S.best_axis == 0) return 0;
  if (fabsf(dot(S.nrm, S.ax)) < 0.9f) return S.clip_side(maxc, c);
  return S.clip_face(maxc, c);
}

// ray-cylinder: ray.cpp:500-656 -- the flat caps when the ray runs inside
// the infinite cylinder parallel to the axis, else the curved side
__device__ inline int coll_ray_cylinder(const GeomView &ray, const GeomView &cyl, CPoint *c) {
  const float len = ray.prm.x, radius = cyl.prm.x, hl = cyl.prm.y * 0.5f;
  const float3 dir = capsule_axis(ray.R), ax = capsule_axis(cyl.R);
  float3 r = ray.p - cyl.p;
  float d = ax.x * r.x + ax.y * r.y + ax.z * r.z;
  float3 q = mk3(d * ax.x - r.x, d * ax.y - r.y, d * ax.z - r.z);
  const float C = dot(q, q) - radius * radius;
  const float uv = ax.x * dir.x + ax.y * dir.y + ax.z * dir.z;
  r = mk3(uv * ax.x - dir.x, uv * ax.y - dir.y, uv * ax.z - dir.z);
  float A = dot(r, r);
  const float B = 2.f * dot(q, r);
  float k = B * B - 4.f * A * C;
  if (k < 1.1920929e-07f && C <= 0.f) {
    const float uvsign = (uv < 0.f) ? -1.0f : 1.0f;
    const float internal = (d >= -hl && d <= +hl) ? -1.0f : 1.0f;
    if ((uv > 0.f && d + uvsign * len < hl * internal) || (uv < 0.f && d + uvsign * len > hl * internal)) return 0;
    c->depth = (-uvsign * d) - (internal * hl);
    c->pos = mk3(ray.p.x + c->depth * dir.x, ray.p.y + c->depth * dir.y, ray.p.z + c->depth * dir.z);
    c->n = uvsign * ax;
    return 1;
  }
  if (k > 0.f) {
    k = sqrtf(k);
    A = 1.f / (2.f * A);
    float alpha = (-B - k) * A;
    if (alpha < 0.f) alpha = (-B + k) * A;
    if (alpha >= 0.f && alpha <= len) {
      c->pos = ray.p + alpha * dir;
      q = c->pos - cyl.p;
      d = q.x * ax.x + q.y * ax.y + q.z * ax.z;
      if (d >= -hl && d <= +hl) {
        const float nsign = (C < 0.f) ? -1.0f : 1.0f;
        c->n = safe_normalize(nsign * (c->pos - (cyl.p + d * ax)));
        c->depth = alpha;
        return 1;
      }
    }
  }
  return 0;
}

// the class that comes first in a collider's arguments (the colliders table,
// collision_kernel.cpp:181-219): ray, then capsule, cylinder, sphere, box,
// plane
__device__ inline int collider_rank(int type) {
  return type == GEOM_RAY ? 0
         : type == GEOM_CAPSULE ? 1
         : type == GEOM_CYLINDER ? 2
         : type == GEOM_SPHERE ? 3
         : type == GEOM_BOX ? 4 : 5;
}

__device__ inline int collide_pair(const GeomView &a, const GeomView &b, int maxc, CPoint *c) {
  if (maxc <= 0) return 0;
  bool rev = collider_rank(a.type) > collider_rank(b.type);
  const GeomView &x = rev ? b : a;
  const GeomView &y = rev ? a : b;
  int n = 0;
  if (x.type == GEOM_RAY) {
    if (y.type == GEOM_SPHERE) n = ray_sphere_hit(x, y.p, y.prm.x, 0, c);
    else if (y.type == GEOM_BOX) n = coll_ray_box(x, y, c);
    else if (y.type == GEOM_CAPSULE) n = coll_ray_capsule(x, y, c);
    else if (y.type == GEOM_PLANE) n = coll_ray_plane(x, y.prm, c);
    else if (y.type == GEOM_CYLINDER) n = coll_ray_cylinder(x, y, c);
  } else if (x.type == GEOM_CAPSULE) {
    if (y.type == GEOM_SPHERE) n = coll_capsule_sphere(x, y.p, y.prm.x, c);
    else if (y.type == GEOM_BOX) n = coll_capsule_box(x, y, c);
    else if (y.type == GEOM_CAPSULE) n = coll_capsule_capsule(x, y, maxc, c);
    else if (y.type == GEOM_PLANE) n = coll_capsule_plane(x, y.prm, maxc, c);
  } else if (x.type == GEOM_CYLINDER) {
    if (y.type == GEOM_SPHERE) n = coll_cylinder_sphere(x, y.p, y.prm.x, c);
    else if (y.type == GEOM_PLANE) n = coll_cylinder_plane(x, y.prm, maxc, c);
    else if (y.type == GEOM_BOX) n = coll_cylinder_box(x, y, maxc, c);
  } else if (x.type == GEOM_SPHERE) {
    if (y.type == GEOM_SPHERE) n = coll_sphere_sphere(x.p, x.prm.x, y.p, y.prm.x, c);
    else if (y.type == GEOM_BOX) n = coll_sphere_box(x.p, x.prm.x, y.p, y.R, mk3(y.prm.x, y.prm.y, y.prm.z), c);
    else n = coll_sphere_plane(x.p, x.prm.x, y.prm, c);
  } else if (x.type == GEOM_BOX) {
    if (y.type == GEOM_BOX)
      n = coll_box_box(x.p, x.R, mk3(x.prm.x, x.prm.y, x.prm.z), y.p, y.R, mk3(y.prm.x, y.prm.y, y.prm.z), maxc, c);
    else n = coll_box_plane(x.p, x.R, mk3(x.prm.x, x.prm.y, x.prm.z), y.prm, maxc, c);
  }
  if (rev)
    for (int i = 0; i < n; i++) c[i].n = -c[i].n;
  return n;
}

}  // namespace odemi
