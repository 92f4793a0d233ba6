// orc_collide.hpp -- CPU restatement of the narrowphase colliders on the path
// (sphere-sphere, sphere-box, sphere-plane, box-box, box-plane, the capsule,
// ray and cylinder colliders), the geom AABBs and the collider dispatch.
//
// TEST INFRASTRUCTURE (oracle).  Each function cites the reference code whose
// behaviour it restates.  The contact arrays hold up to `maxc` contacts of
// pos[3], normal[3], depth.
#pragma once
#include <cmath>
#include "orc_math.hpp"

namespace orc {

enum { G_SPHERE = 0, G_BOX = 1, G_PLANE = 2, G_CAPSULE = 3, G_RAY = 4, G_CYLINDER = 5 };
static const int CONTACTS_UNIMPORTANT = (int)0x80000000;

template <class R> struct ContactGeom {
  R pos[3], normal[3], depth;
  int g1, g2;
};

// geom pose and shape; for planes `param` is the unit normal and offset
template <class R> struct GeomShape {
  int type;
  const R *pos;   // 3 (4-stride not needed)
  const R *Rm;    // 3x4 rotation
  R param[4];
};

// World AABB = (minx,maxx,miny,maxy,minz,maxz).
// box: ode/src/box.cpp:60-77, sphere: sphere.cpp:59-67, plane: plane.cpp:80-107
template <class R> void geom_aabb(const GeomShape<R> &g, R *aabb) {
  const R inf = INFINITY;
  if (g.type == G_SPHERE) {
    R r = g.param[0];
    for (int k = 0; k < 3; k++) { aabb[2 * k] = g.pos[k] - r; aabb[2 * k + 1] = g.pos[k] + r; }
  } else if (g.type == G_BOX) {
    const R *M = g.Rm;
    const R *s = g.param;
    for (int k = 0; k < 3; k++) {
      R ext = R(0.5) * (std::fabs(M[4 * k] * s[0]) + std::fabs(M[4 * k + 1] * s[1]) + std::fabs(M[4 * k + 2] * s[2]));
      aabb[2 * k] = g.pos[k] - ext;
      aabb[2 * k + 1] = g.pos[k] + ext;
    }
  } else if (g.type == G_CAPSULE) {
    // capsule.cpp:60-74: half the axis (column 2 of R) times the length,
    // plus the radius
    const R *M = g.Rm;
    for (int k = 0; k < 3; k++) {
      R ext = std::fabs(M[4 * k + 2] * g.param[1]) * R(0.5) + g.param[0];
      aabb[2 * k] = g.pos[k] - ext;
      aabb[2 * k + 1] = g.pos[k] + ext;
    }
  } else if (g.type == G_CYLINDER) {
    // cylinder.cpp:58-74: the radius along columns 0 and 1 of R, half the
    // length along column 2
    const R *M = g.Rm;
    for (int k = 0; k < 3; k++) {
      R ext = std::fabs(M[4 * k] * g.param[0]) + std::fabs(M[4 * k + 1] * g.param[0]) +
              R(0.5) * std::fabs(M[4 * k + 2] * g.param[1]);
      aabb[2 * k] = g.pos[k] - ext;
      aabb[2 * k + 1] = g.pos[k] + ext;
    }
  } else if (g.type == G_RAY) {
    // ray.cpp:56-88: the box of the start point and the end point
    for (int k = 0; k < 3; k++) {
      R e = g.pos[k] + g.Rm[4 * k + 2] * g.param[0];
      if (g.pos[k] < e) { aabb[2 * k] = g.pos[k]; aabb[2 * k + 1] = e; }
      else { aabb[2 * k] = e; aabb[2 * k + 1] = g.pos[k]; }
    }
  } else {
    const R *p = g.param;
    for (int k = 0; k < 3; k++) { aabb[2 * k] = -inf; aabb[2 * k + 1] = inf; }
    // an axis-aligned plane bounds one half space (plane.cpp:94-106)
    int ax = -1;
    if (p[1] == 0 && p[2] == 0) ax = 0;
    else if (p[0] == 0 && p[2] == 0) ax = 1;
    else if (p[0] == 0 && p[1] == 0) ax = 2;
    if (ax >= 0) {
      aabb[2 * ax] = (p[ax] > 0) ? -inf : -p[3];
      aabb[2 * ax + 1] = (p[ax] > 0) ? p[3] : inf;
    }
  }
}

// sphere-sphere: ode/src/collision_util.cpp:38-67 (via sphere.cpp:110-128)
template <class R>
int collide_sphere_sphere(const R *p1, R r1, const R *p2, R r2, ContactGeom<R> *c) {
  R d3[3] = {p1[0] - p2[0], p1[1] - p2[1], p1[2] - p2[2]};
  R d = std::sqrt(dot3(d3, d3));
  if (d > r1 + r2) return 0;
  if (d <= 0) {
    for (int k = 0; k < 3; k++) c->pos[k] = p1[k];
    c->normal[0] = 1; c->normal[1] = 0; c->normal[2] = 0;
    c->depth = r1 + r2;
  } else {
    R inv = recip(d);
    for (int k = 0; k < 3; k++) c->normal[k] = (p1[k] - p2[k]) * inv;
    R k2 = R(0.5) * (r2 - r1 - d);
    for (int k = 0; k < 3; k++) c->pos[k] = p1[k] + c->normal[k] * k2;
    c->depth = r1 + r2 - d;
  }
  return 1;
}

// sphere-box: ode/src/sphere.cpp:131-219.  Clamp the sphere centre into the
// box frame; inside -> push out through the nearest face, else use the
// closest surface point.
template <class R>
int collide_sphere_box(const R *ps, R rad, const R *pb, const R *Rb, const R *side, ContactGeom<R> *c) {
  R p[3] = {ps[0] - pb[0], ps[1] - pb[1], ps[2] - pb[2]};
  R l[3], t[4];
  bool border = false;
  for (int k = 0; k < 3; k++) {
    l[k] = side[k] * R(0.5);
    t[k] = dot3(p, Rb + k, 1, 4);
    if (t[k] < -l[k]) { t[k] = -l[k]; border = true; }
    if (t[k] > l[k]) { t[k] = l[k]; border = true; }
  }
  if (!border) {
    R best = l[0] - std::fabs(t[0]);
    int bi = 0;
    for (int k = 1; k < 3; k++) {
      R fd = l[k] - std::fabs(t[k]);
      if (fd < best) { best = fd; bi = k; }
    }
    for (int k = 0; k < 3; k++) c->pos[k] = ps[k];
    R axis[3] = {0, 0, 0};
    axis[bi] = (t[bi] > 0) ? R(1) : R(-1);
    mv(c->normal, Rb, axis);
    c->depth = best + rad;
    return 1;
  }
  t[3] = 0;
  R q[3], r[3];
  mv(q, Rb, t);
  for (int k = 0; k < 3; k++) r[k] = p[k] - q[k];
  R depth = rad - std::sqrt(dot3(r, r));
  if (depth < 0) return 0;
  for (int k = 0; k < 3; k++) c->pos[k] = q[k] + pb[k];
  R n[4] = {r[0], r[1], r[2], 0};
  normalize3(n);
  for (int k = 0; k < 3; k++) c->normal[k] = n[k];
  c->depth = depth;
  return 1;
}

// sphere-plane: ode/src/sphere.cpp:222-251
template <class R>
int collide_sphere_plane(const R *ps, R rad, const R *plane, ContactGeom<R> *c) {
  R k = dot3(ps, plane);
  R depth = plane[3] - k + rad;
  if (depth < 0) return 0;
  for (int i = 0; i < 3; i++) {
    c->normal[i] = plane[i];
    c->pos[i] = ps[i] - plane[i] * rad;
  }
  c->depth = depth;
  return 1;
}

// box-plane: ode/src/box.cpp:745-878.  Deepest vertex first, then vertices
// reached along the box edges with the smallest projections onto the plane
// normal, and (for 4 requested contacts) the fourth vertex of that face.
template <class R>
int collide_box_plane(const R *pb, const R *Rb, const R *side, const R *plane, int flags, ContactGeom<R> *c) {
  const R *n = plane;
  R Q[3], A[3], B[3];
  for (int k = 0; k < 3; k++) {
    Q[k] = dot3(n, Rb + k, 1, 4);
    A[k] = side[k] * Q[k];
    B[k] = std::fabs(A[k]);
  }
  R depth = plane[3] + R(0.5) * (B[0] + B[1] + B[2]) - dot3(n, pb);
  if (depth < 0) return 0;
  int maxc = flags & 0xffff;
  if (maxc > 4) maxc = 4;
  R p[3] = {pb[0], pb[1], pb[2]};
  for (int k = 0; k < 3; k++) {
    R h = R(0.5) * side[k];
    if (A[k] > 0) for (int i = 0; i < 3; i++) p[i] -= h * Rb[4 * i + k];
    else for (int i = 0; i < 3; i++) p[i] += h * Rb[4 * i + k];
  }
  for (int i = 0; i < 3; i++) c[0].pos[i] = p[i];
  c[0].depth = depth;
  int ret = 1;
  // the axis order used for contacts 2 and 3 (box.cpp:835-851)
  int first, second;
  if (B[0] < B[1]) {
    if (B[2] < B[0]) { first = 2; second = (B[0] < B[1]) ? 0 : 1; }
    else { first = 0; second = (B[1] < B[2]) ? 1 : 2; }
  } else {
    if (B[2] < B[1]) { first = 2; second = (B[0] < B[1]) ? 0 : 1; }
    else { first = 1; second = (B[0] < B[2]) ? 0 : 2; }
  }
  auto add = [&](int slot, int ax) -> bool {
    if (depth - B[ax] < 0) return false;
    for (int i = 0; i < 3; i++) c[slot].pos[i] = (A[ax] > 0) ? p[i] + side[ax] * Rb[4 * i + ax] : p[i] - side[ax] * Rb[4 * i + ax];
    c[slot].depth = depth - B[ax];
    ret++;
    return true;
  };
  if (maxc > 1) {
    if (add(1, first) && maxc > 2) add(2, second);
  }
  if (maxc == 4 && ret == 3) {
    R d4 = c[1].depth + c[2].depth - depth;
    if (d4 > 0) {
      for (int i = 0; i < 3; i++) c[3].pos[i] = c[1].pos[i] + c[2].pos[i] - p[i];
      c[3].depth = d4;
      ret++;
    }
  }
  for (int j = 0; j < ret; j++)
    for (int i = 0; i < 3; i++) c[j].normal[i] = n[i];
  return ret;
}

// closest approach of two lines: ode/src/collision_util.cpp:70-92
template <class R>
void line_closest_approach(const R *pa, const R *ua, const R *pb, const R *ub, R *alpha, R *beta) {
  R p[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
  R uaub = dot3(ua, ub);
  R q1 = dot3(ua, p);
  R q2 = -dot3(ub, p);
  R d = 1 - uaub * uaub;
  if (d <= R(0.0001)) { *alpha = 0; *beta = 0; return; }
  d = recip(d);
  *alpha = (q1 + uaub * q2) * d;
  *beta = (uaub * q1 + q2) * d;
}

// Clip the quadrilateral q (4 points) against the rectangle |x|<h0,|y|<h1.
// ode/src/box.cpp:187-238: four half-plane chops alternating between the
// output array and a scratch buffer; the whole clip stops as soon as a chop
// has produced 8 points.  Returns the number of points (<= 8) in ret.
template <class R> int clip_quad_rect(const R *h, const R *quad, R *ret) {
  R scratch[16], in[8];
  std::memcpy(in, quad, sizeof(in));
  const R *q = in;
  R *r = ret;
  int nq = 4, nr = 0;
  for (int dir = 0; dir <= 1; dir++) {
    for (int sign = -1; sign <= 1; sign += 2) {
      R *pr = r;
      nr = 0;
      for (int i = 0; i < nq; i++) {
        const R *a = q + 2 * i;
        const R *b = (i + 1 < nq) ? a + 2 : q;
        bool ain = sign * a[dir] < h[dir];
        if (ain) {
          pr[0] = a[0];
          pr[1] = a[1];
          pr += 2;
          if (++nr & 8) { q = r; goto done; }
        }
        if (ain ^ (sign * b[dir] < h[dir])) {
          pr[1 - dir] = a[1 - dir] + (b[1 - dir] - a[1 - dir]) / (b[dir] - a[dir]) * (sign * h[dir] - a[dir]);
          pr[dir] = sign * h[dir];
          pr += 2;
          if (++nr & 8) { q = r; goto done; }
        }
      }
      q = r;
      r = (q == ret) ? scratch : ret;
      nq = nr;
    }
  }
done:
  if (q != ret) std::memcpy(ret, q, sizeof(R) * 2 * nr);
  return nr;
}

// Choose m of n (<=8) polygon points spread around the centroid, starting
// with point i0.  ode/src/box.cpp:249-311.
template <class R> void cull_points(int n, const R *p, int m, int i0, int *iret) {
  R cx, cy;
  if (n == 1) { cx = p[0]; cy = p[1]; }
  else if (n == 2) { cx = R(0.5) * (p[0] + p[2]); cy = R(0.5) * (p[1] + p[3]); }
  else {
    R a = 0, q;
    cx = 0; cy = 0;
    for (int i = 0; i < n - 1; i++) {
      q = p[2 * i] * p[2 * i + 3] - p[2 * i + 2] * p[2 * i + 1];
      a += q;
      cx += q * (p[2 * i] + p[2 * i + 2]);
      cy += q * (p[2 * i + 1] + p[2 * i + 3]);
    }
    q = p[2 * n - 2] * p[1] - p[0] * p[2 * n - 1];
    a = recip(R(3.0) * (a + q));
    cx = a * (cx + q * (p[2 * n - 2] + p[0]));
    cy = a * (cy + q * (p[2 * n - 1] + p[1]));
  }
  R ang[8];
  for (int i = 0; i < n; i++) ang[i] = std::atan2(p[2 * i + 1] - cy, p[2 * i] - cx);
  int avail[8];
  for (int i = 0; i < n; i++) avail[i] = 1;
  avail[i0] = 0;
  iret[0] = i0;
  const double PI = 3.14159265358979323846;
  for (int j = 1; j < m; j++) {
    R a = (R)(R(j) * (2 * PI / m) + ang[i0]);
    if (a > PI) a -= (R)(2 * PI);
    R best = 1e9;
    int bi = i0;
    for (int i = 0; i < n; i++) {
      if (!avail[i]) continue;
      R diff = std::fabs(ang[i] - a);
      if (diff > PI) diff = (R)(2 * PI - diff);
      if (diff < best) { best = diff; bi = i; }
    }
    iret[j] = bi;
    avail[bi] = 0;
  }
}

// box-box separating-axis test and contact generation:
// ode/src/box.cpp:331-712 (dBoxBox) and 716-742 (dCollideBoxBox, which
// negates the normal so that it points into box 1).
template <class R>
int collide_box_box(const R *p1, const R *R1, const R *side1, const R *p2, const R *R2, const R *side2,
                    int flags, ContactGeom<R> *out) {
  const R fudge = R(1.05);
  R p[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  R pp[3];
  mtv(pp, R1, p);
  R A[3] = {side1[0] * R(0.5), side1[1] * R(0.5), side1[2] * R(0.5)};
  R B[3] = {side2[0] * R(0.5), side2[1] * R(0.5), side2[2] * R(0.5)};
  R Rr[3][3], Q[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      Rr[i][j] = dot3(R1 + i, R2 + j, 4, 4);
      Q[i][j] = std::fabs(Rr[i][j]);
    }
  R s = -INFINITY;
  bool invert = false;
  int code = 0;
  const R *normalR = nullptr;  // column of R1 or R2
  R normalC[3] = {0, 0, 0};     // relative to box 1
  bool unimportant = (flags & CONTACTS_UNIMPORTANT) != 0;
  bool done = false;
  auto face_test = [&](R e1, R e2, const R *col, int cc) -> int {
    R s2 = std::fabs(e1) - e2;
    if (s2 > 0) return -1;
    if (s2 > s) {
      s = s2; normalR = col; invert = e1 < 0; code = cc;
      if (unimportant) return 1;
    }
    return 0;
  };
  auto edge_test = [&](R e1, R e2, R n1, R n2, R n3, int cc) -> int {
    R s2 = std::fabs(e1) - e2;
    if (s2 > 0) return -1;
    R l = std::sqrt(n1 * n1 + n2 * n2 + n3 * n3);
    if (l > 0) {
      s2 /= l;
      if (s2 * fudge > s) {
        s = s2; normalR = nullptr;
        normalC[0] = n1 / l; normalC[1] = n2 / l; normalC[2] = n3 / l;
        invert = e1 < 0; code = cc;
        if (unimportant) return 1;
      }
    }
    return 0;
  };
#define ORC_T(x) { int r_ = (x); if (r_ < 0) return 0; if (r_ > 0) { done = true; } }
  do {
    ORC_T(face_test(pp[0], A[0] + B[0] * Q[0][0] + B[1] * Q[0][1] + B[2] * Q[0][2], R1 + 0, 1)); if (done) break;
    ORC_T(face_test(pp[1], A[1] + B[0] * Q[1][0] + B[1] * Q[1][1] + B[2] * Q[1][2], R1 + 1, 2)); if (done) break;
    ORC_T(face_test(pp[2], A[2] + B[0] * Q[2][0] + B[1] * Q[2][1] + B[2] * Q[2][2], R1 + 2, 3)); if (done) break;
    ORC_T(face_test(dot3(R2 + 0, p, 4, 1), A[0] * Q[0][0] + A[1] * Q[1][0] + A[2] * Q[2][0] + B[0], R2 + 0, 4)); if (done) break;
    ORC_T(face_test(dot3(R2 + 1, p, 4, 1), A[0] * Q[0][1] + A[1] * Q[1][1] + A[2] * Q[2][1] + B[1], R2 + 1, 5)); if (done) break;
    ORC_T(face_test(dot3(R2 + 2, p, 4, 1), A[0] * Q[0][2] + A[1] * Q[1][2] + A[2] * Q[2][2] + B[2], R2 + 2, 6)); if (done) break;
    const R(&M)[3][3] = Rr;
    ORC_T(edge_test(pp[2] * M[1][0] - pp[1] * M[2][0], A[1] * Q[2][0] + A[2] * Q[1][0] + B[1] * Q[0][2] + B[2] * Q[0][1], 0, -M[2][0], M[1][0], 7)); if (done) break;
    ORC_T(edge_test(pp[2] * M[1][1] - pp[1] * M[2][1], A[1] * Q[2][1] + A[2] * Q[1][1] + B[0] * Q[0][2] + B[2] * Q[0][0], 0, -M[2][1], M[1][1], 8)); if (done) break;
    ORC_T(edge_test(pp[2] * M[1][2] - pp[1] * M[2][2], A[1] * Q[2][2] + A[2] * Q[1][2] + B[0] * Q[0][1] + B[1] * Q[0][0], 0, -M[2][2], M[1][2], 9)); if (done) break;
    ORC_T(edge_test(pp[0] * M[2][0] - pp[2] * M[0][0], A[0] * Q[2][0] + A[2] * Q[0][0] + B[1] * Q[1][2] + B[2] * Q[1][1], M[2][0], 0, -M[0][0], 10)); if (done) break;
    ORC_T(edge_test(pp[0] * M[2][1] - pp[2] * M[0][1], A[0] * Q[2][1] + A[2] * Q[0][1] + B[0] * Q[1][2] + B[2] * Q[1][0], M[2][1], 0, -M[0][1], 11)); if (done) break;
    ORC_T(edge_test(pp[0] * M[2][2] - pp[2] * M[0][2], A[0] * Q[2][2] + A[2] * Q[0][2] + B[0] * Q[1][1] + B[1] * Q[1][0], M[2][2], 0, -M[0][2], 12)); if (done) break;
    ORC_T(edge_test(pp[1] * M[0][0] - pp[0] * M[1][0], A[0] * Q[1][0] + A[1] * Q[0][0] + B[1] * Q[2][2] + B[2] * Q[2][1], -M[1][0], M[0][0], 0, 13)); if (done) break;
    ORC_T(edge_test(pp[1] * M[0][1] - pp[0] * M[1][1], A[0] * Q[1][1] + A[1] * Q[0][1] + B[0] * Q[2][2] + B[2] * Q[2][0], -M[1][1], M[0][1], 0, 14)); if (done) break;
    ORC_T(edge_test(pp[1] * M[0][2] - pp[0] * M[1][2], A[0] * Q[1][2] + A[1] * Q[0][2] + B[0] * Q[2][1] + B[1] * Q[2][0], -M[1][2], M[0][2], 0, 15)); if (done) break;
  } while (0);
#undef ORC_T
  if (!code) return 0;
  R normal[3];
  if (normalR) { normal[0] = normalR[0]; normal[1] = normalR[4]; normal[2] = normalR[8]; }
  else mv(normal, R1, normalC);
  if (invert) for (int k = 0; k < 3; k++) normal[k] = -normal[k];
  R depth = -s;
  int n_out = 0;
  if (code > 6) {
    // edge-edge: one contact halfway between the closest points of the two
    // edges (box.cpp:465-503)
    R pa[3] = {p1[0], p1[1], p1[2]};
    for (int j = 0; j < 3; j++) {
      R sg = (dot3(normal, R1 + j, 1, 4) > 0) ? R(1) : R(-1);
      for (int i = 0; i < 3; i++) pa[i] += sg * A[j] * R1[4 * i + j];
    }
    R pb[3] = {p2[0], p2[1], p2[2]};
    for (int j = 0; j < 3; j++) {
      R sg = (dot3(normal, R2 + j, 1, 4) > 0) ? R(-1) : R(1);
      for (int i = 0; i < 3; i++) pb[i] += sg * B[j] * R2[4 * i + j];
    }
    R ua[3], ub[3], al, be;
    for (int i = 0; i < 3; i++) ua[i] = R1[(code - 7) / 3 + 4 * i];
    for (int i = 0; i < 3; i++) ub[i] = R2[(code - 7) % 3 + 4 * i];
    line_closest_approach(pa, ua, pb, ub, &al, &be);
    for (int i = 0; i < 3; i++) pa[i] += ua[i] * al;
    for (int i = 0; i < 3; i++) pb[i] += ub[i] * be;
    for (int i = 0; i < 3; i++) out[0].pos[i] = R(0.5) * (pa[i] + pb[i]);
    out[0].depth = depth;
    n_out = 1;
  } else {
    // face-something: clip the incident face against the reference face
    // (box.cpp:505-708)
    const R *Ra, *Rb, *pa, *pb, *Sa, *Sb;
    if (code <= 3) { Ra = R1; Rb = R2; pa = p1; pb = p2; Sa = A; Sb = B; }
    else { Ra = R2; Rb = R1; pa = p2; pb = p1; Sa = B; Sb = A; }
    R n2[3], nr[3], anr[3];
    for (int k = 0; k < 3; k++) n2[k] = (code <= 3) ? normal[k] : -normal[k];
    mtv(nr, Rb, n2);
    for (int k = 0; k < 3; k++) anr[k] = std::fabs(nr[k]);
    int lanr, a1, a2;
    if (anr[1] > anr[0]) {
      if (anr[1] > anr[2]) { a1 = 0; lanr = 1; a2 = 2; }
      else { a1 = 0; a2 = 1; lanr = 2; }
    } else {
      if (anr[0] > anr[2]) { lanr = 0; a1 = 1; a2 = 2; }
      else { a1 = 0; a2 = 1; lanr = 2; }
    }
    R center[3];
    for (int i = 0; i < 3; i++)
      center[i] = (nr[lanr] < 0) ? pb[i] - pa[i] + Sb[lanr] * Rb[4 * i + lanr] : pb[i] - pa[i] - Sb[lanr] * Rb[4 * i + lanr];
    int codeN = (code <= 3) ? code - 1 : code - 4;
    int code1 = (codeN == 0) ? 1 : 0;
    int code2 = (codeN == 2) ? 1 : 2;
    R c1 = dot3(center, Ra + code1, 1, 4);
    R c2 = dot3(center, Ra + code2, 1, 4);
    R m11 = dot3(Ra + code1, Rb + a1, 4, 4);
    R m12 = dot3(Ra + code1, Rb + a2, 4, 4);
    R m21 = dot3(Ra + code2, Rb + a1, 4, 4);
    R m22 = dot3(Ra + code2, Rb + a2, 4, 4);
    R quad[8];
    {
      R k1 = m11 * Sb[a1], k2 = m21 * Sb[a1], k3 = m12 * Sb[a2], k4 = m22 * Sb[a2];
      quad[0] = c1 - k1 - k3; quad[1] = c2 - k2 - k4;
      quad[2] = c1 - k1 + k3; quad[3] = c2 - k2 + k4;
      quad[4] = c1 + k1 + k3; quad[5] = c2 + k2 + k4;
      quad[6] = c1 + k1 - k3; quad[7] = c2 + k2 - k4;
    }
    R rect[2] = {Sa[code1], Sa[code2]};
    R ret[16];
    int n = clip_quad_rect(rect, quad, ret);
    if (n < 1) return 0;
    R point[24], dep[8];
    R det1 = recip(m11 * m22 - m12 * m21);
    m11 *= det1; m12 *= det1; m21 *= det1; m22 *= det1;
    int cnum = 0;
    for (int j = 0; j < n; j++) {
      R k1 = m22 * (ret[2 * j] - c1) - m12 * (ret[2 * j + 1] - c2);
      R k2 = -m21 * (ret[2 * j] - c1) + m11 * (ret[2 * j + 1] - c2);
      for (int i = 0; i < 3; i++) point[3 * cnum + i] = center[i] + k1 * Rb[4 * i + a1] + k2 * Rb[4 * i + a2];
      dep[cnum] = Sa[codeN] - dot3(n2, point + 3 * cnum);
      if (dep[cnum] >= 0) {
        ret[2 * cnum] = ret[2 * j];
        ret[2 * cnum + 1] = ret[2 * j + 1];
        cnum++;
        if ((cnum | CONTACTS_UNIMPORTANT) == (flags & (0xffff | CONTACTS_UNIMPORTANT))) break;
      }
    }
    if (cnum < 1) return 0;
    int maxc = flags & 0xffff;
    if (maxc > cnum) maxc = cnum;
    if (maxc < 1) maxc = 1;
    if (cnum <= maxc) {
      for (int j = 0; j < cnum; j++) {
        for (int i = 0; i < 3; i++) out[j].pos[i] = point[3 * j + i] + pa[i];
        out[j].depth = dep[j];
      }
      n_out = cnum;
    } else {
      int i1 = 0;
      R md = dep[0];
      for (int i = 1; i < cnum; i++)
        if (dep[i] > md) { md = dep[i]; i1 = i; }
      int iret[8];
      cull_points(cnum, ret, maxc, i1, iret);
      for (int j = 0; j < maxc; j++) {
        for (int i = 0; i < 3; i++) out[j].pos[i] = point[3 * iret[j] + i] + pa[i];
        out[j].depth = dep[iret[j]];
      }
      n_out = maxc;
    }
  }
  for (int j = 0; j < n_out; j++)
    for (int k = 0; k < 3; k++) out[j].normal[k] = -normal[k];
  return n_out;
}

// dCollide dispatch (ode/src/collision_kernel.cpp:292-339 with the table of
// 167-190): the sphere-box, sphere-plane and box-plane colliders take the
// sphere/box first; a reversed call swaps the arguments, negates the normals
// and swaps g1/g2.  g1/g2 are written as the caller's ids.
// ------------------------------------------------------------ capsules
// A capsule (param: radius, length) lies along column 2 of its rotation.

// dClosestLineSegmentPoints (collision_util.cpp:109-240): the closest points
// of segments a1-a2 and b1-b2 by the Voronoi clipping of their vertex and
// edge features, in the reference's order of tests
template <class R>
void closest_segment_points(const R *a1, const R *a2, const R *b1, const R *b2, R *cp1, R *cp2) {
  R a1a2[3], b1b2[3], a1b1[3], a1b2[3], a2b1[3], a2b2[3], n[3];
  auto sub = [](R *o, const R *x, const R *y) { for (int k = 0; k < 3; k++) o[k] = x[k] - y[k]; };
  auto set = [](R *o, const R *x) { for (int k = 0; k < 3; k++) o[k] = x[k]; };
  sub(a1a2, a2, a1);
  sub(b1b2, b2, b1);
  sub(a1b1, b1, a1);
  R da1 = dot3(a1a2, a1b1), db1 = dot3(b1b2, a1b1);
  if (da1 <= 0 && db1 >= 0) { set(cp1, a1); set(cp2, b1); return; }
  sub(a1b2, b2, a1);
  R da2 = dot3(a1a2, a1b2), db2 = dot3(b1b2, a1b2);
  if (da2 <= 0 && db2 <= 0) { set(cp1, a1); set(cp2, b2); return; }
  sub(a2b1, b1, a2);
  R da3 = dot3(a1a2, a2b1), db3 = dot3(b1b2, a2b1);
  if (da3 >= 0 && db3 >= 0) { set(cp1, a2); set(cp2, b1); return; }
  sub(a2b2, b2, a2);
  R da4 = dot3(a1a2, a2b2), db4 = dot3(b1b2, a2b2);
  if (da4 >= 0 && db4 <= 0) { set(cp1, a2); set(cp2, b2); return; }
  R la = dot3(a1a2, a1a2);
  if (da1 >= 0 && da3 <= 0) {
    R k = da1 / la;
    for (int i = 0; i < 3; i++) n[i] = a1b1[i] - k * a1a2[i];
    if (dot3(b1b2, n) >= 0) {
      for (int i = 0; i < 3; i++) cp1[i] = a1[i] + k * a1a2[i];
      set(cp2, b1);
      return;
    }
  }
  if (da2 >= 0 && da4 <= 0) {
    R k = da2 / la;
    for (int i = 0; i < 3; i++) n[i] = a1b2[i] - k * a1a2[i];
    if (dot3(b1b2, n) <= 0) {
      for (int i = 0; i < 3; i++) cp1[i] = a1[i] + k * a1a2[i];
      set(cp2, b2);
      return;
    }
  }
  R lb = dot3(b1b2, b1b2);
  if (db1 <= 0 && db2 >= 0) {
    R k = -db1 / lb;
    for (int i = 0; i < 3; i++) n[i] = -a1b1[i] - k * b1b2[i];
    if (dot3(a1a2, n) >= 0) {
      set(cp1, a1);
      for (int i = 0; i < 3; i++) cp2[i] = b1[i] + k * b1b2[i];
      return;
    }
  }
  if (db3 <= 0 && db4 >= 0) {
    R k = -db3 / lb;
    for (int i = 0; i < 3; i++) n[i] = -a2b1[i] - k * b1b2[i];
    if (dot3(a1a2, n) <= 0) {
      set(cp1, a2);
      for (int i = 0; i < 3; i++) cp2[i] = b1[i] + k * b1b2[i];
      return;
    }
  }
  R k = dot3(a1a2, b1b2);
  R det = la * lb - k * k;
  if (det <= 0) { set(cp1, a1); set(cp2, b1); return; }
  det = recip(det);
  R alpha = (lb * da1 - k * db1) * det;
  R beta = (k * da1 - la * db1) * det;
  for (int i = 0; i < 3; i++) cp1[i] = a1[i] + alpha * a1a2[i];
  for (int i = 0; i < 3; i++) cp2[i] = b1[i] + beta * b1b2[i];
}

// dClosestLineBoxPoints (collision_util.cpp:247-375): the point of segment
// p1-p2 closest to the box (c, Rb, side) and the box point closest to it, by
// walking the piecewise-linear derivative of the squared distance along the
// segment in box coordinates
template <class R>
void closest_line_box_points(const R *p1, const R *p2, const R *c, const R *Rb, const R *side, R *lret, R *bret) {
  R tmp[3], s[3], v[3], sign[3], v2[3], h[3], tanchor[3];
  int region[3];
  for (int i = 0; i < 3; i++) tmp[i] = p1[i] - c[i];
  mtv(s, Rb, tmp);
  for (int i = 0; i < 3; i++) tmp[i] = p2[i] - p1[i];
  mtv(v, Rb, tmp);
  for (int i = 0; i < 3; i++) {
    if (v[i] < 0) { s[i] = -s[i]; v[i] = -v[i]; sign[i] = -1; }
    else sign[i] = 1;
  }
  for (int i = 0; i < 3; i++) v2[i] = v[i] * v[i];
  for (int i = 0; i < 3; i++) h[i] = R(0.5) * side[i];
  const R eps = sizeof(R) == 4 ? R(1e-19) : R(1e-307);
  for (int i = 0; i < 3; i++) {
    if (v[i] > eps) {
      if (s[i] < -h[i]) { region[i] = -1; tanchor[i] = (-h[i] - s[i]) / v[i]; }
      else { region[i] = (s[i] > h[i]); tanchor[i] = (h[i] - s[i]) / v[i]; }
    } else {
      region[i] = 0;
      tanchor[i] = 2;
    }
  }
  R t = 0, dd2dt = 0;
  for (int i = 0; i < 3; i++) dd2dt -= (region[i] ? v2[i] : R(0)) * tanchor[i];
  bool done = dd2dt >= 0;
  while (!done) {
    R next_t = 1;
    for (int i = 0; i < 3; i++)
      if (tanchor[i] > t && tanchor[i] < 1 && tanchor[i] < next_t) next_t = tanchor[i];
    R next_dd2dt = 0;
    for (int i = 0; i < 3; i++) next_dd2dt += (region[i] ? v2[i] : R(0)) * (next_t - tanchor[i]);
    if (next_dd2dt >= 0) {
      R m = (next_dd2dt - dd2dt) / (next_t - t);
      t -= dd2dt / m;
      done = true;
      break;
    }
    for (int i = 0; i < 3; i++)
      if (tanchor[i] == next_t) { tanchor[i] = (h[i] - s[i]) / v[i]; region[i]++; }
    t = next_t;
    dd2dt = next_dd2dt;
    if (!(t < 1)) { t = 1; break; }
  }
  for (int i = 0; i < 3; i++) lret[i] = p1[i] + t * tmp[i];
  for (int i = 0; i < 3; i++) {
    tmp[i] = sign[i] * (s[i] + t * v[i]);
    if (tmp[i] < -h[i]) tmp[i] = -h[i];
    else if (tmp[i] > h[i]) tmp[i] = h[i];
  }
  mv(s, Rb, tmp);
  for (int i = 0; i < 3; i++) bret[i] = s[i] + c[i];
}

// capsule-sphere: capsule.cpp:130-162 (the sphere against the closest point
// of the capsule's axis segment)
template <class R>
int collide_capsule_sphere(const R *pc, const R *Rc, R rad, R len, const R *ps, R rs, ContactGeom<R> *c) {
  R alpha = Rc[2] * (ps[0] - pc[0]) + Rc[6] * (ps[1] - pc[1]) + Rc[10] * (ps[2] - pc[2]);
  R lz2 = len * R(0.5);
  if (alpha > lz2) alpha = lz2;
  if (alpha < -lz2) alpha = -lz2;
  R p[3] = {pc[0] + alpha * Rc[2], pc[1] + alpha * Rc[6], pc[2] + alpha * Rc[10]};
  return collide_sphere_sphere(p, rad, ps, rs, c);
}

// capsule-box: capsule.cpp:178-231
template <class R>
int collide_capsule_box(const R *pc, const R *Rc, R rad, R len, const R *pb, const R *Rb, const R *side,
                        ContactGeom<R> *c) {
  R clen = len * R(0.5);
  R p1[3] = {pc[0] + clen * Rc[2], pc[1] + clen * Rc[6], pc[2] + clen * Rc[10]};
  R p2[3] = {pc[0] - clen * Rc[2], pc[1] - clen * Rc[6], pc[2] - clen * Rc[10]};
  R pl[3], pbx[3];
  closest_line_box_points(p1, p2, pb, Rb, side, pl, pbx);
  const R mindist = sizeof(R) == 4 ? R(1e-9) : R(1e-18);
  R d[3] = {pl[0] - pbx[0], pl[1] - pbx[1], pl[2] - pbx[2]};
  if (std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) < mindist) {
    // the axis point is inside the box: the normal points from the box
    // centre (dCollideSpheresZeroDist, capsule.cpp:165-176)
    R n[3] = {pbx[0] - pb[0], pbx[1] - pb[1], pbx[2] - pb[2]};
    normalize3(n);
    for (int k = 0; k < 3; k++) c->normal[k] = n[k];
    c->depth = rad + R(0);
    R k2 = R(0.5) * (R(0) - rad);
    for (int k = 0; k < 3; k++) c->pos[k] = pl[k] + c->normal[k] * k2;
    return 1;
  }
  return collide_sphere_sphere(pl, rad, pbx, R(0), c);
}

// capsule-capsule: capsule.cpp:234-349 (up to two contacts along nearly
// parallel axes, else the closest points of the axis segments)
template <class R>
int collide_capsule_capsule(const R *pos1, const R *R1, R rad1, R len1, const R *pos2, const R *R2, R rad2, R len2,
                            int flags, ContactGeom<R> *c) {
  const R tolerance = R(1e-5);
  R lz1 = len1 * R(0.5), lz2 = len2 * R(0.5);
  R axis1[3] = {R1[2], R1[6], R1[10]}, axis2[3] = {R2[2], R2[6], R2[10]};
  R sphere1[3], sphere2[3];
  R a1a2 = dot3(axis1, axis2);
  R det = R(1.0) - a1a2 * a1a2;
  if (det < tolerance) {
    if (a1a2 < 0) for (int i = 0; i < 3; i++) axis2[i] = -axis2[i];
    R q[3];
    for (int i = 0; i < 3; i++) q[i] = pos1[i] - pos2[i];
    R k = dot3(axis1, q);
    R a1lo = -lz1, a1hi = lz1, a2lo = -lz2 - k, a2hi = lz2 - k;
    R lo = (a1lo > a2lo) ? a1lo : a2lo;
    R hi = (a1hi < a2hi) ? a1hi : a2hi;
    if (lo <= hi) {
      int num = flags & 0xffff;
      if (num >= 2 && lo < hi) {
        for (int i = 0; i < 3; i++) sphere1[i] = pos1[i] + lo * axis1[i];
        for (int i = 0; i < 3; i++) sphere2[i] = pos2[i] + (lo + k) * axis2[i];
        int n1 = collide_sphere_sphere(sphere1, rad1, sphere2, rad2, c);
        if (n1) {
          for (int i = 0; i < 3; i++) sphere1[i] = pos1[i] + hi * axis1[i];
          for (int i = 0; i < 3; i++) sphere2[i] = pos2[i] + (hi + k) * axis2[i];
          int n2 = collide_sphere_sphere(sphere1, rad1, sphere2, rad2, c + 1);
          if (n2) return 2;
        }
      }
      R alpha1 = (lo + hi) * R(0.5);
      R alpha2 = alpha1 + k;
      for (int i = 0; i < 3; i++) sphere1[i] = pos1[i] + alpha1 * axis1[i];
      for (int i = 0; i < 3; i++) sphere2[i] = pos2[i] + alpha2 * axis2[i];
      return collide_sphere_sphere(sphere1, rad1, sphere2, rad2, c);
    }
  }
  R a1[3], a2[3], b1[3], b2[3];
  for (int i = 0; i < 3; i++) {
    a1[i] = pos1[i] + axis1[i] * lz1;
    a2[i] = pos1[i] - axis1[i] * lz1;
    b1[i] = pos2[i] + axis2[i] * lz2;
    b2[i] = pos2[i] - axis2[i] * lz2;
  }
  closest_segment_points(a1, a2, b1, b2, sphere1, sphere2);
  return collide_sphere_sphere(sphere1, rad1, sphere2, rad2, c);
}

// capsule-plane: capsule.cpp:352-409 (the deeper end cap, then the other)
template <class R>
int collide_capsule_plane(const R *pc, const R *Rc, R rad, R len, const R *plane, int flags, ContactGeom<R> *c) {
  R ax[3] = {Rc[2], Rc[6], Rc[10]};
  R sign = (dot3(plane, ax) > 0) ? R(-1.0) : R(1.0);
  R p[3];
  for (int i = 0; i < 3; i++) p[i] = pc[i] + ax[i] * len * R(0.5) * sign;
  R k = dot3(p, plane);
  R depth = plane[3] - k + rad;
  if (depth < 0) return 0;
  for (int i = 0; i < 3; i++) {
    c[0].normal[i] = plane[i];
    c[0].pos[i] = p[i] - plane[i] * rad;
  }
  c[0].depth = depth;
  int n = 1;
  if ((flags & 0xffff) >= 2) {
    for (int i = 0; i < 3; i++) p[i] = pc[i] - ax[i] * len * R(0.5) * sign;
    k = dot3(p, plane);
    depth = plane[3] - k + rad;
    if (depth >= 0) {
      for (int i = 0; i < 3; i++) {
        c[1].normal[i] = plane[i];
        c[1].pos[i] = p[i] - plane[i] * rad;
      }
      c[1].depth = depth;
      n = 2;
    }
  }
  return n;
}

// ------------------------------------------------------------ rays
// A ray (param: length) starts at its position along column 2 of its
// rotation; the contact depth is the distance along the ray.

// ray.cpp:196-232: the entry point (or, for mode 1 from outside, the exit
// point) of the ray through a sphere
template <class R>
int ray_sphere_hit(const R *pr, const R *Rr, R len, const R *ps, R rad, int mode, ContactGeom<R> *c) {
  R q[3] = {pr[0] - ps[0], pr[1] - ps[1], pr[2] - ps[2]};
  R B = q[0] * Rr[2] + q[1] * Rr[6] + q[2] * Rr[10];
  R C = dot3(q, q) - rad * rad;
  R k = B * B - C;
  if (k < 0) return 0;
  k = std::sqrt(k);
  R alpha;
  if (mode && C >= 0) {
    alpha = -B + k;
    if (alpha < 0) return 0;
  } else {
    alpha = -B - k;
    if (alpha < 0) {
      alpha = -B + k;
      if (alpha < 0) return 0;
    }
  }
  if (alpha > len) return 0;
  for (int i = 0; i < 3; i++) c->pos[i] = pr[i] + alpha * Rr[4 * i + 2];
  R nsign = (C < 0 || mode) ? R(-1.0) : R(1.0);
  for (int i = 0; i < 3; i++) c->normal[i] = nsign * (c->pos[i] - ps[i]);
  normalize3(c->normal);
  c->depth = alpha;
  return 1;
}

// ray-box: ray.cpp:253-349 (slab intersection in box coordinates)
template <class R>
int collide_ray_box(const R *pr, const R *Rr, R len, const R *pb, const R *Rb, const R *side, ContactGeom<R> *c) {
  R tmp[3], s[3], v[3], sign[3], h[3];
  for (int i = 0; i < 3; i++) tmp[i] = pr[i] - pb[i];
  mtv(s, Rb, tmp);
  for (int i = 0; i < 3; i++) tmp[i] = Rr[4 * i + 2];
  mtv(v, Rb, tmp);
  for (int i = 0; i < 3; i++) {
    if (v[i] < 0) { s[i] = -s[i]; v[i] = -v[i]; sign[i] = 1; }
    else sign[i] = -1;
  }
  for (int i = 0; i < 3; i++) h[i] = R(0.5) * side[i];
  if ((s[0] < -h[0] && v[0] <= 0) || s[0] > h[0] || (s[1] < -h[1] && v[1] <= 0) || s[1] > h[1] ||
      (s[2] < -h[2] && v[2] <= 0) || s[2] > h[2] || (v[0] == 0 && v[1] == 0 && v[2] == 0))
    return 0;
  R lo = -INFINITY, hi = INFINITY;
  int nlo = 0, nhi = 0;
  for (int i = 0; i < 3; i++) {
    if (v[i] != 0) {
      R k = (-h[i] - s[i]) / v[i];
      if (k > lo) { lo = k; nlo = i; }
      k = (h[i] - s[i]) / v[i];
      if (k < hi) { hi = k; nhi = i; }
    }
  }
  if (lo > hi) return 0;
  R alpha;
  int n;
  if (lo >= 0) { alpha = lo; n = nlo; }
  else { alpha = hi; n = nhi; }
  if (alpha < 0 || alpha > len) return 0;
  for (int i = 0; i < 3; i++) c->pos[i] = pr[i] + alpha * Rr[4 * i + 2];
  for (int i = 0; i < 3; i++) c->normal[i] = Rb[4 * i + n] * sign[n];
  c->depth = alpha;
  return 1;
}

// ray-capsule: ray.cpp:352-464 (the infinite cylinder, then the end caps)
template <class R>
int collide_ray_capsule(const R *pr, const R *Rr, R len, const R *pc, const R *Rc, R rad, R clen,
                        ContactGeom<R> *c) {
  R lz2 = clen * R(0.5);
  R cs[3] = {pr[0] - pc[0], pr[1] - pc[1], pr[2] - pc[2]};
  R k = Rc[2] * cs[0] + Rc[6] * cs[1] + Rc[10] * cs[2];
  R q[3], r[3];
  for (int i = 0; i < 3; i++) q[i] = k * Rc[4 * i + 2] - cs[i];
  R C = dot3(q, q) - rad * rad;
  int inside = 0;
  if (C < 0) {
    if (k < -lz2) k = -lz2;
    else if (k > lz2) k = lz2;
    for (int i = 0; i < 3; i++) r[i] = pc[i] + k * Rc[4 * i + 2];
    if ((pr[0] - r[0]) * (pr[0] - r[0]) + (pr[1] - r[1]) * (pr[1] - r[1]) + (pr[2] - r[2]) * (pr[2] - r[2]) <
        rad * rad)
      inside = 1;
  }
  if (!inside && C < 0) {
    k = (k < 0) ? -lz2 : lz2;
  } else {
    R uv = Rc[2] * Rr[2] + Rc[6] * Rr[6] + Rc[10] * Rr[10];
    for (int i = 0; i < 3; i++) r[i] = uv * Rc[4 * i + 2] - Rr[4 * i + 2];
    R A = dot3(r, r);
    R B = 2 * dot3(q, r);
    k = B * B - 4 * A * C;
    if (k < 0) {
      if (!inside) return 0;
      k = (uv < 0) ? -lz2 : lz2;
    } else {
      k = std::sqrt(k);
      A = recip(2 * A);
      R alpha = (-B - k) * A;
      if (alpha < 0) {
        alpha = (-B + k) * A;
        if (alpha < 0) return 0;
      }
      if (alpha > len) return 0;
      for (int i = 0; i < 3; i++) c->pos[i] = pr[i] + alpha * Rr[4 * i + 2];
      for (int i = 0; i < 3; i++) q[i] = c->pos[i] - pc[i];
      k = q[0] * Rc[2] + q[1] * Rc[6] + q[2] * Rc[10];
      R nsign = inside ? R(-1.0) : R(1.0);
      if (k >= -lz2 && k <= lz2) {
        for (int i = 0; i < 3; i++) c->normal[i] = nsign * (c->pos[i] - (pc[i] + k * Rc[4 * i + 2]));
        normalize3(c->normal);
        c->depth = alpha;
        return 1;
      }
      k = (k < 0) ? -lz2 : lz2;
    }
  }
  for (int i = 0; i < 3; i++) q[i] = pc[i] + k * Rc[4 * i + 2];
  return ray_sphere_hit(pr, Rr, len, q, rad, inside, c);
}

// ray-plane: ray.cpp:467-497
template <class R>
int collide_ray_plane(const R *pr, const R *Rr, R len, const R *plane, ContactGeom<R> *c) {
  R alpha = plane[3] - dot3(plane, pr);
  R nsign = (alpha > 0) ? R(-1.0) : R(1.0);
  R k = plane[0] * Rr[2] + plane[1] * Rr[6] + plane[2] * Rr[10];
  if (k == 0) return 0;
  alpha /= k;
  if (alpha < 0 || alpha > len) return 0;
  for (int i = 0; i < 3; i++) c->pos[i] = pr[i] + alpha * Rr[4 * i + 2];
  for (int i = 0; i < 3; i++) c->normal[i] = nsign * plane[i];
  c->depth = alpha;
  return 1;
}

// ------------------------------------------------------------ cylinders
// A cylinder (param: radius, length) has its axis along column 2 of its
// rotation.  Without libccd (the reference's config.h) the colliders are
// cylinder-sphere, cylinder-plane, cylinder-box and ray-cylinder;
// cylinder-cylinder and capsule-cylinder pairs have none (no contacts).

// the tolerance of the parallel tests, per precision (collision_cylinder_
// sphere.cpp:66-71, collision_cylinder_plane.cpp:51-56)
template <class R> inline R cyl_tol() { return sizeof(R) == 4 ? R(0.0001) : R(0.0000001); }

// cylinder-sphere: collision_cylinder_sphere.cpp:57-278 -- s along the axis
// from the first disc, t from the axis; the side (nappe), a disc or a disc's
// rim
template <class R>
int collide_cylinder_sphere(const R *pc, const R *Rc, R radius, R length, const R *ps, R radius2,
                            ContactGeom<R> *c) {
  const R tol = cyl_tol<R>();
  R ax[3] = {Rc[2], Rc[6], Rc[10]};
  R h = length * R(0.5);
  R top[3], bot[3];
  for (int i = 0; i < 3; i++) {
    top[i] = ax[i] * h + pc[i];
    bot[i] = ax[i] * -h + pc[i];
  }
  R s = (ps[0] - bot[0]) * ax[0] - (bot[1] - ps[1]) * ax[1] - (bot[2] - ps[2]) * ax[2];
  if (s < -radius2 || s > length + radius2) return 0;
  R C[3];
  for (int i = 0; i < 3; i++) C[i] = s * ax[i] + bot[i] - ps[i];
  R t = std::sqrt(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]);
  if (t > radius + radius2) return 0;
  if (t > radius && (s < 0 || s > length)) {
    // the rim of a disc
    const R *disc = (s <= 0) ? bot : top;
    R e = (s <= 0) ? s : s - length;
    c->depth = radius2 - std::sqrt(e * e + (t - radius) * (t - radius));
    if (c->depth < 0) return 0;
    for (int i = 0; i < 3; i++) c->pos[i] = C[i] / t * -radius + disc[i];
    for (int i = 0; i < 3; i++) c->normal[i] = (c->pos[i] - ps[i]) / (radius2 - c->depth);
    return 1;
  }
  if ((radius - t) <= s && (radius - t) <= (length - s)) {
    // the side
    c->depth = (radius2 + radius) - t;
    if (c->depth < 0) return 0;
    if (t > radius2 + tol) {
      for (int i = 0; i < 3; i++) C[i] /= t;
      for (int i = 0; i < 3; i++) c->pos[i] = C[i] * radius2 + ps[i];
      for (int i = 0; i < 3; i++) c->normal[i] = C[i];
    } else {
      for (int i = 0; i < 3; i++) c->pos[i] = C[i] + ps[i];
      for (int i = 0; i < 3; i++) c->normal[i] = C[i] / t;
    }
    return 1;
  }
  // a disc
  if (s <= length * R(0.5)) {
    c->depth = s + radius2;
    if (c->depth < 0) return 0;
    for (int i = 0; i < 3; i++) c->pos[i] = radius2 * ax[i] + ps[i];
    for (int i = 0; i < 3; i++) c->normal[i] = ax[i];
  } else {
    c->depth = radius2 + length - s;
    if (c->depth < 0) return 0;
    for (int i = 0; i < 3; i++) c->pos[i] = radius2 * -ax[i] + ps[i];
    for (int i = 0; i < 3; i++) c->normal[i] = -ax[i];
  }
  return 1;
}

// cylinder-plane: collision_cylinder_plane.cpp:40-267 -- a disc parallel to
// the plane gives up to four rim points, otherwise the deepest rim point of
// each disc
template <class R>
int collide_cylinder_plane(const R *pc, const R *Rc, R radius, R length, const R *pl, int flags,
                           ContactGeom<R> *c) {
  const R tol = cyl_tol<R>();
  const int maxc = flags & 0xffff;
  R ax[3] = {Rc[2], Rc[6], Rc[10]};
  R h = length * R(0.5);
  R top[3], bot[3];
  for (int i = 0; i < 3; i++) {
    top[i] = ax[i] * h + pc[i];
    bot[i] = ax[i] * -h + pc[i];
  }
  int n = 0;
  auto emit = [&](const R *p, bool strict, bool plain_dot) -> bool {
    ContactGeom<R> &k = c[n];
    for (int i = 0; i < 3; i++) k.pos[i] = p[i];
    k.depth = plain_dot ? pl[3] - dot3(pl, k.pos) : pl[3] - pl[0] * k.pos[0] - pl[1] * k.pos[1] - pl[2] * k.pos[2];
    if (strict ? k.depth > 0 : k.depth >= 0) {
      for (int i = 0; i < 3; i++) k.normal[i] = pl[i];
      n++;
      return n >= maxc;
    }
    return false;
  };
  R s = ax[0] * pl[0] + ax[1] * pl[1] + ax[2] * pl[2];
  if (s < 0) s += R(1.0);
  else s -= R(1.0);
  if (s < tol && s > -tol) {
    R d1 = pl[3] - dot3(pl, bot);
    R d2 = pl[3] - dot3(pl, top);
    const R *P;
    if (d1 >= d2) {
      if (!(d1 >= 0)) return 0;
      P = bot;
    } else {
      if (!(d2 >= 0)) return 0;
      P = top;
    }
    R V1[3], V2[3];
    if (ax[0] < tol && ax[0] > -tol) { V1[0] = ax[0] + R(1.0); V1[1] = ax[1]; V1[2] = ax[2]; }
    else { V1[0] = ax[0]; V1[1] = ax[1] + R(1.0); V1[2] = ax[2]; }
    cross3(V2, V1, ax);
    R l = std::sqrt(V2[0] * V2[0] + V2[1] * V2[1] + V2[2] * V2[2]);
    l = radius / l;
    for (int i = 0; i < 3; i++) V2[i] *= l;
    cross3(V1, V2, ax);
    R q[3];
    for (int i = 0; i < 3; i++) q[i] = P[i] + V1[i];
    if (emit(q, true, true)) return n;
    for (int i = 0; i < 3; i++) q[i] = P[i] - V1[i];
    if (emit(q, true, true)) return n;
    for (int i = 0; i < 3; i++) q[i] = P[i] + V2[i];
    if (emit(q, true, true)) return n;
    for (int i = 0; i < 3; i++) q[i] = P[i] - V2[i];
    emit(q, true, true);
    return n;
  }
  R t = dot3(pl, ax);
  R C[3];
  for (int i = 0; i < 3; i++) C[i] = ax[i] * t - pl[i];
  R l = std::sqrt(C[0] * C[0] + C[1] * C[1] + C[2] * C[2]);
  l = radius / l;
  for (int i = 0; i < 3; i++) C[i] *= l;
  R q[3];
  for (int i = 0; i < 3; i++) q[i] = C[i] + bot[i];
  if (emit(q, false, true)) return n;
  for (int i = 0; i < 3; i++) q[i] = C[i] + top[i];
  emit(q, false, false);
  return n;
}

// cylinder-box: collision_cylinder_box.cpp (separating axes over the box
// faces, the cylinder axis, their cross products, the box vertices and the
// box edges against both rims; then the cylinder's side segment clipped to
// the box, or the box face clipped to an 8-gon of the disc)
template <class R> struct CylBox {
  const R *Rc, *pc, *Rb, *pb;
  R radius, size, ax[3], half[3], vert[8][3], diff[3], normal[3], best_depth, best_rb, best_rc;
  int best_axis;

  CylBox(const R *pc_, const R *Rc_, R radius_, R size_, const R *pb_, const R *Rb_, const R *side)
      : Rc(Rc_), pc(pc_), Rb(Rb_), pb(pb_), radius(radius_), size(size_) {
    for (int i = 0; i < 3; i++) { ax[i] = Rc[4 * i + 2]; half[i] = side[i] * R(0.5); }
    // box vertices (collision_cylinder_box.cpp:130-180 numbering) in world space
    static const int sg[8][3] = {{-1, 1, -1}, {1, 1, -1}, {-1, -1, -1}, {1, -1, -1},
                                 {1, 1, 1}, {1, -1, 1}, {-1, -1, 1}, {-1, 1, 1}};
    for (int v = 0; v < 8; v++) {
      R loc[3] = {sg[v][0] * half[0], sg[v][1] * half[1], sg[v][2] * half[2]}, w[3];
      mv(w, Rb, loc);
      for (int i = 0; i < 3; i++) vert[v][i] = w[i] + pb[i];
    }
    for (int i = 0; i < 3; i++) { diff[i] = pc[i] - pb[i]; normal[i] = 0; }
    best_depth = INFINITY;
    best_rb = best_rc = 0;
    best_axis = 0;
  }

  // collision_cylinder_box.cpp:206-272
  bool test_axis(R *n, int id) {
    R l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (l < R(1e-5)) return true;
    normalize3(n);
    R d1 = dot3(ax, n);
    R rc;
    if (d1 > R(1.0) || d1 < R(-1.0)) rc = size * R(0.5);
    else rc = std::fabs(d1 * (size * R(0.5))) + radius * std::sqrt(R(1.0) - d1 * d1);
    R rb = std::fabs(dot3(Rb, n, 4, 1)) * half[0];
    rb += std::fabs(dot3(Rb + 1, n, 4, 1)) * half[1];
    rb += std::fabs(dot3(Rb + 2, n, 4, 1)) * half[2];
    R fd = dot3(diff, n);
    R depth = rc + rb;
    if (std::fabs(fd) > depth) return false;
    depth -= std::fabs(fd);
    if (depth < best_depth) {
      best_depth = depth;
      for (int i = 0; i < 3; i++) normal[i] = n[i];
      best_axis = id;
      best_rb = rb;
      best_rc = rc;
      if (fd > 0) for (int i = 0; i < 3; i++) normal[i] = -normal[i];
    }
    return true;
  }

  // collision_cylinder_box.cpp:275-320: the axis normal to a box edge and the
  // rim's tangent where the edge's line meets the rim's plane
  bool test_edge_rim(const R *centre, const R *v0, const R *v1, int id) {
    R dir[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
    normalize3(dir);
    R d2 = dot3(dir, ax);
    if (std::fabs(d2) < R(1e-5)) return true;
    R t[3] = {centre[0] - v0[0], centre[1] - v0[1], centre[2] - v0[2]};
    R d1 = dot3(t, ax);
    R p[3];
    for (int i = 0; i < 3; i++) p[i] = v0[i] + dir[i] * (d1 / d2);
    for (int i = 0; i < 3; i++) t[i] = centre[i] - p[i];
    R tan[3], a[3];
    cross3(tan, t, ax);
    cross3(a, tan, dir);
    return test_axis(a, id);
  }

  // collision_cylinder_box.cpp:323-560
  bool separated() {
    R a[3];
    for (int k = 0; k < 3; k++) {
      for (int i = 0; i < 3; i++) a[i] = Rb[4 * i + k];
      if (!test_axis(a, 1 + k)) return true;
    }
    for (int i = 0; i < 3; i++) a[i] = ax[i];
    if (!test_axis(a, 4)) return true;
    const R eps = R(1e-6);
    for (int k = 0; k < 3; k++) {
      cross3(a, ax, Rb + k, 1, 1, 4);
      if (dot3(a, a) > eps && !test_axis(a, 5 + k)) return true;
    }
    for (int v = 0; v < 8; v++) {
      R t1[3] = {vert[v][0] - pc[0], vert[v][1] - pc[1], vert[v][2] - pc[2]}, t2[3];
      cross3(t2, ax, t1);
      cross3(a, ax, t2);
      if (dot3(a, a) > eps && !test_axis(a, 8 + v)) return true;
    }
    // the twelve box edges against the top rim (ids 16-27), then the bottom
    // rim (ids 28-39), in the reference's edge order
    static const int edges[12][2] = {{1, 0}, {1, 3}, {2, 3}, {2, 0}, {4, 1}, {4, 7},
                                     {0, 7}, {5, 3}, {5, 6}, {2, 6}, {4, 5}, {6, 7}};
    for (int side = 0; side < 2; side++) {
      R h = size * R(0.5), cc[3];
      for (int i = 0; i < 3; i++) cc[i] = side == 0 ? pc[i] + ax[i] * h : pc[i] - ax[i] * h;
      for (int e = 0; e < 12; e++)
        if (!test_edge_rim(cc, vert[edges[e][0]], vert[edges[e][1]], 16 + 12 * side + e)) return true;
    }
    return false;
  }

  // collision_util.cpp:471-509
  static bool clip_edge(R *e0, R *e1, const R *pl) {
    R d0 = pl[0] * e0[0] + pl[1] * e0[1] + pl[2] * e0[2] + pl[3];
    R d1 = pl[0] * e1[0] + pl[1] * e1[1] + pl[2] * e1[2] + pl[3];
    if (d0 < 0 && d1 < 0) return false;
    if (d0 > 0 && d1 > 0) return true;
    if ((d0 > 0 && d1 < 0) || (d0 < 0 && d1 > 0)) {
      R x[3];
      for (int i = 0; i < 3; i++) x[i] = e0[i] - (e0[i] - e1[i]) * d0 / (d0 - d1);
      R *dst = d0 < 0 ? e0 : e1;
      for (int i = 0; i < 3; i++) dst[i] = x[i];
    }
    return true;
  }

  // collision_util.cpp:512-555
  static void clip_poly(const R (*in)[3], int nin, R (*out)[3], int &nout, const R *pl) {
    nout = 0;
    for (int i1 = 0, i0 = nin - 1; i1 < nin; i0 = i1, i1++) {
      R d0 = pl[0] * in[i0][0] + pl[1] * in[i0][1] + pl[2] * in[i0][2] + pl[3];
      R d1 = pl[0] * in[i1][0] + pl[1] * in[i1][1] + pl[2] * in[i1][2] + pl[3];
      if (d0 >= 0) {
        for (int i = 0; i < 3; i++) out[nout][i] = in[i0][i];
        nout++;
      }
      if ((d0 > 0 && d1 < 0) || (d0 < 0 && d1 > 0)) {
        for (int i = 0; i < 3; i++) out[nout][i] = in[i0][i] - (in[i0][i] - in[i1][i]) * d0 / (d0 - d1);
        nout++;
      }
    }
  }

  void put(ContactGeom<R> *c, int &n, const R *p, R depth) {
    for (int i = 0; i < 3; i++) { c[n].pos[i] = p[i]; c[n].normal[i] = -normal[i]; }
    c[n].depth = depth;
    n++;
  }

  // collision_cylinder_box.cpp:563-680: the side segment of the cylinder
  // facing the box, clipped by the six face planes
  int clip_side(int maxc, ContactGeom<R> *c) {
    R t = dot3(ax, normal), vn[3];
    for (int i = 0; i < 3; i++) vn[i] = normal[i] - ax[i] * t;
    normalize3(vn);
    R h = size * R(0.5), e0[3], e1[3];
    for (int i = 0; i < 3; i++) {
      R q = pc[i] + vn[i] * radius;
      e0[i] = q + ax[i] * h;
      e1[i] = q - ax[i] * h;
    }
    for (int i = 0; i < 3; i++) { e0[i] -= pb[i]; e1[i] -= pb[i]; }
    for (int sgn = 0; sgn < 2; sgn++)
      for (int k = 0; k < 3; k++) {
        R pl[4] = {Rb[k], Rb[4 + k], Rb[8 + k], half[k]};
        if (sgn) { pl[0] = -pl[0]; pl[1] = -pl[1]; pl[2] = -pl[2]; }
        if (!clip_edge(e0, e1, pl)) return 0;
      }
    R dep0 = best_rb + dot3(e0, normal), dep1 = best_rb + dot3(e1, normal);
    if (dep0 < 0) dep0 = 0;
    if (dep1 < 0) dep1 = 0;
    for (int i = 0; i < 3; i++) { e0[i] += pb[i]; e1[i] += pb[i]; }
    int n = 0;
    put(c, n, e0, dep0);
    if (n != maxc) put(c, n, e1, dep1);
    return n;
  }

  // collision_util.h:176-210 (the determinant and the inverse as written
  // there, division binding to the second product of the diagonal and
  // off-diagonal terms the way the expression parses)
  static void inv3(const R *m, R *d) {
    R det = m[0] * (m[5] * m[10] - m[9] * m[6]) - m[1] * (m[4] * m[10] - m[8] * m[6]) +
            m[2] * (m[4] * m[9] - m[8] * m[5]);
    for (int i = 0; i < 12; i++) d[i] = 0;
    if (std::fabs(det) < R(0.0005)) {
      d[0] = d[5] = d[10] = 1;
      return;
    }
    d[0] = m[5] * m[10] - m[6] * m[9] / det;
    d[1] = -(m[1] * m[10] - m[9] * m[2]) / det;
    d[2] = m[1] * m[6] - m[5] * m[2] / det;
    d[4] = -(m[4] * m[10] - m[6] * m[8]) / det;
    d[5] = m[0] * m[10] - m[8] * m[2] / det;
    d[6] = -(m[0] * m[6] - m[4] * m[2]) / det;
    d[8] = m[4] * m[9] - m[8] * m[5] / det;
    d[9] = -(m[0] * m[9] - m[8] * m[1]) / det;
    d[10] = m[0] * m[5] - m[1] * m[4] / det;
  }

  // collision_cylinder_box.cpp:683-1000: the box face most facing the
  // normal, in the disc's frame, clipped by the disc plane and the eight
  // planes of an 8-gon around the rim
  int clip_face(int maxc, ContactGeom<R> *c) {
    R h = size * R(0.5), cpos[3], cn[3] = {0, 0, 0};
    if (dot3(ax, normal) > R(0.0)) {
      for (int i = 0; i < 3; i++) cpos[i] = pc[i] + ax[i] * h;
      cn[2] = R(-1.0);
    } else {
      for (int i = 0; i < 3; i++) cpos[i] = pc[i] - ax[i] * h;
      cn[2] = R(1.0);
    }
    R binv[12], nr[3], an[3];
    inv3(Rb, binv);
    mv(nr, binv, normal);
    for (int i = 0; i < 3; i++) an[i] = std::fabs(nr[i]);
    int b0, b1, b2;
    if (an[1] > an[0]) {
      if (an[0] > an[2]) { b0 = 1; b1 = 0; b2 = 2; }
      else if (an[1] > an[2]) { b0 = 1; b1 = 2; b2 = 0; }
      else { b0 = 2; b1 = 1; b2 = 0; }
    } else {
      if (an[1] > an[2]) { b0 = 0; b1 = 1; b2 = 2; }
      else if (an[0] > an[2]) { b0 = 0; b1 = 2; b2 = 1; }
      else { b0 = 2; b1 = 0; b2 = 1; }
    }
    R ctr[3];
    for (int i = 0; i < 3; i++)
      ctr[i] = nr[b0] > 0 ? pb[i] - half[b0] * Rb[4 * i + b0] : pb[i] + half[b0] * Rb[4 * i + b0];
    R pts[4][3], A[16][3], B[16][3];
    for (int k = 0; k < 16; k++)
      for (int i = 0; i < 3; i++) A[k][i] = B[k][i] = 0;
    static const int s1[4] = {1, -1, -1, 1}, s2[4] = {-1, -1, 1, 1};
    for (int k = 0; k < 4; k++)
      for (int i = 0; i < 3; i++) {
        R u = half[b1] * Rb[4 * i + b1], v = half[b2] * Rb[4 * i + b2];
        R x = s1[k] > 0 ? ctr[i] + u : ctr[i] - u;
        pts[k][i] = s2[k] > 0 ? x + v : x - v;
      }
    R cinv[12];
    inv3(Rc, cinv);
    for (int k = 0; k < 4; k++) {
      R t[3] = {pts[k][0] - cpos[0], pts[k][1] - cpos[1], pts[k][2] - cpos[2]};
      mv(pts[k], cinv, t);
    }
    int n1 = 0, n2 = 0;
    R pl[4] = {cn[0], cn[1], cn[2], R(0.0)};
    clip_poly(pts, 4, A, n1, pl);
    R ang = (R)(M_PI / 8);
    const R inc = ang * R(2.0);
    int seg = 0;
    for (; seg < 8; seg++) {
      R rn[4] = {-std::cos(ang), -std::sin(ang), 0, radius};
      ang += inc;
      if (seg % 2 == 0) clip_poly(A, n1, B, n2, rn);
      else clip_poly(B, n2, A, n1, rn);
    }
    // seg == 8: the last clip wrote A
    int n = 0;
    for (int k = 0; k < n1; k++) {
      R p[3], t[3];
      mv(p, Rc, A[k]);
      for (int i = 0; i < 3; i++) p[i] += cpos[i];
      for (int i = 0; i < 3; i++) t[i] = p[i] - pc[i];
      R d = best_rc - dot3(t, normal);
      if (d > R(0.0)) {
        put(c, n, p, d);
        if (n == maxc) break;
      }
    }
    return n;
  }

  int run(int flags, ContactGeom<R> *c) {
    if (separated() || best_axis == 0) return 0;
    const int maxc = flags & 0xffff;
    R d = dot3(normal, ax);
    if (std::fabs(d) < R(0.9)) return clip_side(maxc, c);
    return clip_face(maxc, c);
  }
};

// ray-cylinder: ray.cpp:500-656 -- the flat caps when the ray runs inside
// the infinite cylinder parallel to the axis, else the curved side
template <class R>
int collide_ray_cylinder(const R *pr, const R *Rr, R len, const R *pc, const R *Rc, R radius, R clen,
                         ContactGeom<R> *c) {
  const R eps = sizeof(R) == 4 ? R(1.1920929e-07) : R(2.220446049250313e-16);
  R hl = clen * R(0.5);
  R r[3] = {pr[0] - pc[0], pr[1] - pc[1], pr[2] - pc[2]};
  R d = dot3(Rc + 2, r, 4, 1);
  R q[3];
  for (int i = 0; i < 3; i++) q[i] = d * Rc[4 * i + 2] - r[i];
  R C = dot3(q, q) - radius * radius;
  R uv = dot3(Rc + 2, Rr + 2, 4, 4);
  for (int i = 0; i < 3; i++) r[i] = uv * Rc[4 * i + 2] - Rr[4 * i + 2];
  R A = dot3(r, r);
  R B = 2 * dot3(q, r);
  R k = B * B - 4 * A * C;
  if (k < eps && C <= 0) {
    R uvsign = (uv < 0) ? R(-1.0) : R(1.0);
    R internal = (d >= -hl && d <= +hl) ? R(-1.0) : R(1.0);
    if ((uv > 0 && d + uvsign * len < hl * internal) || (uv < 0 && d + uvsign * len > hl * internal)) return 0;
    c->depth = (-uvsign * d) - (internal * hl);
    for (int i = 0; i < 3; i++) c->pos[i] = pr[i] + c->depth * Rr[4 * i + 2];
    for (int i = 0; i < 3; i++) c->normal[i] = uvsign * Rc[4 * i + 2];
    return 1;
  }
  if (k > 0) {
    k = std::sqrt(k);
    A = recip(2 * A);
    R alpha = (-B - k) * A;
    if (alpha < 0) alpha = (-B + k) * A;
    if (alpha >= 0 && alpha <= len) {
      for (int i = 0; i < 3; i++) c->pos[i] = pr[i] + alpha * Rr[4 * i + 2];
      for (int i = 0; i < 3; i++) q[i] = c->pos[i] - pc[i];
      d = dot3(q, Rc + 2, 1, 4);
      if (d >= -hl && d <= +hl) {
        R nsign = (C < 0) ? R(-1.0) : R(1.0);
        for (int i = 0; i < 3; i++) c->normal[i] = nsign * (c->pos[i] - (pc[i] + d * Rc[4 * i + 2]));
        normalize3(c->normal);
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
inline int collider_rank(int type) {
  return type == G_RAY ? 0
         : type == G_CAPSULE ? 1
         : type == G_CYLINDER ? 2
         : type == G_SPHERE ? 3
         : type == G_BOX ? 4 : 5;
}

template <class R>
int collide_geoms(const GeomShape<R> &a, int ida, const GeomShape<R> &b, int idb, int flags, ContactGeom<R> *c) {
  if ((flags & 0xffff) == 0) return 0;
  auto run = [&](const GeomShape<R> &x, const GeomShape<R> &y) -> int {
    if (x.type == G_RAY) {
      if (y.type == G_SPHERE) return ray_sphere_hit(x.pos, x.Rm, x.param[0], y.pos, y.param[0], 0, c);
      if (y.type == G_BOX) return collide_ray_box(x.pos, x.Rm, x.param[0], y.pos, y.Rm, y.param, c);
      if (y.type == G_CAPSULE)
        return collide_ray_capsule(x.pos, x.Rm, x.param[0], y.pos, y.Rm, y.param[0], y.param[1], c);
      if (y.type == G_PLANE) return collide_ray_plane(x.pos, x.Rm, x.param[0], y.param, c);
      if (y.type == G_CYLINDER)
        return collide_ray_cylinder(x.pos, x.Rm, x.param[0], y.pos, y.Rm, y.param[0], y.param[1], c);
      return 0;
    }
    if (x.type == G_CYLINDER) {
      if (y.type == G_SPHERE) return collide_cylinder_sphere(x.pos, x.Rm, x.param[0], x.param[1], y.pos, y.param[0], c);
      if (y.type == G_PLANE) return collide_cylinder_plane(x.pos, x.Rm, x.param[0], x.param[1], y.param, flags, c);
      if (y.type == G_BOX) return CylBox<R>(x.pos, x.Rm, x.param[0], x.param[1], y.pos, y.Rm, y.param).run(flags, c);
      return 0;
    }
    if (x.type == G_CAPSULE) {
      if (y.type == G_SPHERE) return collide_capsule_sphere(x.pos, x.Rm, x.param[0], x.param[1], y.pos, y.param[0], c);
      if (y.type == G_BOX)
        return collide_capsule_box(x.pos, x.Rm, x.param[0], x.param[1], y.pos, y.Rm, y.param, c);
      if (y.type == G_CAPSULE)
        return collide_capsule_capsule(x.pos, x.Rm, x.param[0], x.param[1], y.pos, y.Rm, y.param[0], y.param[1], flags,
                                       c);
      if (y.type == G_PLANE) return collide_capsule_plane(x.pos, x.Rm, x.param[0], x.param[1], y.param, flags, c);
      return 0;
    }
    if (x.type == G_SPHERE && y.type == G_SPHERE) return collide_sphere_sphere(x.pos, x.param[0], y.pos, y.param[0], c);
    if (x.type == G_SPHERE && y.type == G_BOX) return collide_sphere_box(x.pos, x.param[0], y.pos, y.Rm, y.param, c);
    if (x.type == G_SPHERE && y.type == G_PLANE) return collide_sphere_plane(x.pos, x.param[0], y.param, c);
    if (x.type == G_BOX && y.type == G_BOX) return collide_box_box(x.pos, x.Rm, x.param, y.pos, y.Rm, y.param, flags, c);
    if (x.type == G_BOX && y.type == G_PLANE) return collide_box_plane(x.pos, x.Rm, x.param, y.param, flags, c);
    return 0;
  };
  bool reverse = collider_rank(a.type) > collider_rank(b.type);
  int n = reverse ? run(b, a) : run(a, b);
  for (int i = 0; i < n; i++) {
    if (reverse) for (int k = 0; k < 3; k++) c[i].normal[k] = -c[i].normal[k];
    c[i].g1 = ida;
    c[i].g2 = idb;
  }
  return n;
}

}  // namespace orc
