// orc_collide.hpp -- CPU restatement of the narrowphase colliders on the path
// (sphere-sphere, sphere-box, sphere-plane, box-box, box-plane), the geom
// AABBs and the collider dispatch.
//
// TEST INFRASTRUCTURE (oracle).  Each function cites the reference code whose
// behaviour it restates.  The contact arrays hold up to `maxc` contacts of
// pos[3], normal[3], depth.
#pragma once
#include <cmath>
#include "orc_math.hpp"

namespace orc {

enum { G_SPHERE = 0, G_BOX = 1, G_PLANE = 2, G_CAPSULE = 3 };
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

// the class that comes first in a collider's arguments (the colliders table,
// collision_kernel.cpp:181-189): capsule, then sphere, box, plane
inline int collider_rank(int type) {
  return type == G_CAPSULE ? 0 : (type == G_SPHERE ? 1 : (type == G_BOX ? 2 : 3));
}

template <class R>
int collide_geoms(const GeomShape<R> &a, int ida, const GeomShape<R> &b, int idb, int flags, ContactGeom<R> *c) {
  if ((flags & 0xffff) == 0) return 0;
  auto run = [&](const GeomShape<R> &x, const GeomShape<R> &y) -> int {
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
