// adj_colliders.hpp -- the narrowphase of colliders.hpp written over a
// generic scalar, so that evaluating it on dual numbers yields the
// derivatives of every contact's position, normal and depth with respect to
// the two geoms' poses.  Branch decisions are taken on the values and follow
// colliders.hpp line for line (same reference colliders: collision_util.cpp,
// sphere.cpp, box.cpp), so the contacts come out in the same number and
// order as in the forward step.
#pragma once
#include "adj_math.hpp"

namespace odemi {
namespace adj {

template <class S>
struct CP {
  V3<S> pos, n;
  S depth;
};

template <class S>
struct GeomD {
  int type;
  V3<S> p;
  M3<S> R;
  float prm[4];
};

template <class S> OD_HD inline int d_sphere_sphere(const V3<S> &p1, float r1, const V3<S> &p2, float r2, CP<S> *c) {
  V3<S> d = p1 - p2;
  S dist = sqrt_(dot(d, d));
  if (val(dist) > r1 + r2) return 0;
  if (val(dist) <= 0) {
    c->pos = p1;
    c->n = v3(S(1), S(0), S(0));
    c->depth = S(r1 + r2);
  } else {
    c->n = (S(1) / dist) * d;
    c->pos = p1 + (S(0.5f) * (S(r2 - r1) - dist)) * c->n;
    c->depth = S(r1 + r2) - dist;
  }
  return 1;
}

template <class S>
OD_HD inline int d_sphere_box(const V3<S> &ps, float rad, const V3<S> &pb, const M3<S> &R, const float *side, CP<S> *c) {
  V3<S> p = ps - pb;
  float l[3] = {0.5f * side[0], 0.5f * side[1], 0.5f * side[2]};
  S t[3];
  bool border = false;
  for (int k = 0; k < 3; k++) {
    S v = dot(p, R.col(k));
    if (val(v) < -l[k]) { v = S(-l[k]); border = true; }
    if (val(v) > l[k]) { v = S(l[k]); border = true; }
    t[k] = v;
  }
  if (!border) {
    S best = S(l[0]) - fabs_(t[0]);
    int bi = 0;
    S f1 = S(l[1]) - fabs_(t[1]);
    if (val(f1) < val(best)) { best = f1; bi = 1; }
    S f2 = S(l[2]) - fabs_(t[2]);
    if (val(f2) < val(best)) { best = f2; bi = 2; }
    c->pos = ps;
    float sgn = (val(t[bi]) > 0) ? 1.f : -1.f;
    c->n = S(sgn) * R.col(bi);
    c->depth = best + S(rad);
    return 1;
  }
  V3<S> q = R.mul(v3(t[0], t[1], t[2]));
  V3<S> r = p - q;
  S len = sqrt_(dot(r, r));
  S depth = S(rad) - len;
  if (val(depth) < 0) return 0;
  c->pos = q + pb;
  c->n = (S(1) / len) * r;
  c->depth = depth;
  return 1;
}

template <class S> OD_HD inline int d_sphere_plane(const V3<S> &ps, float rad, const float *pl, CP<S> *c) {
  V3<S> n = v3(S(pl[0]), S(pl[1]), S(pl[2]));
  S depth = S(pl[3]) - dot(ps, n) + S(rad);
  if (val(depth) < 0) return 0;
  c->n = n;
  c->pos = ps - S(rad) * n;
  c->depth = depth;
  return 1;
}

template <class S>
OD_HD inline int d_box_plane(const V3<S> &pb, const M3<S> &R, const float *side, const float *pl, int maxc, CP<S> *c) {
  V3<S> n = v3(S(pl[0]), S(pl[1]), S(pl[2]));
  V3<S> ax[3] = {R.col(0), R.col(1), R.col(2)};
  S A[3] = {S(side[0]) * dot(n, ax[0]), S(side[1]) * dot(n, ax[1]), S(side[2]) * dot(n, ax[2])};
  S B[3] = {fabs_(A[0]), fabs_(A[1]), fabs_(A[2])};
  S depth = S(pl[3]) + S(0.5f) * (B[0] + B[1] + B[2]) - dot(n, pb);
  if (val(depth) < 0) return 0;
  if (maxc > 4) maxc = 4;
  V3<S> p = pb;
  for (int k = 0; k < 3; k++)
    p = (val(A[k]) > 0) ? p - S(0.5f * side[k]) * ax[k] : p + S(0.5f * side[k]) * ax[k];
  c[0].pos = p;
  c[0].depth = depth;
  int ret = 1;
  int first, second;
  float b0 = (float)val(B[0]), b1 = (float)val(B[1]), b2 = (float)val(B[2]);
  if (b0 < b1) {
    if (b2 < b0) { first = 2; second = 0; }
    else { first = 0; second = (b1 < b2) ? 1 : 2; }
  } else {
    if (b2 < b1) { first = 2; second = (b0 < b1) ? 0 : 1; }
    else { first = 1; second = (b0 < b2) ? 0 : 2; }
  }
  if (maxc > 1) {
    for (int s = 0; s < 2; s++) {
      if (s == 1 && maxc <= 2) break;
      int a = s == 0 ? first : second;
      if (val(depth) - val(B[a]) < 0) break;
      c[1 + s].pos = (val(A[a]) > 0) ? p + S(side[a]) * ax[a] : p - S(side[a]) * ax[a];
      c[1 + s].depth = depth - B[a];
      ret++;
    }
  }
  if (maxc == 4 && ret == 3) {
    S d4 = c[1].depth + c[2].depth - depth;
    if (val(d4) > 0) {
      c[3].pos = c[1].pos + c[2].pos - p;
      c[3].depth = d4;
      ret++;
    }
  }
  for (int j = 0; j < ret; j++) c[j].n = n;
  return ret;
}

template <class S>
OD_HD inline void d_line_closest(const V3<S> &pa, const V3<S> &ua, const V3<S> &pb, const V3<S> &ub, S &al, S &be) {
  V3<S> p = pb - pa;
  S uaub = dot(ua, ub);
  S q1 = dot(ua, p);
  S q2 = -dot(ub, p);
  S d = S(1) - uaub * uaub;
  if (val(d) <= 0.0001f) { al = S(0); be = S(0); return; }
  d = S(1) / d;
  al = (q1 + uaub * q2) * d;
  be = (uaub * q1 + q2) * d;
}

// rectangle clipping of a quad as clip_rect_quad, on generic scalars
template <class S> OD_HD inline int d_clip_rect_quad(float h0, float h1, const S *quad, S *ret) {
  S bufA[16], bufB[16];
  for (int i = 0; i < 8; i++) bufA[i] = quad[i];
  S *q = bufA, *r = bufB;
  int nq = 4, nr = 0;
  for (int chop = 0; chop < 4; chop++) {
    int dir = chop >> 1;
    float sign = (chop & 1) ? 1.f : -1.f;
    float h = dir ? h1 : h0;
    nr = 0;
    bool full = false;
    for (int i = 0; i < nq && !full; i++) {
      S ax = q[2 * i], ay = q[2 * i + 1];
      int j = (i + 1 < nq) ? i + 1 : 0;
      S bx = q[2 * j], by = q[2 * j + 1];
      S ad = dir ? ay : ax, bd = dir ? by : bx;
      bool ain = sign * val(ad) < h, bin = sign * val(bd) < h;
      if (ain) {
        r[2 * nr] = ax; r[2 * nr + 1] = ay; nr++;
        if (nr & 8) { full = true; break; }
      }
      if (ain != bin) {
        S ao = dir ? ax : ay, bo = dir ? bx : by;
        S o = ao + (bo - ao) / (bd - ad) * (S(sign * h) - ad);
        S d = S(sign * h);
        r[2 * nr] = dir ? o : d;
        r[2 * nr + 1] = dir ? d : o;
        nr++;
        if (nr & 8) { full = true; break; }
      }
    }
    S *t = q; q = r; r = t;
    nq = nr;
    if (full) break;
  }
  for (int i = 0; i < 2 * nq; i++) ret[i] = q[i];
  return nq;
}

// culling on the values only (selection of indices)
template <class T> OD_HD inline void d_cull_points(int n, const T *p, int m, int i0, int *iret) {
  T cx, cy;
  if (n == 1) { cx = p[0]; cy = p[1]; }
  else if (n == 2) { cx = T(0.5) * (p[0] + p[2]); cy = T(0.5) * (p[1] + p[3]); }
  else {
    T a = 0, q;
    cx = 0; cy = 0;
    for (int i = 0; i < n - 1; i++) {
      q = p[2 * i] * p[2 * i + 3] - p[2 * i + 2] * p[2 * i + 1];
      a += q;
      cx += q * (p[2 * i] + p[2 * i + 2]);
      cy += q * (p[2 * i + 1] + p[2 * i + 3]);
    }
    q = p[2 * n - 2] * p[1] - p[0] * p[2 * n - 1];
    a = T(1) / (T(3) * (a + q));
    cx = a * (cx + q * (p[2 * n - 2] + p[0]));
    cy = a * (cy + q * (p[2 * n - 1] + p[1]));
  }
  T ang[8];
  for (int i = 0; i < n; i++) ang[i] = std::atan2(p[2 * i + 1] - cy, p[2 * i] - cx);
  unsigned avail = (1u << n) - 1u;
  avail &= ~(1u << i0);
  iret[0] = i0;
  const T PI = T(3.14159265358979323846);
  for (int j = 1; j < m; j++) {
    T a = T(j) * (T(2) * PI / T(m)) + ang[i0];
    if (a > PI) a -= T(2) * PI;
    T best = T(1e9);
    int bi = i0;
    for (int i = 0; i < n; i++) {
      if (!(avail & (1u << i))) continue;
      T diff = std::fabs(ang[i] - a);
      if (diff > PI) diff = T(2) * PI - diff;
      if (diff < best) { best = diff; bi = i; }
    }
    iret[j] = bi;
    avail &= ~(1u << bi);
  }
}

template <class S>
OD_HD inline int d_box_box(const V3<S> &p1, const M3<S> &R1, const float *side1, const V3<S> &p2, const M3<S> &R2,
                           const float *side2, int maxc, CP<S> *out) {
  using T = decltype(val(p1.x));
  const T fudge = T(1.05);
  V3<S> u[3] = {R1.col(0), R1.col(1), R1.col(2)};
  V3<S> v[3] = {R2.col(0), R2.col(1), R2.col(2)};
  V3<S> p = p2 - p1;
  S pp[3] = {dot(u[0], p), dot(u[1], p), dot(u[2], p)};
  float A[3] = {0.5f * side1[0], 0.5f * side1[1], 0.5f * side1[2]};
  float B[3] = {0.5f * side2[0], 0.5f * side2[1], 0.5f * side2[2]};
  S Rm[3][3], Q[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      Rm[i][j] = dot(u[i], v[j]);
      Q[i][j] = fabs_(Rm[i][j]);
    }
  T s = -T(INFINITY);
  bool invert = false;
  int code = 0;
  V3<S> nrm = v3(S(0), S(0), S(0));
  S sS = S(0);   // separation of the chosen axis (differentiable copy of s)
  auto face = [&](const S &e1, const S &e2, const V3<S> &ax, int cc) -> bool {
    S s2 = fabs_(e1) - e2;
    if (val(s2) > 0) return false;
    if (val(s2) > s) { s = val(s2); sS = s2; nrm = ax; invert = val(e1) < 0; code = cc; }
    return true;
  };
  auto edge = [&](const S &e1, const S &e2, const S &n1, const S &n2, const S &n3, int cc) -> bool {
    S s2 = fabs_(e1) - e2;
    if (val(s2) > 0) return false;
    S l = sqrt_(n1 * n1 + n2 * n2 + n3 * n3);
    if (val(l) > 0) {
      s2 = s2 / l;
      if (val(s2) * fudge > s) {
        s = val(s2); sS = s2;
        nrm = v3(n1 / l, n2 / l, n3 / l);
        invert = val(e1) < 0;
        code = cc;
      }
    }
    return true;
  };
  const S z = S(0);
  if (!face(pp[0], S(A[0]) + B[0] * Q[0][0] + B[1] * Q[0][1] + B[2] * Q[0][2], u[0], 1)) return 0;
  if (!face(pp[1], S(A[1]) + B[0] * Q[1][0] + B[1] * Q[1][1] + B[2] * Q[1][2], u[1], 2)) return 0;
  if (!face(pp[2], S(A[2]) + B[0] * Q[2][0] + B[1] * Q[2][1] + B[2] * Q[2][2], u[2], 3)) return 0;
  if (!face(dot(v[0], p), A[0] * Q[0][0] + A[1] * Q[1][0] + A[2] * Q[2][0] + S(B[0]), v[0], 4)) return 0;
  if (!face(dot(v[1], p), A[0] * Q[0][1] + A[1] * Q[1][1] + A[2] * Q[2][1] + S(B[1]), v[1], 5)) return 0;
  if (!face(dot(v[2], p), A[0] * Q[0][2] + A[1] * Q[1][2] + A[2] * Q[2][2] + S(B[2]), v[2], 6)) return 0;
  if (!edge(pp[2] * Rm[1][0] - pp[1] * Rm[2][0], A[1] * Q[2][0] + A[2] * Q[1][0] + B[1] * Q[0][2] + B[2] * Q[0][1], z, -Rm[2][0], Rm[1][0], 7)) return 0;
  if (!edge(pp[2] * Rm[1][1] - pp[1] * Rm[2][1], A[1] * Q[2][1] + A[2] * Q[1][1] + B[0] * Q[0][2] + B[2] * Q[0][0], z, -Rm[2][1], Rm[1][1], 8)) return 0;
  if (!edge(pp[2] * Rm[1][2] - pp[1] * Rm[2][2], A[1] * Q[2][2] + A[2] * Q[1][2] + B[0] * Q[0][1] + B[1] * Q[0][0], z, -Rm[2][2], Rm[1][2], 9)) return 0;
  if (!edge(pp[0] * Rm[2][0] - pp[2] * Rm[0][0], A[0] * Q[2][0] + A[2] * Q[0][0] + B[1] * Q[1][2] + B[2] * Q[1][1], Rm[2][0], z, -Rm[0][0], 10)) return 0;
  if (!edge(pp[0] * Rm[2][1] - pp[2] * Rm[0][1], A[0] * Q[2][1] + A[2] * Q[0][1] + B[0] * Q[1][2] + B[2] * Q[1][0], Rm[2][1], z, -Rm[0][1], 11)) return 0;
  if (!edge(pp[0] * Rm[2][2] - pp[2] * Rm[0][2], A[0] * Q[2][2] + A[2] * Q[0][2] + B[0] * Q[1][1] + B[1] * Q[1][0], Rm[2][2], z, -Rm[0][2], 12)) return 0;
  if (!edge(pp[1] * Rm[0][0] - pp[0] * Rm[1][0], A[0] * Q[1][0] + A[1] * Q[0][0] + B[1] * Q[2][2] + B[2] * Q[2][1], -Rm[1][0], Rm[0][0], z, 13)) return 0;
  if (!edge(pp[1] * Rm[0][1] - pp[0] * Rm[1][1], A[0] * Q[1][1] + A[1] * Q[0][1] + B[0] * Q[2][2] + B[2] * Q[2][0], -Rm[1][1], Rm[0][1], z, 14)) return 0;
  if (!edge(pp[1] * Rm[0][2] - pp[0] * Rm[1][2], A[0] * Q[1][2] + A[1] * Q[0][2] + B[0] * Q[2][1] + B[1] * Q[2][0], -Rm[1][2], Rm[0][2], z, 15)) return 0;
  if (!code) return 0;
  V3<S> normal = (code <= 6) ? nrm : R1.mul(nrm);
  if (invert) normal = -normal;
  S depth = -sS;
  int n_out;
  if (code > 6) {
    V3<S> pa = p1;
    for (int k = 0; k < 3; k++) pa = pa + S((val(dot(normal, u[k])) > 0) ? A[k] : -A[k]) * u[k];
    V3<S> pb = p2;
    for (int k = 0; k < 3; k++) pb = pb + S((val(dot(normal, v[k])) > 0) ? -B[k] : B[k]) * v[k];
    V3<S> ua = u[(code - 7) / 3];
    V3<S> ub = v[(code - 7) % 3];
    S al, be;
    d_line_closest(pa, ua, pb, ub, al, be);
    pa = pa + al * ua;
    pb = pb + be * ub;
    out[0].pos = S(0.5f) * (pa + pb);
    out[0].depth = depth;
    n_out = 1;
  } else {
    bool aIs1 = code <= 3;
    const V3<S> *ra = aIs1 ? u : v;
    const V3<S> *rb = aIs1 ? v : u;
    V3<S> pa = aIs1 ? p1 : p2, pb = aIs1 ? p2 : p1;
    const float *Sa = aIs1 ? A : B;
    const float *Sb = aIs1 ? B : A;
    V3<S> n2 = aIs1 ? normal : -normal;
    S nr[3] = {dot(rb[0], n2), dot(rb[1], n2), dot(rb[2], n2)};
    T an0 = std::fabs(val(nr[0])), an1 = std::fabs(val(nr[1])), an2 = std::fabs(val(nr[2]));
    int lanr, a1, a2;
    if (an1 > an0) {
      if (an1 > an2) { a1 = 0; lanr = 1; a2 = 2; }
      else { a1 = 0; a2 = 1; lanr = 2; }
    } else {
      if (an0 > an2) { lanr = 0; a1 = 1; a2 = 2; }
      else { a1 = 0; a2 = 1; lanr = 2; }
    }
    V3<S> center = (val(nr[lanr]) < 0) ? (pb - pa) + S(Sb[lanr]) * rb[lanr] : (pb - pa) - S(Sb[lanr]) * rb[lanr];
    int codeN = aIs1 ? code - 1 : code - 4;
    int code1 = (codeN == 0) ? 1 : 0;
    int code2 = (codeN == 2) ? 1 : 2;
    V3<S> rc1 = ra[code1], rc2 = ra[code2];
    V3<S> rba1 = rb[a1], rba2 = rb[a2];
    S c1 = dot(center, rc1), c2 = dot(center, rc2);
    S m11 = dot(rc1, rba1), m12 = dot(rc1, rba2), m21 = dot(rc2, rba1), m22 = dot(rc2, rba2);
    S quad[8];
    {
      S k1 = m11 * Sb[a1], k2 = m21 * Sb[a1], k3 = m12 * Sb[a2], k4 = m22 * Sb[a2];
      quad[0] = c1 - k1 - k3; quad[1] = c2 - k2 - k4;
      quad[2] = c1 - k1 + k3; quad[3] = c2 - k2 + k4;
      quad[4] = c1 + k1 + k3; quad[5] = c2 + k2 + k4;
      quad[6] = c1 + k1 - k3; quad[7] = c2 + k2 - k4;
    }
    S ret[16];
    int n = d_clip_rect_quad(Sa[code1], Sa[code2], quad, ret);
    if (n < 1) return 0;
    S det1 = S(1) / (m11 * m22 - m12 * m21);
    m11 = m11 * det1; m12 = m12 * det1; m21 = m21 * det1; m22 = m22 * det1;
    V3<S> point[8];
    S dep[8];
    T retv[16];
    int cnum = 0;
    for (int j = 0; j < n; j++) {
      S k1 = m22 * (ret[2 * j] - c1) - m12 * (ret[2 * j + 1] - c2);
      S k2 = -m21 * (ret[2 * j] - c1) + m11 * (ret[2 * j + 1] - c2);
      V3<S> pt = center + k1 * rba1 + k2 * rba2;
      S d = S(Sa[codeN]) - dot(n2, pt);
      if (val(d) >= 0) {
        point[cnum] = pt;
        dep[cnum] = d;
        retv[2 * cnum] = val(ret[2 * j]);
        retv[2 * cnum + 1] = val(ret[2 * j + 1]);
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
      T md = val(dep[0]);
      for (int i = 1; i < cnum; i++)
        if (val(dep[i]) > md) { md = val(dep[i]); i1 = i; }
      int iret[8];
      d_cull_points(cnum, retv, mc, i1, iret);
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

// dCollide dispatch with the reversal rule of collide_pair
template <class S> OD_HD inline int d_collide(const GeomD<S> &a, const GeomD<S> &b, int maxc, CP<S> *c) {
  if (maxc <= 0) return 0;
  bool rev = a.type > b.type;
  const GeomD<S> &x = rev ? b : a;
  const GeomD<S> &y = rev ? a : b;
  int n = 0;
  if (x.type == GEOM_SPHERE) {
    if (y.type == GEOM_SPHERE) n = d_sphere_sphere(x.p, x.prm[0], y.p, y.prm[0], c);
    else if (y.type == GEOM_BOX) n = d_sphere_box(x.p, x.prm[0], y.p, y.R, y.prm, c);
    else n = d_sphere_plane(x.p, x.prm[0], y.prm, c);
  } else if (x.type == GEOM_BOX) {
    if (y.type == GEOM_BOX) n = d_box_box(x.p, x.R, x.prm, y.p, y.R, y.prm, maxc, c);
    else n = d_box_plane(x.p, x.R, x.prm, y.prm, maxc, c);
  }
  if (rev)
    for (int i = 0; i < n; i++) c[i].n = -c[i].n;
  return n;
}

}  // namespace adj
}  // namespace odemi
