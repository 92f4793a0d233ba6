// oracle_capi.cpp -- C entry points of the CPU restatement.
//
// TEST INFRASTRUCTURE (oracle).  Built twice by oracle/Makefile: with
// ORC_REAL=double (liboracle_f64.so) and ORC_REAL=float (liboracle_f32.so).
// Only tests/, the smoke check and the CPU baseline leg of bench.py load it.
//
// scene_run() follows the same SceneDesc/SceneOut contract as the scene
// driver (oracle/scene_driver.cpp), so the restatement, the reference build
// and the MI355X library are interchangeable behind one harness.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "orc_world.hpp"
#include "scene.h"

#ifndef ORC_REAL
#define ORC_REAL double
#endif
typedef ORC_REAL R;
using namespace orc;

namespace {

inline bool has(double v) { return !std::isnan(v); }

void apply_limot_params(Joint<R> &j, const double *p, bool hinge2) {
  auto set = [](LimMot<R> &lm, int which, double v) {
    if (!has(v)) return;
    R x = (R)v;
    switch (which) {  // dxJointLimitMotor::set, joint.cpp:472-509
      case SCJ_LO: lm.lostop = x; break;
      case SCJ_HI: lm.histop = x; break;
      case SCJ_VEL: lm.vel = x; break;
      case SCJ_FMAX: if (x >= 0) lm.fmax = x; break;
      case SCJ_FUDGE: if (x >= 0 && x <= 1) lm.fudge = x; break;
      case SCJ_BOUNCE: lm.bounce = x; break;
      case SCJ_CFM: lm.normal_cfm = x; break;
      case SCJ_STOP_ERP: lm.stop_erp = x; break;
      case SCJ_STOP_CFM: lm.stop_cfm = x; break;
    }
  };
  for (int k : {SCJ_LO, SCJ_HI, SCJ_VEL, SCJ_FMAX, SCJ_FUDGE, SCJ_BOUNCE, SCJ_CFM, SCJ_STOP_ERP, SCJ_STOP_CFM})
    set(j.limot, k, p[k]);
  if (hinge2) {
    if (has(p[SCJ_SUSP_ERP])) j.susp_erp = (R)p[SCJ_SUSP_ERP];
    if (has(p[SCJ_SUSP_CFM])) j.susp_cfm = (R)p[SCJ_SUSP_CFM];
    set(j.limot2, SCJ_VEL, p[SCJ_VEL2]);
    set(j.limot2, SCJ_FMAX, p[SCJ_FMAX2]);
  }
}

void record(const World<R> &w, double *dst) {
  for (size_t i = 0; i < w.bodies.size(); i++) {
    const Body<R> &b = w.bodies[i];
    double *r = dst + 13 * i;
    for (int k = 0; k < 3; k++) r[k] = b.pos[k];
    for (int k = 0; k < 4; k++) r[3 + k] = b.q[k];
    for (int k = 0; k < 3; k++) r[7 + k] = b.lvel[k];
    for (int k = 0; k < 3; k++) r[10 + k] = b.avel[k];
  }
}

}  // namespace

// 1: the following runs solve in the coloured order of the MI355X coloured
// mode instead of the reference order (not part of the reference).
static int g_colored = 0;
extern "C" void orc_set_colored(int on) { g_colored = on ? 1 : 0; }
// coloured mode: most members of a colour, the workgroup width of the MI355X
// launch being restated (default 256, the library's default width)
static int g_colour_width = 256;
extern "C" void orc_set_colour_width(int w) { g_colour_width = w > 0 ? w : 256; }
// coloured mode: when on, the colouring of the last coloured solve is kept,
// five ints per constraint (cid, b1, b2, colour, round) in the order the
// colours were assigned; orc_colour_log copies up to cap entries and returns
// the number kept
static int g_colour_log_on = 0;
static std::vector<int> g_colour_log;
extern "C" void orc_set_colour_log(int on) { g_colour_log_on = on; g_colour_log.clear(); }
extern "C" long long orc_colour_log(int *out, long long cap) {
  const long long n = (long long)g_colour_log.size() / 5;
  for (long long k = 0; k < 5 * std::min(n, cap); k++) out[k] = g_colour_log[k];
  return n;
}
// coloured mode: iterate to a residual tolerance, optionally with the
// conjugate-gradient acceleration; the sweep count of the runs since the
// last call is returned by orc_colored_sweeps
static double g_tol = 0;
static int g_maxit = 0, g_check = 10, g_accel = 0;
static double g_warm = 0;
static long long g_sweeps = 0, g_unconv = 0;
extern "C" void orc_set_colored_solver(double tol, int max_iterations, int check_every, int accel) {
  g_tol = tol; g_maxit = max_iterations; g_check = check_every > 0 ? check_every : 1; g_accel = accel;
}
// coloured mode: start each solve from warm x the previous step's impulses
extern "C" void orc_set_warm_start(double warm) { g_warm = warm; }
extern "C" long long orc_colored_sweeps(void) { long long s = g_sweeps; g_sweeps = 0; return s; }
// tolerance-mode solves that ended at the iteration cap since the last call
extern "C" long long orc_colored_unconverged(void) { long long s = g_unconv; g_unconv = 0; return s; }

// the coloured mode's pair order: (g1, g2) with g1 < g2, sorted
static void colored_pair_order(std::vector<std::pair<int, int>> &pairs) {
  for (auto &p : pairs)
    if (p.first > p.second) std::swap(p.first, p.second);
  std::sort(pairs.begin(), pairs.end());
}

// Optional body state applied after the world is built (joints keep the
// anchors of the scene's own poses): 13 doubles per body, pos q lvel avel.
static std::vector<double> g_state_override;
extern "C" void orc_set_state_override(const double *st, int nb) {
  if (st && nb > 0) g_state_override.assign(st, st + 13 * (size_t)nb);
  else g_state_override.clear();
}

// Builds the world exactly as the scene driver does through the ODE API.
static void build_world(const SceneDesc *s, World<R> &w) {
  w.prm.colored = g_colored;
  w.prm.colour_width = g_colour_width;
  if (g_colour_log_on) w.colour_log = &g_colour_log;
  w.prm.exact = s->world_step ? 1 : 0;
  w.prm.tol = (R)g_tol;
  w.prm.max_iterations = g_maxit;
  w.prm.check_every = g_check;
  w.prm.accel = g_accel;
  w.prm.warm = (R)g_warm;
  for (int k = 0; k < 3; k++) w.prm.gravity[k] = (R)s->gravity[k];
  w.prm.erp = (R)s->erp;
  w.prm.cfm = (R)s->cfm;
  w.prm.iterations = s->iterations;
  w.prm.w = (R)s->sor_w;
  w.prm.max_vel = (R)s->max_vel;
  w.prm.min_depth = (R)s->min_depth;
  w.prm.hash_min = s->hash_minlevel;
  w.prm.hash_max = s->hash_maxlevel;
  for (int i = 0; i < s->nb; i++) {
    int b = w.add_body();
    R I9[9];
    for (int k = 0; k < 9; k++) I9[k] = (R)s->body_I[9 * (size_t)i + k];
    w.set_mass(b, (R)s->body_mass[i], I9);
    for (int k = 0; k < 3; k++) w.bodies[b].pos[k] = (R)s->body_pos[3 * (size_t)i + k];
    R q[4];
    for (int k = 0; k < 4; k++) q[k] = (R)s->body_q[4 * (size_t)i + k];
    w.set_quat(b, q);
    for (int k = 0; k < 3; k++) w.bodies[b].lvel[k] = (R)s->body_lvel[3 * (size_t)i + k];
    for (int k = 0; k < 3; k++) w.bodies[b].avel[k] = (R)s->body_avel[3 * (size_t)i + k];
  }
  for (int i = 0; i < s->ng; i++) {
    R p[4];
    for (int k = 0; k < 4; k++) p[k] = (R)s->geom_param[4 * (size_t)i + k];
    int g = w.add_geom(s->geom_type[i], p);
    if (s->geom_space && s->geom_space[i] == 1) w.move_to_space2(g);
    if (s->geom_type[i] != SC_PLANE) {
      if (s->geom_body[i] >= 0) w.geom_set_body(g, s->geom_body[i]);
      else {
        for (int k = 0; k < 3; k++) w.geoms[g].pos[k] = (R)s->geom_pos[3 * (size_t)i + k];
        R q[4];
        for (int k = 0; k < 4; k++) q[k] = (R)s->geom_q[4 * (size_t)i + k];
        quat_to_rot(w.geoms[g].Rm, q);  // dGeomSetQuaternion, collision_kernel.cpp:599-613
      }
    }
  }
  for (int i = 0; i < s->nj; i++) {
    Joint<R> j;
    j.type = s->joint_type[i];
    j.erp = w.prm.erp;
    j.cfm = w.prm.cfm;
    j.limot.init(w.prm.erp, w.prm.cfm);
    j.limot2.init(w.prm.erp, w.prm.cfm);
    j.susp_erp = w.prm.erp;
    j.susp_cfm = w.prm.cfm;
    if (j.type == SC_HINGE2) { j.axis1[0] = 1; j.axis1[1] = 0; j.axis2[0] = 0; j.axis2[1] = 1; }
    w.joints.push_back(j);
    int ji = (int)w.joints.size() - 1;
    w.attach(ji, s->joint_b1[i], s->joint_b2[i]);
    Joint<R> &jr = w.joints[ji];
    const double *a = s->joint_anchor + 3 * (size_t)i;
    const double *x1 = s->joint_axis1 + 3 * (size_t)i;
    const double *x2 = s->joint_axis2 + 3 * (size_t)i;
    const double *prm = s->joint_param + SC_NJP * (size_t)i;
    if (jr.type == SC_BALL) {
      w.set_anchors(jr, (R)a[0], (R)a[1], (R)a[2]);
      if (has(prm[SCJ_ERP])) jr.erp = (R)prm[SCJ_ERP];
      if (has(prm[SCJ_CFM])) jr.cfm = (R)prm[SCJ_CFM];
    } else if (jr.type == SC_HINGE) {
      w.set_anchors(jr, (R)a[0], (R)a[1], (R)a[2]);
      w.hinge_qrel(jr);
      w.set_axes(jr, (R)x1[0], (R)x1[1], (R)x1[2], true, true);
      w.hinge_qrel(jr);
      apply_limot_params(jr, prm, false);
    } else {
      // hinge2 setters: hinge2.cpp:214-256 (the final axes decide s0, c0, v1, v2)
      w.set_anchors(jr, (R)a[0], (R)a[1], (R)a[2]);
      w.hinge2_make_v1v2(jr);
      R ax1[4], ax2[4], axc[4];
      w.set_axes(jr, (R)x1[0], (R)x1[1], (R)x1[2], true, false);
      w.hinge2_axis_info(jr, ax1, ax2, axc, jr.s0, jr.c0);
      w.hinge2_make_v1v2(jr);
      w.set_axes(jr, (R)x2[0], (R)x2[1], (R)x2[2], false, true);
      w.hinge2_axis_info(jr, ax1, ax2, axc, jr.s0, jr.c0);
      w.hinge2_make_v1v2(jr);
      apply_limot_params(jr, prm, true);
    }
  }
  w.n_perm = (int)w.joints.size();
  if (g_state_override.size() == 13 * (size_t)s->nb) {
    for (int i = 0; i < s->nb; i++) {
      const double *r = &g_state_override[13 * (size_t)i];
      for (int k = 0; k < 3; k++) w.bodies[i].pos[k] = (R)r[k];
      R q[4] = {(R)r[3], (R)r[4], (R)r[5], (R)r[6]};
      w.set_quat(i, q);
      for (int k = 0; k < 3; k++) { w.bodies[i].lvel[k] = (R)r[7 + k]; w.bodies[i].avel[k] = (R)r[10 + k]; }
    }
  }
}

extern "C" int scene_run(const SceneDesc *s, SceneOut *o) {
  using clk = std::chrono::steady_clock;
  World<R> w;
  build_world(s, w);
  Surface<R> surf;
  std::memset(&surf, 0, sizeof(surf));
  surf.mode = s->surf_mode;
  surf.mu = (R)s->mu;
  surf.mu2 = (R)s->mu2;
  surf.bounce = (R)s->bounce;
  surf.bounce_vel = (R)s->bounce_vel;
  surf.soft_erp = (R)s->soft_erp;
  surf.soft_cfm = (R)s->soft_cfm;
  o->total_pairs = 0;
  o->total_contacts = 0;
  o->sec_collide = 0;
  o->sec_step = 0;
  o->status = 0;
  if (o->feedback) w.fbout.assign(12 * (size_t)o->max_contacts, 0);
  long long npairs_diag = 0, ncont_diag = 0;
  int rec = 0;
  if (o->traj) record(w, o->traj);
  rec++;
  w.rng.seed = s->seed;
  std::vector<std::pair<int, int>> pairs;
  std::vector<ContactGeom<R>> cbuf(s->max_contacts > 0 ? s->max_contacts : 1);
  for (int step = 0; step < o->nsteps; step++) {
    bool diag = step < o->diag_steps;
    if (s->step_ctrl) {
      const double *ctl = s->step_ctrl + (size_t)step * s->nj * 4;
      for (int i = 0; i < s->nj; i++) {
        const double *u = ctl + 4 * (size_t)i;
        Joint<R> &j = w.joints[i];
        if (j.type == SC_HINGE2 || j.type == SC_HINGE) {
          if (has(u[0])) j.limot.lostop = (R)u[0];
          if (has(u[1])) j.limot.histop = (R)u[1];
        }
        if (j.type == SC_HINGE2 && (u[2] != 0 || u[3] != 0)) w.hinge2_add_torques(i, (R)u[2], (R)u[3]);
      }
    }
    auto t0 = clk::now();
    w.hash_pairs(pairs);
    if (g_colored) colored_pair_order(pairs);
    if (!w.space2.empty()) {
      // dSpaceCollide2(space 0, space 1) runs first in the frame
      std::vector<std::pair<int, int>> p2;
      w.collide2_pairs(p2);
      p2.insert(p2.end(), pairs.begin(), pairs.end());
      pairs.swap(p2);
    }
    int step_contacts = 0;
    for (auto &pr : pairs) {
      int g1 = pr.first, g2 = pr.second;
      if (diag && npairs_diag < o->max_pairs) {
        o->pairs[2 * npairs_diag] = g1;
        o->pairs[2 * npairs_diag + 1] = g2;
        npairs_diag++;
      }
      int b1 = w.geoms[g1].body, b2 = w.geoms[g2].body;
      if (s->skip_connected && b1 >= 0 && b2 >= 0 && w.connected_non_contact(b1, b2)) continue;
      int n = collide_geoms(w.shape(g1), g1, w.shape(g2), g2, s->max_contacts, cbuf.data());
      for (int i = 0; i < n; i++) {
        int ji = w.add_contact_joint(surf, cbuf[i], b1, b2);
        if (diag && ncont_diag < o->max_contacts) {
          double *r = o->contacts + 9 * ncont_diag;
          for (int k = 0; k < 3; k++) r[k] = cbuf[i].pos[k];
          for (int k = 0; k < 3; k++) r[3 + k] = cbuf[i].normal[k];
          r[6] = cbuf[i].depth;
          r[7] = cbuf[i].g1;
          r[8] = cbuf[i].g2;
          if (o->feedback) w.joints[ji].fb = ncont_diag;
          ncont_diag++;
        }
        step_contacts++;
      }
    }
    auto t1 = clk::now();
    w.track_residual = o->residual != nullptr;
    w.quickstep((R)s->dt);
    if (o->residual) o->residual[step] = (double)(g_colored ? w.last_res : w.step_res);
    w.empty_contacts();
    auto t2 = clk::now();
    if (step >= o->time_from) {
      o->sec_collide += std::chrono::duration<double>(t1 - t0).count();
      o->sec_step += std::chrono::duration<double>(t2 - t1).count();
    }
    if (diag) {
      if (o->pair_step_count) o->pair_step_count[step] = (int)pairs.size();
      if (o->contact_step_count) o->contact_step_count[step] = step_contacts;
    }
    o->total_pairs += (long long)pairs.size();
    o->total_contacts += step_contacts;
    if (o->traj && o->record_every > 0 && (step + 1) % o->record_every == 0) {
      record(w, o->traj + (size_t)rec * s->nb * 13);
      rec++;
    }
  }
  if (o->feedback)
    for (size_t k = 0; k < 12 * (size_t)o->max_contacts; k++) o->feedback[k] = (double)w.fbout[k];
  g_sweeps += w.prm.sweeps;
  g_unconv += w.prm.unconverged;
  return o->status;
}

extern "C" int scene_dreal_size(void) { return (int)sizeof(R); }

// ---------------------------------------------------------------- components

// dTestRand known answer (ode/src/misc.cpp:52-62)
extern "C" int orc_rand_test(void) {
  Rng r;
  r.seed = 0;
  const unsigned long want[5] = {0x3c6ef35fUL, 0x47502932UL, 0xd1ccf6e9UL, 0xaaf95334UL, 0x6252e503UL};
  for (int i = 0; i < 5; i++)
    if (r.next() != want[i]) return 0;
  return 1;
}

// n draws of dRandInt(bound[i]) from `seed`; writes the results and returns
// the final seed.
extern "C" unsigned long orc_rand_ints(unsigned long seed, int n, const int *bound, int *out) {
  Rng r;
  r.seed = seed;
  for (int i = 0; i < n; i++) out[i] = r.below(bound[i]);
  return r.seed;
}

// One dCollide call between two geoms given their type, pose and shape.
// pos: 3 doubles, rot: 12 doubles (3x4), param: 4 doubles.  Writes up to
// (flags & 0xffff) contacts of 7 doubles (pos3 normal3 depth).
extern "C" int orc_collide(int t1, const double *pos1, const double *rot1, const double *prm1, int t2,
                           const double *pos2, const double *rot2, const double *prm2, int flags, double *out) {
  R p1[4], p2[4], r1[12], r2[12];
  GeomShape<R> a, b;
  a.type = t1; b.type = t2;
  for (int k = 0; k < 3; k++) { p1[k] = (R)pos1[k]; p2[k] = (R)pos2[k]; }
  for (int k = 0; k < 12; k++) { r1[k] = (R)rot1[k]; r2[k] = (R)rot2[k]; }
  for (int k = 0; k < 4; k++) { a.param[k] = (R)prm1[k]; b.param[k] = (R)prm2[k]; }
  a.pos = p1; a.Rm = r1; b.pos = p2; b.Rm = r2;
  int maxc = flags & 0xffff;
  std::vector<ContactGeom<R>> c(maxc > 0 ? maxc : 1);
  int n = collide_geoms(a, 0, b, 1, flags, c.data());
  for (int i = 0; i < n; i++) {
    for (int k = 0; k < 3; k++) out[7 * i + k] = c[i].pos[k];
    for (int k = 0; k < 3; k++) out[7 * i + 3 + k] = c[i].normal[k];
    out[7 * i + 6] = c[i].depth;
  }
  return n;
}

// AABB of one geom (6 doubles).
extern "C" void orc_aabb(int t, const double *pos, const double *rot, const double *prm, double *out) {
  R p[4], r[12];
  GeomShape<R> g;
  g.type = t;
  for (int k = 0; k < 3; k++) p[k] = (R)pos[k];
  for (int k = 0; k < 12; k++) r[k] = (R)rot[k];
  for (int k = 0; k < 4; k++) g.param[k] = (R)prm[k];
  g.pos = p;
  g.Rm = r;
  R a[6];
  geom_aabb(g, a);
  for (int k = 0; k < 6; k++) out[k] = a[k];
}

// dPlaneSpace of a unit normal.
extern "C" void orc_plane_space(const double *n, double *p, double *q) {
  R nn[3] = {(R)n[0], (R)n[1], (R)n[2]}, pp[3], qq[3];
  plane_space(nn, pp, qq);
  for (int k = 0; k < 3; k++) { p[k] = pp[k]; q[k] = qq[k]; }
}

// Step-by-step record of a run for the adjoint's checks: per step the body
// state before the step (13 doubles per body: pos, q, lvel, avel) and every
// contact of the step in creation order, 16 doubles each: pos 3, normal 3
// (as the collider returned it), depth, impulses of its rows 3, g1, g2, b1,
// b2 (after attach), flags, rows.  states holds (nsteps+1) records (the
// last is the final state).  Returns the total contact count, or -1 when
// max_total is exceeded.
// The adjoint's record of permanent joint j after a step (JREC doubles):
// impulses[6], bounds lo[6] and hi[6], cfm/h[6], anchor1[3], anchor2[3],
// axis1[3], axis2[3] (body frames), Info2 erp, suspension erp, c0, s0, v1[3],
// v2[3], per limit/motor k: stop erp[2], stop[2], motor velocity[2]; the
// hinge2 torques of the step[2]; type, body 1, body 2, rows, row of limot
// k[2], its kind[2] (1 limit, 2 motor), 1 when the adjoint covers every row.
// Row bounds and cfm follow the joints' getInfo2 (hinge.cpp:76-149,
// hinge2.cpp:127-180, joint.cpp:556-733).
static const int JREC = 64;
static void joint_record(World<R> &w, int ji, double h, const double *ctl, double *o) {
  const Joint<R> &j = w.joints[ji];
  for (int k = 0; k < JREC; k++) o[k] = 0;
  const int m = j.nrows;
  for (int r = 0; r < 6; r++) {
    o[r] = r < m ? (double)j.lam[r] : 0.0;
    o[6 + r] = -INFINITY;
    o[12 + r] = INFINITY;
    o[18 + r] = (double)w.prm.cfm / h;
  }
  for (int k = 0; k < 3; k++) {
    o[24 + k] = (double)j.anchor1[k];
    o[27 + k] = (double)j.anchor2[k];
    o[30 + k] = (double)j.axis1[k];
    o[33 + k] = (double)j.axis2[k];
    o[40 + k] = (double)j.v1[k];
    o[43 + k] = (double)j.v2[k];
  }
  o[36] = (double)j.used_erp;
  o[37] = (double)j.susp_erp;
  o[38] = (double)j.c0;
  o[39] = (double)j.s0;
  o[54] = j.type;
  o[55] = j.node0;
  o[56] = j.node1;
  o[57] = m;
  o[58] = o[59] = -1;
  bool covered = true;
  int row = 3;
  if (j.type == SC_BALL) {
    for (int r = 0; r < 3; r++) o[18 + r] = (double)j.cfm / h;
  } else if (j.type == SC_HINGE) {
    row = 5;
  } else {
    o[18] = (double)j.susp_cfm / h;
    row = 4;
  }
  for (int l = 0; l < 2 && j.type != SC_BALL; l++) {
    const LimMot<R> &lm = l == 0 ? j.limot : j.limot2;
    if (l == 1 && j.type != SC_HINGE2) break;
    bool powered = lm.fmax > 0;
    if (!(powered || lm.limit)) continue;
    if (row >= m) { covered = false; break; }
    o[58 + l] = row;
    if (lm.limit && lm.lostop == lm.histop) powered = false;
    if (powered && !lm.limit) {
      o[60 + l] = 2;
      o[50 + l] = (double)lm.vel;
      o[6 + row] = -(double)lm.fmax;
      o[12 + row] = (double)lm.fmax;
      o[18 + row] = (double)lm.normal_cfm / h;
    } else {
      // a limit row; a motor pushing against a one-sided limit, a bouncy
      // limit and the hinge's angle are not covered
      if (powered || (lm.bounce > 0 && lm.lostop != lm.histop) || j.type == SC_HINGE) covered = false;
      o[60 + l] = 1;
      o[46 + l] = (double)lm.stop_erp;
      o[48 + l] = (double)(lm.limit == 1 ? lm.lostop : lm.histop);
      o[18 + row] = (double)lm.stop_cfm / h;
      if (lm.lostop != lm.histop) {
        if (lm.limit == 1) o[6 + row] = 0;
        else o[12 + row] = 0;
      }
    }
    row++;
  }
  if (j.type == SC_HINGE2 && ctl) {
    o[52] = ctl[2];
    o[53] = ctl[3];
  }
  o[62] = covered ? 1 : 0;
}

extern "C" long orc_tape_run(const SceneDesc *s, int nsteps, double *states, int *nk, double *contacts,
                             long max_total, double *jrec) {
  World<R> w;
  build_world(s, w);
  Surface<R> surf;
  std::memset(&surf, 0, sizeof(surf));
  surf.mode = s->surf_mode;
  surf.mu = (R)s->mu;
  surf.mu2 = (R)s->mu2;
  surf.bounce = (R)s->bounce;
  surf.bounce_vel = (R)s->bounce_vel;
  surf.soft_erp = (R)s->soft_erp;
  surf.soft_cfm = (R)s->soft_cfm;
  w.rng.seed = s->seed;
  std::vector<std::pair<int, int>> pairs;
  std::vector<ContactGeom<R>> cbuf(s->max_contacts > 0 ? s->max_contacts : 1);
  long total = 0;
  for (int step = 0; step < nsteps; step++) {
    record(w, states + (size_t)step * s->nb * 13);
    if (s->step_ctrl) {
      // per-step joint controls, as scene_run applies them
      const double *ctl = s->step_ctrl + (size_t)step * s->nj * 4;
      for (int i = 0; i < s->nj; i++) {
        const double *u = ctl + 4 * (size_t)i;
        Joint<R> &j = w.joints[i];
        if (j.type == SC_HINGE2 || j.type == SC_HINGE) {
          if (has(u[0])) j.limot.lostop = (R)u[0];
          if (has(u[1])) j.limot.histop = (R)u[1];
        }
        if (j.type == SC_HINGE2 && (u[2] != 0 || u[3] != 0)) w.hinge2_add_torques(i, (R)u[2], (R)u[3]);
      }
    }
    w.hash_pairs(pairs);
    if (g_colored) colored_pair_order(pairs);
    std::vector<int> cj;
    std::vector<ContactGeom<R>> cg;
    for (auto &pr : pairs) {
      int g1 = pr.first, g2 = pr.second;
      int b1 = w.geoms[g1].body, b2 = w.geoms[g2].body;
      if (s->skip_connected && b1 >= 0 && b2 >= 0 && w.connected_non_contact(b1, b2)) continue;
      int n = collide_geoms(w.shape(g1), g1, w.shape(g2), g2, s->max_contacts, cbuf.data());
      for (int i = 0; i < n; i++) {
        cj.push_back(w.add_contact_joint(surf, cbuf[i], b1, b2));
        cg.push_back(cbuf[i]);
      }
    }
    w.quickstep((R)s->dt);
    nk[step] = (int)cj.size();
    for (int j = 0; jrec && j < w.n_perm; j++)
      joint_record(w, j, s->dt, s->step_ctrl ? s->step_ctrl + ((size_t)step * s->nj + j) * 4 : nullptr,
                   jrec + JREC * ((size_t)step * w.n_perm + j));
    if (total + (long)cj.size() > max_total) return -1;
    for (size_t i = 0; i < cj.size(); i++) {
      const Joint<R> &j = w.joints[cj[i]];
      double *o = contacts + 16 * (size_t)(total + (long)i);
      for (int k = 0; k < 3; k++) o[k] = cg[i].pos[k];
      for (int k = 0; k < 3; k++) o[3 + k] = cg[i].normal[k];
      o[6] = cg[i].depth;
      for (int k = 0; k < 3; k++) o[7 + k] = k < j.nrows ? (double)j.lam[k] : 0.0;
      o[10] = cg[i].g1;
      o[11] = cg[i].g2;
      o[12] = j.node0;
      o[13] = j.node1;
      o[14] = j.flags;
      o[15] = j.nrows;
    }
    total += (long)cj.size();
    w.empty_contacts();
  }
  record(w, states + (size_t)nsteps * s->nb * 13);
  return total;
}

// The exact box-constrained LCP of the dWorldStep restatement on one dense
// problem (A n x n row-major, symmetric positive definite): x in [lo,hi],
// w = A x - b complementary to the bounds.  Returns the pivoting rounds, or
// -1 at the cap or for an A that is not positive definite.
extern "C" int orc_blcp(int n, const double *A, const double *b, const double *lo, const double *hi, double *x) {
  std::vector<R> a(A, A + (size_t)n * n), bb(b, b + n), l(lo, lo + n), h(hi, hi + n), xx(n, 0);
  std::vector<char> in(n, 1);
  std::vector<int> st(n, 0);
  int r = World<R>::blcp(n, a, bb, l, h, in, xx, st);
  for (int i = 0; i < n; i++) x[i] = (double)xx[i];
  return r;
}
