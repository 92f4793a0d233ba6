// scene_driver.cpp -- an ordinary ODE-API application that builds a SceneDesc
// and runs the canonical frame loop (collide -> contact joints -> QuickStep ->
// empty the contact group), recording trajectories, the near-callback pair
// order, contacts and contact-joint feedback.
//
// Test infrastructure.  The same source is compiled
//   * against the reference headers and sources (oracle/ref.mk), and
//   * against the MI355X library's <ode/ode.h> (tests: the drop-in check),
// so both run literally the same application code.
//
// The frame loop is the one of the reference demos
// (ode/demo/demo_space_stress.cpp:74-104 near callback, 370-376 frame;
// demos/raycar/ode.cpp:67-130).
#include <ode/ode.h>
#include "scene.h"

#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Ctx {
  const SceneDesc *s;
  SceneOut *o;
  dWorldID world;
  dJointGroupID cgroup;
  std::vector<dContact> buf;
  std::vector<dJointFeedback> fb;
  int step;
  bool diag;
  long long npairs_diag, ncont_diag;
  int step_pairs, step_contacts;
  std::vector<dJointID> fbjoints;
};

inline int geom_index(dGeomID g) { return (int)(intptr_t)dGeomGetData(g) - 1; }

void near_callback(void *data, dGeomID o1, dGeomID o2) {
  // a space met in dSpaceCollide2: descend (demos/raycar/ode.cpp:50-61)
  if (dGeomIsSpace(o1)) {
    dSpaceCollide2(o1, o2, data, &near_callback);
    return;
  }
  if (dGeomIsSpace(o2)) {
    dSpaceCollide2(o2, o1, data, &near_callback);
    return;
  }
  Ctx *c = (Ctx *)data;
  const SceneDesc *s = c->s;
  SceneOut *o = c->o;
  if (c->diag && c->npairs_diag < o->max_pairs) {
    o->pairs[2 * c->npairs_diag + 0] = geom_index(o1);
    o->pairs[2 * c->npairs_diag + 1] = geom_index(o2);
    c->npairs_diag++;
  }
  c->step_pairs++;
  dBodyID b1 = dGeomGetBody(o1), b2 = dGeomGetBody(o2);
  if (s->skip_connected && b1 && b2 && dAreConnectedExcluding(b1, b2, dJointTypeContact)) return;
  int maxc = s->max_contacts;
  if ((int)c->buf.size() < maxc) c->buf.resize(maxc);
  int n = dCollide(o1, o2, maxc, &c->buf[0].geom, sizeof(dContact));
  for (int i = 0; i < n; i++) {
    dContact &ct = c->buf[i];
    ct.surface.mode = s->surf_mode;
    ct.surface.mu = (dReal)s->mu;
    ct.surface.mu2 = (dReal)s->mu2;
    ct.surface.bounce = (dReal)s->bounce;
    ct.surface.bounce_vel = (dReal)s->bounce_vel;
    ct.surface.soft_erp = (dReal)s->soft_erp;
    ct.surface.soft_cfm = (dReal)s->soft_cfm;
    dJointID j = dJointCreateContact(c->world, c->cgroup, &ct);
    dJointAttach(j, b1, b2);
    if (c->diag && c->ncont_diag < o->max_contacts) {
      double *r = o->contacts + 9 * c->ncont_diag;
      for (int k = 0; k < 3; k++) r[k] = ct.geom.pos[k];
      for (int k = 0; k < 3; k++) r[3 + k] = ct.geom.normal[k];
      r[6] = ct.geom.depth;
      r[7] = geom_index(ct.geom.g1);
      r[8] = geom_index(ct.geom.g2);
      if (o->feedback) {
        dJointSetFeedback(j, &c->fb[c->ncont_diag]);
        c->fbjoints.push_back(j);
      }
      c->ncont_diag++;
    }
    c->step_contacts++;
  }
}

inline bool has(double v) { return !std::isnan(v); }

void set_limot(dJointID j, int type, const double *p) {
  auto setp = [&](int param, double v) {
    if (!has(v)) return;
    if (type == SC_HINGE) dJointSetHingeParam(j, param, (dReal)v);
    else if (type == SC_HINGE2) dJointSetHinge2Param(j, param, (dReal)v);
  };
  setp(dParamLoStop, p[SCJ_LO]);
  setp(dParamHiStop, p[SCJ_HI]);
  setp(dParamVel, p[SCJ_VEL]);
  setp(dParamFMax, p[SCJ_FMAX]);
  setp(dParamFudgeFactor, p[SCJ_FUDGE]);
  setp(dParamBounce, p[SCJ_BOUNCE]);
  setp(dParamCFM, p[SCJ_CFM]);
  setp(dParamStopERP, p[SCJ_STOP_ERP]);
  setp(dParamStopCFM, p[SCJ_STOP_CFM]);
  if (type == SC_HINGE2) {
    setp(dParamSuspensionERP, p[SCJ_SUSP_ERP]);
    setp(dParamSuspensionCFM, p[SCJ_SUSP_CFM]);
    setp(dParamVel2, p[SCJ_VEL2]);
    setp(dParamFMax2, p[SCJ_FMAX2]);
  }
}

void record_state(const SceneDesc *s, dBodyID *bodies, double *dst) {
  for (int i = 0; i < s->nb; i++) {
    const dReal *p = dBodyGetPosition(bodies[i]);
    const dReal *q = dBodyGetQuaternion(bodies[i]);
    const dReal *v = dBodyGetLinearVel(bodies[i]);
    const dReal *w = dBodyGetAngularVel(bodies[i]);
    double *r = dst + 13 * (size_t)i;
    for (int k = 0; k < 3; k++) r[k] = p[k];
    for (int k = 0; k < 4; k++) r[3 + k] = q[k];
    for (int k = 0; k < 3; k++) r[7 + k] = v[k];
    for (int k = 0; k < 3; k++) r[10 + k] = w[k];
  }
}

}  // namespace

extern "C" int scene_run(const SceneDesc *s, SceneOut *o) {
  using clk = std::chrono::steady_clock;
  dInitODE2(0);
  dWorldID world = dWorldCreate();
  dWorldSetGravity(world, (dReal)s->gravity[0], (dReal)s->gravity[1], (dReal)s->gravity[2]);
  dWorldSetERP(world, (dReal)s->erp);
  dWorldSetCFM(world, (dReal)s->cfm);
  dWorldSetQuickStepNumIterations(world, s->iterations);
  dWorldSetQuickStepW(world, (dReal)s->sor_w);
  dWorldSetContactMaxCorrectingVel(world, (dReal)s->max_vel);
  dWorldSetContactSurfaceLayer(world, (dReal)s->min_depth);
  dSpaceID space = dHashSpaceCreate(0);
  dHashSpaceSetLevels(space, s->hash_minlevel, s->hash_maxlevel);
  // a static terrain space collided against the dynamic one
  // (demos/raycar/ode.cpp:42-48, 118-123)
  dSpaceID space2 = 0;
  for (int i = 0; s->geom_space && i < s->ng && !space2; i++)
    if (s->geom_space[i] == 1) space2 = dHashSpaceCreate(0);
  if (space2) dHashSpaceSetLevels(space2, s->hash_minlevel, s->hash_maxlevel);
  dJointGroupID cgroup = dJointGroupCreate(0);

  std::vector<dBodyID> bodies(s->nb);
  for (int i = 0; i < s->nb; i++) {
    dBodyID b = dBodyCreate(world);
    const double *I = s->body_I + 9 * (size_t)i;
    dMass m;
    dMassSetZero(&m);
    dMassSetParameters(&m, (dReal)s->body_mass[i], 0, 0, 0, (dReal)I[0], (dReal)I[4], (dReal)I[8],
                       (dReal)I[1], (dReal)I[2], (dReal)I[5]);
    dBodySetMass(b, &m);
    const double *p = s->body_pos + 3 * (size_t)i;
    const double *q = s->body_q + 4 * (size_t)i;
    dBodySetPosition(b, (dReal)p[0], (dReal)p[1], (dReal)p[2]);
    dQuaternion qq = {(dReal)q[0], (dReal)q[1], (dReal)q[2], (dReal)q[3]};
    dBodySetQuaternion(b, qq);
    const double *v = s->body_lvel + 3 * (size_t)i;
    const double *w = s->body_avel + 3 * (size_t)i;
    dBodySetLinearVel(b, (dReal)v[0], (dReal)v[1], (dReal)v[2]);
    dBodySetAngularVel(b, (dReal)w[0], (dReal)w[1], (dReal)w[2]);
    bodies[i] = b;
  }
  for (int i = 0; i < s->ng; i++) {
    const double *p = s->geom_param + 4 * (size_t)i;
    dSpaceID gs = (space2 && s->geom_space[i] == 1) ? space2 : space;
    dGeomID g = 0;
    switch (s->geom_type[i]) {
      case SC_SPHERE: g = dCreateSphere(gs, (dReal)p[0]); break;
      case SC_BOX: g = dCreateBox(gs, (dReal)p[0], (dReal)p[1], (dReal)p[2]); break;
      case SC_PLANE: g = dCreatePlane(gs, (dReal)p[0], (dReal)p[1], (dReal)p[2], (dReal)p[3]); break;
      case SC_CAPSULE: g = dCreateCapsule(gs, (dReal)p[0], (dReal)p[1]); break;
      case SC_CYLINDER: g = dCreateCylinder(gs, (dReal)p[0], (dReal)p[1]); break;
    }
    dGeomSetData(g, (void *)(intptr_t)(i + 1));
    if (s->geom_type[i] != SC_PLANE) {
      if (s->geom_body[i] >= 0) {
        dGeomSetBody(g, bodies[s->geom_body[i]]);
      } else {
        const double *gp = s->geom_pos + 3 * (size_t)i;
        const double *gq = s->geom_q + 4 * (size_t)i;
        dGeomSetPosition(g, (dReal)gp[0], (dReal)gp[1], (dReal)gp[2]);
        dQuaternion qq = {(dReal)gq[0], (dReal)gq[1], (dReal)gq[2], (dReal)gq[3]};
        dGeomSetQuaternion(g, qq);
      }
    }
  }
  std::vector<dJointID> joints(s->nj);
  for (int i = 0; i < s->nj; i++) {
    dJointID j = 0;
    int t = s->joint_type[i];
    if (t == SC_BALL) j = dJointCreateBall(world, 0);
    else if (t == SC_HINGE) j = dJointCreateHinge(world, 0);
    else j = dJointCreateHinge2(world, 0);
    dBodyID b1 = bodies[s->joint_b1[i]];
    dBodyID b2 = s->joint_b2[i] >= 0 ? bodies[s->joint_b2[i]] : 0;
    dJointAttach(j, b1, b2);
    const double *a = s->joint_anchor + 3 * (size_t)i;
    const double *x1 = s->joint_axis1 + 3 * (size_t)i;
    const double *x2 = s->joint_axis2 + 3 * (size_t)i;
    const double *prm = s->joint_param + SC_NJP * (size_t)i;
    if (t == SC_BALL) {
      dJointSetBallAnchor(j, (dReal)a[0], (dReal)a[1], (dReal)a[2]);
      if (has(prm[SCJ_ERP])) dJointSetBallParam(j, dParamERP, (dReal)prm[SCJ_ERP]);
      if (has(prm[SCJ_CFM])) dJointSetBallParam(j, dParamCFM, (dReal)prm[SCJ_CFM]);
    } else if (t == SC_HINGE) {
      dJointSetHingeAnchor(j, (dReal)a[0], (dReal)a[1], (dReal)a[2]);
      dJointSetHingeAxis(j, (dReal)x1[0], (dReal)x1[1], (dReal)x1[2]);
      set_limot(j, t, prm);
    } else {
      dJointSetHinge2Anchor(j, (dReal)a[0], (dReal)a[1], (dReal)a[2]);
      dJointSetHinge2Axis1(j, (dReal)x1[0], (dReal)x1[1], (dReal)x1[2]);
      dJointSetHinge2Axis2(j, (dReal)x2[0], (dReal)x2[1], (dReal)x2[2]);
      set_limot(j, t, prm);
    }
    joints[i] = j;
  }

  Ctx c;
  c.s = s;
  c.o = o;
  c.world = world;
  c.cgroup = cgroup;
  c.buf.resize(s->max_contacts > 0 ? s->max_contacts : 1);
  c.fb.resize(o->feedback ? (size_t)o->max_contacts : 0);
  c.npairs_diag = 0;
  c.ncont_diag = 0;
  o->total_pairs = 0;
  o->total_contacts = 0;
  o->sec_collide = 0;
  o->sec_step = 0;
  o->status = 0;

  int rec = 0;
  if (o->traj) record_state(s, bodies.data(), o->traj);
  rec++;
  dRandSetSeed(s->seed);
  for (int step = 0; step < o->nsteps; step++) {
    c.step = step;
    c.diag = step < o->diag_steps;
    c.step_pairs = 0;
    c.step_contacts = 0;
    c.fbjoints.clear();
    long long fb_start = c.ncont_diag;
    if (s->step_ctrl) {
      const double *ctl = s->step_ctrl + (size_t)step * s->nj * 4;
      for (int i = 0; i < s->nj; i++) {
        const double *u = ctl + 4 * (size_t)i;
        int t = s->joint_type[i];
        if (t == SC_HINGE2) {
          if (has(u[0])) dJointSetHinge2Param(joints[i], dParamLoStop, (dReal)u[0]);
          if (has(u[1])) dJointSetHinge2Param(joints[i], dParamHiStop, (dReal)u[1]);
          if (u[2] != 0 || u[3] != 0) dJointAddHinge2Torques(joints[i], (dReal)u[2], (dReal)u[3]);
        } else if (t == SC_HINGE) {
          if (has(u[0])) dJointSetHingeParam(joints[i], dParamLoStop, (dReal)u[0]);
          if (has(u[1])) dJointSetHingeParam(joints[i], dParamHiStop, (dReal)u[1]);
        }
      }
    }
#ifndef SCENE_DRIVER_REFERENCE
    if (o->grad_out && step == o->nsteps - 1) dWorldSetTape(world, 1);
#endif
    auto t0 = clk::now();
    if (space2) dSpaceCollide2((dGeomID)space, (dGeomID)space2, &c, &near_callback);
    dSpaceCollide(space, &c, &near_callback);
    auto t1 = clk::now();
    int ok = s->world_step ? dWorldStep(world, (dReal)s->dt) : dWorldQuickStep(world, (dReal)s->dt);
    if (c.diag && o->feedback) {
      for (size_t k = 0; k < c.fbjoints.size(); k++) {
        const dJointFeedback &f = c.fb[fb_start + k];
        double *r = o->feedback + 12 * (fb_start + k);
        for (int m = 0; m < 3; m++) {
          r[m] = f.f1[m];
          r[3 + m] = f.t1[m];
          r[6 + m] = f.f2[m];
          r[9 + m] = f.t2[m];
        }
      }
    }
    dJointGroupEmpty(cgroup);
    auto t2 = clk::now();
    if (step >= o->time_from) {
      o->sec_collide += std::chrono::duration<double>(t1 - t0).count();
      o->sec_step += std::chrono::duration<double>(t2 - t1).count();
    }
    if (!ok && !o->status) o->status = step + 1;
    if (c.diag) {
      if (o->pair_step_count) o->pair_step_count[step] = c.step_pairs;
      if (o->contact_step_count) o->contact_step_count[step] = c.step_contacts;
    }
    o->total_pairs += c.step_pairs;
    o->total_contacts += c.step_contacts;
    if (o->traj && o->record_every > 0 && (step + 1) % o->record_every == 0) {
      record_state(s, bodies.data(), o->traj + (size_t)rec * s->nb * 13);
      rec++;
    }
  }

#ifndef SCENE_DRIVER_REFERENCE
  if (o->grad_out && o->grad_in && o->nsteps > 0) {
    // reverse step of the last dWorldQuickStep (an addition of this library)
    std::vector<dReal> go(13 * (size_t)s->nb), gi(13 * (size_t)s->nb, 0);
    for (size_t k = 0; k < go.size(); k++) go[k] = (dReal)o->grad_out[k];
    if (!dWorldQuickStepBackward(world, go.data(), gi.data()) && !o->status) o->status = o->nsteps + 1;
    for (size_t k = 0; k < gi.size(); k++) o->grad_in[k] = gi[k];
  }
#endif
  dJointGroupDestroy(cgroup);
  dSpaceDestroy(space);
  if (space2) dSpaceDestroy(space2);
  dWorldDestroy(world);
  dCloseODE();
  return o->status;
}

// dReal size of the library this driver was compiled against (4 or 8).
extern "C" int scene_dreal_size(void) { return (int)sizeof(dReal); }
