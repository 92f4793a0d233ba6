// ode_api.cpp -- the ODE 0.12 C API subset on top of the device pipeline.
//
// The host keeps the object model the API exposes (worlds, bodies, spaces,
// geoms, joints, joint groups) with the reference's list semantics; each world
// owns a device mirror (a one-world batch).  dSpaceCollide runs the broadphase
// and a speculative narrowphase on the GPU, then invokes the user's near
// callback for every pair in the reference's report order; dCollide inside the
// callback returns the contacts the GPU computed for that pair (or launches a
// one-pair collide for a different request).  dWorldQuickStep uploads the
// contact joints the callback created and runs the step kernel in the
// reference order (islands, row order, seeded shuffle).
#include <algorithm>
#include <cstdarg>
#include <cstdlib>

#include "adj_step.hpp"
#include "colliders.hpp"
#include "host_common.hpp"
#include "odemi.h"
#include "odemi_internal.hpp"

namespace odemi {
unsigned long &global_seed();
void report_debug(const char *fmt, ...);
hipError_t launch_collide_one(const GeomView &a, const GeomView &b, int maxc, CPoint *out, int *n, hipStream_t st);
void *noalloc_scratch(hipError_t *err);
hipError_t launch_collide_list(const GeomView *a, const GeomView *b, int npairs, int maxc, CPoint *out, int *n,
                               hipStream_t st);
hipError_t launch_api_tape_capture(const DevBatch &B, const ContactIn *cin, int nk, adj::TapeContact<float> *con,
                                   int *nk_out, adj::TapeJoint<float> *jt, hipStream_t st);
hipError_t launch_adjoint_step(const DevBatch &B, const float4 *const st[4], const adj::TapeContact<float> *con,
                               const long long *con_off, const int *nk, const float4 *const gin[4],
                               float4 *const gout[4], float *gmu, float *gbounce, int *stats, float *scr,
                               const long long *scr_off, float act_tol, float lin_tol, int lin_maxit, int wg,
                               const adj::TapeJoint<float> *jt, hipStream_t st_);
hipError_t collide_one_noalloc(void *addr, const GeomView &a, const GeomView &b, int maxc, CPoint *host_pts,
                               int *host_n, hipStream_t st);
}  // namespace odemi

using namespace odemi;

enum { BODY_GYRO = 256 };

struct dxJointGroup;
struct dxSpace;

struct dxWorld {
  float gravity[3] = {0, 0, 0};
  float erp = 0.2f, cfm = 1e-5f;           // ode.cpp:1550-1556 (dSINGLE)
  int iterations = 20;
  float sor_w = 1.3f;
  float max_vel = INFINITY, min_depth = 0;
  std::vector<dxBody *> bodies;            // creation order
  std::vector<dxJoint *> joints;           // creation order
  int attach_seq = 0;
  // device mirror
  BatchImpl *mir = nullptr;
  dxSpace *mir_space = nullptr;
  std::vector<dxGeom *> mgeoms;            // mirror geom index -> geom
  std::vector<dxJoint *> mjoints;          // mirror joint index -> joint
  bool struct_dirty = true, state_dirty = true, joints_dirty = true, rank_dirty = true;
  bool device_newer = false;
  bool tape = false;
  DevBuf cin, cin_start, cin_count, feedback, pairs_out, pcount_out, contacts_out, one_out, one_n;
  // the last dWorldQuickStep recorded for dWorldQuickStepBackward: pre-step
  // state (pos, q, lvel, avel; nb float4 each), contacts with impulses, and
  // the contact policy the adjoint differentiates
  DevBuf tape_st, tape_con, tape_nk, tape_jnt;
  DevBuf lcp_ws;                           // dWorldStep's exact-LCP workspace (grown on demand)
  bool tape_ok = false;
  char tape_why[200] = "no step recorded (dWorldSetTape(w, 1) before dWorldQuickStep)";
  int tape_mode = 0, tape_maxc = 4;
  float tape_mu = 0, tape_bounce = 0, tape_bounce_vel = 0;
  int pcap = 0, kcap = 0;
};

struct dxBody {
  dxWorld *world;
  float pos[4] = {0, 0, 0, 0}, q[4] = {1, 0, 0, 0}, R[12], lvel[4] = {0, 0, 0, 0}, avel[4] = {0, 0, 0, 0};
  float facc[4] = {0, 0, 0, 0}, tacc[4] = {0, 0, 0, 0};
  dMass mass;
  float invI[12];
  float invMass = 1;
  void *data = nullptr;
  std::vector<dxGeom *> geoms;             // newest first
  std::vector<dxJoint *> jlist;            // newest attach first
  int mindex = -1;
};

struct dxGeom {
  int cls;
  dxBody *body = nullptr;
  dxSpace *space = nullptr;                // the space holding this geom
  float param[4] = {0, 0, 0, 0};
  float pos[4] = {0, 0, 0, 0}, R[12];
  unsigned long cat = ~0UL, col = ~0UL;
  void *data = nullptr;
  bool dirty = true;
  int mindex = -1;
  int ray_flags = 0;                       // ray FirstContact / BackfaceCull / ClosestHit bits
  virtual ~dxGeom() {}
};

// A space is a geom (collision_space.cpp:85-104): it can sit in another
// space, and its bounding box is the union of its geoms' boxes.
struct dxSpace : dxGeom {
  int minlevel = -3, maxlevel = 10;
  std::vector<dxGeom *> list;              // space list, front first
  int cached_maxc = 4;
  dxWorld *world = nullptr;
  int sublevel = 0;
  int cleanup = 1;
  dxSpace() { cls = dHashSpaceClass; }
};

inline bool is_space(const dxGeom *g) { return g && g->cls >= dFirstSpaceClass && g->cls <= dLastSpaceClass; }

// device geom type of a geom class (-1: no collider on the device)
inline int device_type(int cls) {
  switch (cls) {
    case dSphereClass: return GEOM_SPHERE;
    case dBoxClass: return GEOM_BOX;
    case dPlaneClass: return GEOM_PLANE;
    case dCapsuleClass: return GEOM_CAPSULE;
    case dRayClass: return GEOM_RAY;
    case dCylinderClass: return GEOM_CYLINDER;
    default: return -1;
  }
}

struct dxJoint {
  int type;                                // dJointType*
  dxWorld *world;
  dxJointGroup *group = nullptr;
  dxBody *node[2] = {nullptr, nullptr};
  int flags = 0;
  JointDev dev;                            // ball / hinge / hinge2 record
  dContact contact;                        // contact joints
  dJointFeedback *fb = nullptr;
  int seq = 0;
  void *data = nullptr;
};

struct dxJointGroup {
  std::vector<dxJoint *> joints;
};

namespace {

struct CollideCtx {
  bool active = false;
  dxSpace *space = nullptr;
  dxGeom *g1 = nullptr, *g2 = nullptr;
  const ContactRec *con = nullptr;
  int n = 0, maxc = 0;
} g_ctx;

int g_init = 0;

// every live space, so that a destroyed world can be unlinked from the spaces
// that cache it
std::vector<dxSpace *> g_spaces;

void body_rot_from_q(dxBody *b) { h_quat_to_rot(b->q, b->R); }

void mark_state(dxWorld *w) { if (w) w->state_dirty = true; }
void mark_struct(dxWorld *w) { if (w) w->struct_dirty = true; }

// geoms move to the front of their space list when first marked dirty
// (collision_space.cpp:47-75, 209-213)
// (collision_space.cpp:47-75): from the geom up the space hierarchy, each
// clean geom or space becomes dirty and moves to the front of its space
void geom_moved(dxGeom *g) {
  while (g && g->space && !g->dirty) {
    auto &L = g->space->list;
    auto it = std::find(L.begin(), L.end(), g);
    if (it != L.end()) L.erase(it);
    L.insert(L.begin(), g);
    g->dirty = true;
    if (g->space->world) g->space->world->rank_dirty = true;
    g = g->space;
  }
  for (; g; g = g->space) g->dirty = true;
}
void body_moved(dxBody *b) {
  for (dxGeom *g : b->geoms) geom_moved(g);
  mark_state(b->world);
}

int pull_state(dxWorld *w);

// dGeomSetBody(g, 0) (collision_kernel.cpp:530-547): the geom keeps the
// body's last pose as its own and leaves the body's geom list; it is not
// marked as moved, since its effective pose does not change
void geom_detach_body(dxGeom *g) {
  dxBody *b = g->body;
  if (!b) return;
  pull_state(b->world);
  std::memcpy(g->pos, b->pos, sizeof(g->pos));
  std::memcpy(g->R, b->R, sizeof(g->R));
  auto &L = b->geoms;
  auto it = std::find(L.begin(), L.end(), g);
  if (it != L.end()) L.erase(it);
  mark_struct(b->world);
  g->body = nullptr;
}

// device state -> host copies
int pull_state(dxWorld *w) {
  if (!w->device_newer || !w->mir) return 1;
  BatchImpl *m = w->mir;
  std::vector<float> st(13 * (size_t)m->nb);
  if (!batch_get_state(m, st.data())) return 0;
  for (dxBody *b : w->bodies) {
    if (b->mindex < 0) continue;
    const float *r = &st[13 * (size_t)b->mindex];
    for (int k = 0; k < 3; k++) b->pos[k] = r[k];
    for (int k = 0; k < 4; k++) b->q[k] = r[3 + k];
    for (int k = 0; k < 3; k++) b->lvel[k] = r[7 + k];
    for (int k = 0; k < 3; k++) b->avel[k] = r[10 + k];
    body_rot_from_q(b);
  }
  w->device_newer = false;
  return 1;
}

JointDev joint_record(const dxJoint *j) {
  JointDev d = j->dev;
  d.b1 = j->node[0] ? j->node[0]->mindex : -1;
  d.b2 = j->node[1] ? j->node[1]->mindex : -1;
  d.seq = j->seq;
  d.flags = j->flags;
  return d;
}

bool is_permanent(const dxJoint *j) { return j->type != dJointTypeContact; }

// (re)creates the device mirror of a world with its space
int build_mirror(dxWorld *w, dxSpace *sp) {
  if (w->tape_ok) {
    w->tape_ok = false;
    std::snprintf(w->tape_why, sizeof(w->tape_why), "the world changed since the recorded step");
  }
  if (w->mir) {
    pull_state(w);
    batch_destroy(w->mir);
    w->mir = nullptr;
  }
  BatchArrays A;
  WorldInfo wi{};
  int nb = (int)w->bodies.size();
  wi.b0 = 0; wi.nb = nb; wi.g0 = 0; wi.j0 = 0;
  w->mgeoms.clear();
  if (sp) {
    // mirror index: creation order within the space (reverse of the list
    // order when nothing has moved); ranks carry the actual list order
    std::vector<dxGeom *> gl = sp->list;
    w->mgeoms = gl;
  }
  for (int i = 0; i < nb; i++) w->bodies[i]->mindex = i;
  w->mjoints.clear();
  for (dxJoint *j : w->joints)
    if (is_permanent(j) && j->node[0]) w->mjoints.push_back(j);
  wi.ng = (int)w->mgeoms.size();
  wi.nj = (int)w->mjoints.size();
  world_capacities(nb, wi.ng, wi.nj, 1.0, wi.pcap, wi.kcap, wi.rcap);
  wi.kcap = std::max(wi.kcap, 4096);
  wi.rcap = 3 * wi.kcap + 6 * wi.nj + 8;
  A.wi.push_back(wi);
  WorldParams P;
  std::memset(&P, 0, sizeof(P));
  for (int k = 0; k < 3; k++) P.gravity[k] = w->gravity[k];
  P.erp = w->erp; P.cfm = w->cfm; P.iterations = w->iterations; P.sor_w = w->sor_w;
  P.max_vel = w->max_vel; P.min_depth = w->min_depth; P.dt = 0.01f;
  P.hash_min = sp ? sp->minlevel : -3;
  P.hash_max = sp ? sp->maxlevel : 10;
  P.max_contacts = sp ? sp->cached_maxc : 4;
  P.cell_scale = 1.0f;
  A.wp.push_back(P);
  for (dxBody *b : w->bodies) {
    A.pos.push_back(make_float4(b->pos[0], b->pos[1], b->pos[2], 0));
    A.q.push_back(make_float4(b->q[0], b->q[1], b->q[2], b->q[3]));
    A.lvel.push_back(make_float4(b->lvel[0], b->lvel[1], b->lvel[2], 0));
    A.avel.push_back(make_float4(b->avel[0], b->avel[1], b->avel[2], 0));
    A.facc.push_back(make_float4(b->facc[0], b->facc[1], b->facc[2], 0));
    A.tacc.push_back(make_float4(b->tacc[0], b->tacc[1], b->tacc[2], 0));
    for (int r = 0; r < 3; r++) {
      A.I.push_back(make_float4(b->mass.I[4 * r], b->mass.I[4 * r + 1], b->mass.I[4 * r + 2], 0));
      A.invI.push_back(make_float4(b->invI[4 * r], b->invI[4 * r + 1], b->invI[4 * r + 2], 0));
    }
    A.mass.push_back(b->mass.mass);
    A.invMass.push_back(b->invMass);
  }
  for (size_t i = 0; i < w->mgeoms.size(); i++) {
    dxGeom *g = w->mgeoms[i];
    g->mindex = (int)i;
    A.gtype.push_back(device_type(g->cls));
    A.gbody.push_back(g->body ? g->body->mindex : -1);
    A.gparam.push_back(make_float4(g->param[0], g->param[1], g->param[2], g->param[3]));
    A.gpos.push_back(make_float4(g->pos[0], g->pos[1], g->pos[2], 0));
    for (int r = 0; r < 3; r++) A.gR.push_back(make_float4(g->R[4 * r], g->R[4 * r + 1], g->R[4 * r + 2], 0));
    A.gcat.push_back((unsigned)g->cat);
    A.gcol.push_back((unsigned)g->col);
  }
  A.grank.assign(w->mgeoms.size(), 0);
  if (sp)
    for (size_t r = 0; r < sp->list.size(); r++) A.grank[sp->list[r]->mindex] = (int)r;
  for (dxJoint *j : w->mjoints) A.joints.push_back(joint_record(j));
  w->mir = batch_create_arrays(A, ODEMI_ORDER_REFERENCE);
  if (!w->mir) return 0;
  w->mir_space = sp;
  w->pcap = wi.pcap;
  w->kcap = wi.kcap;
  if (!w->pairs_out.alloc(sizeof(PairRec) * wi.pcap) || !w->pcount_out.alloc(sizeof(int) * (wi.pcap + 1)) ||
      !w->contacts_out.alloc(sizeof(ContactRec) * wi.kcap) || !w->cin_start.alloc(sizeof(int)) ||
      !w->cin_count.alloc(sizeof(int)) || !w->one_out.alloc(sizeof(CPoint) * MAX_SLOTS) || !w->one_n.alloc(16)) {
    report_error("device allocation failed");
    return 0;
  }
  w->struct_dirty = false;
  w->state_dirty = false;
  w->joints_dirty = false;
  w->rank_dirty = false;
  w->device_newer = false;
  return 1;
}

// brings the device mirror up to date with the host objects
int sync_mirror(dxWorld *w, dxSpace *sp) {
  if (!sp) sp = w->mir_space;
  if (w->struct_dirty || !w->mir || (sp && sp != w->mir_space)) return build_mirror(w, sp);
  BatchImpl *m = w->mir;
  if (w->state_dirty) {
    pull_state(w);
    std::vector<float> st(13 * (size_t)m->nb);
    std::vector<float4> fa(m->nb), ta(m->nb);
    for (dxBody *b : w->bodies) {
      float *r = &st[13 * (size_t)b->mindex];
      for (int k = 0; k < 3; k++) r[k] = b->pos[k];
      for (int k = 0; k < 4; k++) r[3 + k] = b->q[k];
      for (int k = 0; k < 3; k++) r[7 + k] = b->lvel[k];
      for (int k = 0; k < 3; k++) r[10 + k] = b->avel[k];
      fa[b->mindex] = make_float4(b->facc[0], b->facc[1], b->facc[2], 0);
      ta[b->mindex] = make_float4(b->tacc[0], b->tacc[1], b->tacc[2], 0);
    }
    if (!batch_set_state(m, st.data())) return 0;
    if (m->nb) {
      OD_HIP(hipMemcpy(m->d.facc.p, fa.data(), sizeof(float4) * m->nb, hipMemcpyHostToDevice));
      OD_HIP(hipMemcpy(m->d.tacc.p, ta.data(), sizeof(float4) * m->nb, hipMemcpyHostToDevice));
    }
    w->state_dirty = false;
  }
  if (w->joints_dirty) {
    std::vector<JointDev> J;
    for (dxJoint *j : w->mjoints) J.push_back(joint_record(j));
    if (!J.empty()) OD_HIP(hipMemcpy(m->d.joints.p, J.data(), sizeof(JointDev) * J.size(), hipMemcpyHostToDevice));
    w->joints_dirty = false;
  }
  if (w->rank_dirty && sp) {
    std::vector<int> rk(w->mgeoms.size());
    for (size_t r = 0; r < sp->list.size(); r++) rk[sp->list[r]->mindex] = (int)r;
    if (!rk.empty()) OD_HIP(hipMemcpy(m->d.grank.p, rk.data(), sizeof(int) * rk.size(), hipMemcpyHostToDevice));
    w->rank_dirty = false;
  }
  return 1;
}

int world_params_upload(dxWorld *w, float dt, int maxc, int exact) {
  BatchImpl *m = w->mir;
  WorldParams P;
  OD_HIP(hipMemcpy(&P, m->d.wp.p, sizeof(P), hipMemcpyDeviceToHost));
  for (int k = 0; k < 3; k++) P.gravity[k] = w->gravity[k];
  P.erp = w->erp; P.cfm = w->cfm; P.iterations = w->iterations; P.sor_w = w->sor_w;
  P.max_vel = w->max_vel; P.min_depth = w->min_depth;
  if (dt > 0) P.dt = dt;
  if (maxc > 0) P.max_contacts = maxc;
  P.seed = (unsigned)global_seed();
  if (w->mir_space) { P.hash_min = w->mir_space->minlevel; P.hash_max = w->mir_space->maxlevel; }
  P.exact = exact;
  OD_HIP(hipMemcpy(m->d.wp.p, &P, sizeof(P), hipMemcpyHostToDevice));
  return 1;
}

dxWorld *space_world(dxSpace *sp) {
  if (sp->world) return sp->world;
  for (dxGeom *g : sp->list) {
    if (g->body) return sp->world = g->body->world;
    if (is_space(g)) {
      dxWorld *w = space_world(static_cast<dxSpace *>(g));
      if (w) return sp->world = w;
    }
  }
  return nullptr;
}

void detach(dxJoint *j) {
  for (int i = 0; i < 2; i++) {
    dxBody *b = j->node[i];
    if (!b) continue;
    auto it = std::find(b->jlist.begin(), b->jlist.end(), j);
    if (it != b->jlist.end()) b->jlist.erase(it);
  }
  j->node[0] = j->node[1] = nullptr;
}

void destroy_joint(dxJoint *j) {
  detach(j);
  dxWorld *w = j->world;
  if (w) {
    auto it = std::find(w->joints.begin(), w->joints.end(), j);
    if (it != w->joints.end()) w->joints.erase(it);
    if (is_permanent(j)) mark_struct(w);
  }
  delete j;
}

HPose pose_of(dxBody *b) {
  HPose p;
  std::memcpy(p.pos, b->pos, sizeof(p.pos));
  std::memcpy(p.q, b->q, sizeof(p.q));
  std::memcpy(p.R, b->R, sizeof(p.R));
  return p;
}

void joint_changed(dxJoint *j) {
  if (j->world) j->world->joints_dirty = true;
}

void write_contact(dContactGeom *c, const float *pos, const float *n, float depth, dGeomID g1, dGeomID g2) {
  for (int k = 0; k < 3; k++) { c->pos[k] = pos[k]; c->normal[k] = n[k]; }
  c->pos[3] = 0;
  c->normal[3] = 0;
  c->depth = depth;
  c->g1 = g1;
  c->g2 = g2;
  c->side1 = -1;
  c->side2 = -1;
}

// Checks that the coming step is one dWorldQuickStepBackward differentiates
// (the conditions of adj_step.hpp: contact joints, ball joints, hinges
// without limits and hinge2 joints; one contact policy -- dContactApprox1
// friction or none, optional dContactBounce -- for every contact, no forces
// but gravity, every contact's geoms in the world's space) and sizes the
// tape; records the reason when it is not.  Joint rows the adjoint does not
// cover are found by the capture after the step.
bool tape_prepare(dxWorld *w, const std::vector<ContactIn> &cin) {
  auto refuse = [&](const char *fmt, auto... a) {
    std::snprintf(w->tape_why, sizeof(w->tape_why), fmt, a...);
    w->tape_ok = false;
    return false;
  };
  for (dxGeom *g : w->mgeoms)
    if (g->cls != dSphereClass && g->cls != dBoxClass && g->cls != dPlaneClass)
      return refuse("the world has a capsule, cylinder or ray geom (the adjoint colliders are sphere, box and plane)");
  for (dxBody *b : w->bodies)
    for (int k = 0; k < 3; k++)
      if (b->facc[k] != 0 || b->tacc[k] != 0) return refuse("a body carries a force or torque besides gravity");
  const int approx = dContactApprox1;
  for (size_t i = 0; i < cin.size(); i++) {
    const ContactIn &c = cin[i];
    if (c.g1 < 0 || c.g2 < 0) return refuse("contact %zu names a geom outside the world's space", i);
    if (c.mode & ~(approx | dContactBounce)) return refuse("contact %zu uses surface mode bits 0x%x", i, c.mode);
    if (c.mu > 0 && (c.mode & approx) != approx) return refuse("contact %zu: friction without dContactApprox1", i);
    if (i == 0) {
      w->tape_mode = c.mode; w->tape_mu = c.mu; w->tape_bounce = c.bounce; w->tape_bounce_vel = c.bounce_vel;
    } else if (c.mode != w->tape_mode || c.mu != w->tape_mu || c.bounce != w->tape_bounce ||
               c.bounce_vel != w->tape_bounce_vel) {
      return refuse("contacts %d and %zu differ in surface parameters", 0, i);
    }
  }
  w->tape_maxc = w->mir_space ? w->mir_space->cached_maxc : 4;
  const size_t nb = w->mir->nb, kcap = w->mir->wi[0].kcap;
  const size_t nj = w->mir->nj;
  if (!w->tape_st.alloc(4 * sizeof(float4) * (nb + 1)) ||
      !w->tape_con.alloc(sizeof(adj::TapeContact<float>) * (kcap + 1)) || !w->tape_nk.alloc(16) ||
      !w->tape_jnt.alloc(sizeof(adj::TapeJoint<float>) * (nj + 1)))
    return refuse("device allocation of the tape failed");
  return true;
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- init
// library set-up: the device address of dCollideNoAlloc's scratch array on
// the current GPU, when there is one (looked up again on a device's first
// dCollideNoAlloc call otherwise)
static void init_device_state() {
  int n = 0;
  if (hipGetDeviceCount(&n) == hipSuccess && n > 0) (void)noalloc_scratch(nullptr);
}
int dInitODE2(unsigned int) { g_init++; init_device_state(); return 1; }
void dInitODE(void) { g_init++; init_device_state(); }
void dCloseODE(void) { if (g_init > 0) g_init--; }

// ---------------------------------------------------------------- world
dWorldID dWorldCreate(void) { return new dxWorld(); }
// dWorldDestroy (ode.cpp:1585-1612): every body is destroyed as dBodyDestroy
// does (its geoms keep the body's last pose), grouped joints are only
// deactivated (their group still owns them), and spaces that cached this
// world forget it.
void dWorldDestroy(dWorldID w) {
  if (!w) return;
  pull_state(w);
  for (dxBody *b : w->bodies) {
    std::vector<dxGeom *> gs = b->geoms;
    for (dxGeom *g : gs) geom_detach_body(g);
  }
  if (w->mir) batch_destroy(w->mir);
  w->mir = nullptr;
  for (dxJoint *j : w->joints) {
    j->world = nullptr;
    detach(j);
    if (!j->group) delete j;
  }
  for (dxBody *b : w->bodies) delete b;
  for (dxSpace *s : g_spaces)
    if (s->world == w) s->world = nullptr;
  delete w;
}
void dWorldSetGravity(dWorldID w, dReal x, dReal y, dReal z) { w->gravity[0] = x; w->gravity[1] = y; w->gravity[2] = z; }
void dWorldGetGravity(dWorldID w, dVector3 g) { for (int k = 0; k < 3; k++) g[k] = w->gravity[k]; }
void dWorldSetERP(dWorldID w, dReal erp) { w->erp = erp; }
dReal dWorldGetERP(dWorldID w) { return w->erp; }
void dWorldSetCFM(dWorldID w, dReal cfm) { w->cfm = cfm; }
dReal dWorldGetCFM(dWorldID w) { return w->cfm; }
void dWorldSetQuickStepNumIterations(dWorldID w, int n) { w->iterations = n; }
int dWorldGetQuickStepNumIterations(dWorldID w) { return w->iterations; }
void dWorldSetQuickStepW(dWorldID w, dReal v) { w->sor_w = v; }
dReal dWorldGetQuickStepW(dWorldID w) { return w->sor_w; }
void dWorldSetContactMaxCorrectingVel(dWorldID w, dReal v) { w->max_vel = v; }
void dWorldSetContactSurfaceLayer(dWorldID w, dReal d) { w->min_depth = d; }

// dWorldQuickStep (ode.cpp:1807-1825): uploads the contact joints and runs
// the step kernel on the world's mirror.
// One step of the world on its device mirror: dWorldQuickStep (exact =
// false: the reference-order SOR) or dWorldStep (exact: each island's LCP
// solved exactly).
static int world_step_impl(dWorldID w, dReal stepsize, bool exact, const char *who) {
  if (!w || !(stepsize > 0)) return 0;
  if (!sync_mirror(w, nullptr)) return 0;
  BatchImpl *m = w->mir;
  // contact joints in creation order
  std::vector<ContactIn> cin;
  std::vector<dJointFeedback *> fbs;
  for (dxJoint *j : w->joints) {
    if (j->type != dJointTypeContact || !j->node[0]) continue;
    ContactIn c;
    std::memset(&c, 0, sizeof(c));
    const dContact &k = j->contact;
    for (int i = 0; i < 3; i++) { c.pos[i] = k.geom.pos[i]; c.normal[i] = k.geom.normal[i]; c.fdir1[i] = k.fdir1[i]; }
    c.depth = k.geom.depth;
    c.b1 = j->node[0]->mindex;
    c.b2 = j->node[1] ? j->node[1]->mindex : -1;
    c.flags = j->flags;
    c.seq = j->seq;
    c.mode = k.surface.mode;
    c.mu = k.surface.mu; c.mu2 = k.surface.mu2; c.bounce = k.surface.bounce; c.bounce_vel = k.surface.bounce_vel;
    c.soft_erp = k.surface.soft_erp; c.soft_cfm = k.surface.soft_cfm; c.motion1 = k.surface.motion1;
    c.motion2 = k.surface.motion2; c.motionN = k.surface.motionN; c.slip1 = k.surface.slip1;
    c.slip2 = k.surface.slip2;
    c.fb = -1;
    if (j->fb) { c.fb = (int)fbs.size(); fbs.push_back(j->fb); }
    c.g1 = (k.geom.g1 && k.geom.g1->mindex >= 0 && k.geom.g1->mindex < (int)w->mgeoms.size() &&
            w->mgeoms[k.geom.g1->mindex] == k.geom.g1) ? k.geom.g1->mindex : -1;
    c.g2 = (k.geom.g2 && k.geom.g2->mindex >= 0 && k.geom.g2->mindex < (int)w->mgeoms.size() &&
            w->mgeoms[k.geom.g2->mindex] == k.geom.g2) ? k.geom.g2->mindex : -1;
    cin.push_back(c);
  }
  if ((int)cin.size() > m->wi[0].kcap) {
    report_error("%s: %zu contact joints exceed the capacity %d", who, cin.size(), m->wi[0].kcap);
    return 0;
  }
  // feedback slots of permanent joints
  std::vector<JointDev> J;
  bool jfb = false;
  for (dxJoint *j : w->mjoints) {
    JointDev d = joint_record(j);
    d.fb = -1;
    if (j->fb) { d.fb = (int)fbs.size(); fbs.push_back(j->fb); jfb = true; }
    J.push_back(d);
  }
  if (jfb || w->joints_dirty) {
    if (!J.empty()) OD_HIP(hipMemcpy(m->d.joints.p, J.data(), sizeof(JointDev) * J.size(), hipMemcpyHostToDevice));
    w->joints_dirty = false;
  }
  if (!w->cin.alloc(sizeof(ContactIn) * (cin.size() + 1))) return 0;
  if (!cin.empty()) OD_HIP(hipMemcpy(w->cin.p, cin.data(), sizeof(ContactIn) * cin.size(), hipMemcpyHostToDevice));
  int zero = 0, nk = (int)cin.size();
  OD_HIP(hipMemcpy(w->cin_start.p, &zero, sizeof(int), hipMemcpyHostToDevice));
  OD_HIP(hipMemcpy(w->cin_count.p, &nk, sizeof(int), hipMemcpyHostToDevice));
  if (!w->feedback.alloc(sizeof(float) * 12 * (fbs.size() + 1))) return 0;
  if (!fbs.empty()) {
    // the user's current values: the kernel overwrites what the reference
    // writes (quickstep.cpp:918-957: f1/t1 of every joint with rows, f2/t2
    // only when the joint has a second body) and the rest comes back as it was
    std::vector<float> f0(12 * fbs.size());
    for (size_t i = 0; i < fbs.size(); i++)
      for (int k = 0; k < 3; k++) {
        f0[12 * i + k] = fbs[i]->f1[k]; f0[12 * i + 3 + k] = fbs[i]->t1[k];
        f0[12 * i + 6 + k] = fbs[i]->f2[k]; f0[12 * i + 9 + k] = fbs[i]->t2[k];
      }
    OD_HIP(hipMemcpy(w->feedback.p, f0.data(), sizeof(float) * f0.size(), hipMemcpyHostToDevice));
  }
  if (!world_params_upload(w, stepsize, 0, exact ? 1 : 0)) return 0;
  refresh_dev(m);
  DevBatch D = m->dev;
  D.cin = w->cin.as<ContactIn>();
  D.cin_start = w->cin_start.as<int>();
  D.cin_count = w->cin_count.as<int>();
  D.feedback = fbs.empty() ? nullptr : w->feedback.as<float>();
  D.ctrl = nullptr;
  if (exact) {
    // workspace for an island of every row of the world (a contact has at
    // most 3 rows, a joint at most 6)
    const int rows_ub = 3 * (int)cin.size() + 6 * (int)w->mjoints.size() + 1;
    if (!w->lcp_ws.alloc(sizeof(float) * (size_t)lcp_ws_floats(rows_ub))) {
      report_error("%s: device allocation of the LCP workspace (%d rows) failed", who, rows_ub);
      return 0;
    }
    D.lcp_ws = w->lcp_ws.as<float>();
    D.lcp_ws_stride = lcp_ws_floats(rows_ub);
    D.lcp_ws_rows = rows_ub;
  }
  const bool record = !exact && w->tape && tape_prepare(w, cin);
  if (exact && w->tape) {
    std::snprintf(w->tape_why, sizeof(w->tape_why), "the last step was a dWorldStep (the adjoint is of dWorldQuickStep)");
    w->tape_ok = false;
  }
  if (record) {
    const size_t bytes = sizeof(float4) * (size_t)m->nb;
    float4 *st = w->tape_st.as<float4>();
    OD_HIP(hipMemcpyAsync(st, m->d.pos.p, bytes, hipMemcpyDeviceToDevice, m->stream));
    OD_HIP(hipMemcpyAsync(st + m->nb, m->d.q.p, bytes, hipMemcpyDeviceToDevice, m->stream));
    OD_HIP(hipMemcpyAsync(st + 2 * (size_t)m->nb, m->d.lvel.p, bytes, hipMemcpyDeviceToDevice, m->stream));
    OD_HIP(hipMemcpyAsync(st + 3 * (size_t)m->nb, m->d.avel.p, bytes, hipMemcpyDeviceToDevice, m->stream));
  }
  OD_HIP(launch_world_kernel(KM_API_STEP, D, 0, 1, 0, m->wg, 0, m->stream));
  if (record)
    OD_HIP(launch_api_tape_capture(D, w->cin.as<ContactIn>(), nk, w->tape_con.as<adj::TapeContact<float>>(),
                                   w->tape_nk.as<int>(), m->nj ? w->tape_jnt.as<adj::TapeJoint<float>>() : nullptr,
                                   m->stream));
  OD_HIP(hipStreamSynchronize(m->stream));
  long long cnt[5];
  batch_counters(m, cnt);
  if (cnt[4]) {
    report_error("%s: %s (flags %lld)", who,
                 (cnt[4] & OVF_LCP) ? "the LCP solve did not finish" : "device capacity exceeded", cnt[4]);
    return 0;
  }
  if (!exact) w->tape_ok = record;
  if (record && m->nj) {
    int cov[2];
    OD_HIP(hipMemcpy(cov, w->tape_nk.p, sizeof(cov), hipMemcpyDeviceToHost));
    if (cov[1]) {
      std::snprintf(w->tape_why, sizeof(w->tape_why),
                    "a joint had rows the adjoint does not cover (a hinge limit, a bouncy or motor-loaded limit)");
      w->tape_ok = false;
    }
  }
  // RNG state, feedback, joint limit state, space order
  WorldParams P;
  OD_HIP(hipMemcpy(&P, m->d.wp.p, sizeof(P), hipMemcpyDeviceToHost));
  global_seed() = P.seed;
  if (!fbs.empty()) {
    std::vector<float> f(12 * fbs.size());
    OD_HIP(hipMemcpy(f.data(), w->feedback.p, sizeof(float) * f.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < fbs.size(); i++) {
      dJointFeedback *o = fbs[i];
      for (int k = 0; k < 3; k++) {
        o->f1[k] = f[12 * i + k]; o->t1[k] = f[12 * i + 3 + k];
        o->f2[k] = f[12 * i + 6 + k]; o->t2[k] = f[12 * i + 9 + k];
      }
    }
  }
  if (!w->mjoints.empty()) {
    std::vector<JointDev> Jd(w->mjoints.size());
    OD_HIP(hipMemcpy(Jd.data(), m->d.joints.p, sizeof(JointDev) * Jd.size(), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < Jd.size(); i++) {
      w->mjoints[i]->dev.lm1 = Jd[i].lm1;
      w->mjoints[i]->dev.lm2 = Jd[i].lm2;
    }
  }
  if (w->mir_space && !w->mgeoms.empty()) {
    std::vector<int> rk(w->mgeoms.size());
    OD_HIP(hipMemcpy(rk.data(), m->d.grank.p, sizeof(int) * rk.size(), hipMemcpyDeviceToHost));
    std::vector<dxGeom *> L(rk.size());
    for (size_t g = 0; g < rk.size(); g++) L[rk[g]] = w->mgeoms[g];
    w->mir_space->list = L;
    for (dxGeom *g : w->mgeoms)
      if (g->body) g->dirty = true;
  }
  for (dxBody *b : w->bodies) {
    for (int k = 0; k < 3; k++) { b->facc[k] = 0; b->tacc[k] = 0; }
  }
  w->device_newer = true;
  return 1;
}
int dWorldQuickStep(dWorldID w, dReal stepsize) { return world_step_impl(w, stepsize, false, "dWorldQuickStep"); }
int dWorldStep(dWorldID w, dReal stepsize) { return world_step_impl(w, stepsize, true, "dWorldStep"); }

// ---------------------------------------------------------------- bodies
dBodyID dBodyCreate(dWorldID w) {
  dxBody *b = new dxBody();
  b->world = w;
  dMassSetParameters(&b->mass, 1, 0, 0, 0, 1, 1, 1, 0, 0, 0);
  std::memset(b->invI, 0, sizeof(b->invI));
  b->invI[0] = b->invI[5] = b->invI[10] = 1;
  body_rot_from_q(b);
  w->bodies.push_back(b);
  mark_struct(w);
  return b;
}
void dBodyDestroy(dBodyID b) {
  dxWorld *w = b->world;
  pull_state(w);
  // dGeomSetBody(geom, 0) for every geom of the body (ode.cpp:296-309)
  std::vector<dxGeom *> gs = b->geoms;
  for (dxGeom *g : gs) geom_detach_body(g);
  std::vector<dxJoint *> js = b->jlist;
  for (dxJoint *j : js) detach(j);
  w->bodies.erase(std::find(w->bodies.begin(), w->bodies.end(), b));
  mark_struct(w);
  delete b;
}
void dBodySetData(dBodyID b, void *d) { b->data = d; }
void *dBodyGetData(dBodyID b) { return b->data; }
dWorldID dBodyGetWorld(dBodyID b) { return b->world; }
void dBodySetPosition(dBodyID b, dReal x, dReal y, dReal z) {
  pull_state(b->world);
  b->pos[0] = x; b->pos[1] = y; b->pos[2] = z;
  body_moved(b);
}
void dBodySetRotation(dBodyID b, const dMatrix3 R) {
  // quaternion of R (rotation.cpp:262-307), normalised; R follows from it
  pull_state(b->world);
  float tr = R[0] + R[5] + R[10], s, q[4];
  if (tr >= 0) {
    s = std::sqrt(tr + 1);
    q[0] = 0.5f * s;
    s = 0.5f / s;
    q[1] = (R[9] - R[6]) * s; q[2] = (R[2] - R[8]) * s; q[3] = (R[4] - R[1]) * s;
  } else if (R[5] > R[0] && R[5] >= R[10]) {
    s = std::sqrt((R[5] - (R[10] + R[0])) + 1);
    q[2] = 0.5f * s;
    s = 0.5f / s;
    q[3] = (R[6] + R[9]) * s; q[1] = (R[1] + R[4]) * s; q[0] = (R[2] - R[8]) * s;
  } else if (R[10] > R[0]) {
    s = std::sqrt((R[10] - (R[0] + R[5])) + 1);
    q[3] = 0.5f * s;
    s = 0.5f / s;
    q[1] = (R[8] + R[2]) * s; q[2] = (R[6] + R[9]) * s; q[0] = (R[4] - R[1]) * s;
  } else {
    s = std::sqrt((R[0] - (R[5] + R[10])) + 1);
    q[1] = 0.5f * s;
    s = 0.5f / s;
    q[2] = (R[1] + R[4]) * s; q[3] = (R[8] + R[2]) * s; q[0] = (R[9] - R[6]) * s;
  }
  h_normalize4(q);
  std::memcpy(b->q, q, sizeof(q));
  body_rot_from_q(b);
  body_moved(b);
}
void dBodySetQuaternion(dBodyID b, const dQuaternion q) {
  pull_state(b->world);
  for (int k = 0; k < 4; k++) b->q[k] = q[k];
  h_normalize4(b->q);
  body_rot_from_q(b);
  body_moved(b);
}
void dBodySetLinearVel(dBodyID b, dReal x, dReal y, dReal z) {
  pull_state(b->world);
  b->lvel[0] = x; b->lvel[1] = y; b->lvel[2] = z;
  mark_state(b->world);
}
void dBodySetAngularVel(dBodyID b, dReal x, dReal y, dReal z) {
  pull_state(b->world);
  b->avel[0] = x; b->avel[1] = y; b->avel[2] = z;
  mark_state(b->world);
}
const dReal *dBodyGetPosition(dBodyID b) { pull_state(b->world); return b->pos; }
const dReal *dBodyGetRotation(dBodyID b) { pull_state(b->world); return b->R; }
const dReal *dBodyGetQuaternion(dBodyID b) { pull_state(b->world); return b->q; }
const dReal *dBodyGetLinearVel(dBodyID b) { pull_state(b->world); return b->lvel; }
const dReal *dBodyGetAngularVel(dBodyID b) { pull_state(b->world); return b->avel; }
void dBodySetMass(dBodyID b, const dMass *m) {
  // dBodySetMass (ode.cpp:492-509): inverse inertia by Cholesky
  b->mass = *m;
  if (!h_invert_pd3(b->mass.I, b->invI)) {
    report_debug("inertia must be positive definite!");
    std::memset(b->invI, 0, sizeof(b->invI));
    b->invI[0] = b->invI[5] = b->invI[10] = 1;
  }
  b->invMass = 1.0f / m->mass;
  mark_struct(b->world);
}
void dBodyGetMass(dBodyID b, dMass *m) { *m = b->mass; }
void dBodyAddForce(dBodyID b, dReal x, dReal y, dReal z) {
  b->facc[0] += x; b->facc[1] += y; b->facc[2] += z;
  mark_state(b->world);
}
void dBodyAddTorque(dBodyID b, dReal x, dReal y, dReal z) {
  b->tacc[0] += x; b->tacc[1] += y; b->tacc[2] += z;
  mark_state(b->world);
}
const dReal *dBodyGetForce(dBodyID b) { return b->facc; }
const dReal *dBodyGetTorque(dBodyID b) { return b->tacc; }

// ---------------------------------------------------------------- mass
void dMassSetZero(dMass *m) { std::memset(m, 0, sizeof(*m)); }
void dMassSetParameters(dMass *m, dReal mass, dReal cx, dReal cy, dReal cz, dReal I11, dReal I22, dReal I33,
                        dReal I12, dReal I13, dReal I23) {  // mass.cpp:94-113
  dMassSetZero(m);
  m->mass = mass;
  m->c[0] = cx; m->c[1] = cy; m->c[2] = cz;
  m->I[0] = I11; m->I[5] = I22; m->I[10] = I33;
  m->I[1] = I12; m->I[2] = I13; m->I[6] = I23;
  m->I[4] = I12; m->I[8] = I13; m->I[9] = I23;
}
void dMassSetSphereTotal(dMass *m, dReal total, dReal r) {  // mass.cpp:125-140
  dMassSetZero(m);
  m->mass = total;
  dReal II = 0.4f * total * r * r;
  m->I[0] = m->I[5] = m->I[10] = II;
}
void dMassSetSphere(dMass *m, dReal density, dReal r) {
  dMassSetSphereTotal(m, (dReal)((4.0f / 3.0f) * 3.14159265358979323846 * r * r * r * density), r);
}
void dMassSetBoxTotal(dMass *m, dReal total, dReal lx, dReal ly, dReal lz) {  // mass.cpp:208-221
  dMassSetZero(m);
  m->mass = total;
  m->I[0] = total / 12.0f * (ly * ly + lz * lz);
  m->I[5] = total / 12.0f * (lx * lx + lz * lz);
  m->I[10] = total / 12.0f * (lx * lx + ly * ly);
}
void dMassSetBox(dMass *m, dReal density, dReal lx, dReal ly, dReal lz) {
  dMassSetBoxTotal(m, lx * ly * lz * density, lx, ly, lz);
}
// mass.cpp:141-162: a cylinder of the given length plus two hemispherical
// caps, axis along `direction` (1..3)
void dMassSetCapsule(dMass *m, dReal density, int direction, dReal radius, dReal length) {
  if (direction < 1 || direction > 3) {
    report_debug("bad direction number");
    return;
  }
  dMassSetZero(m);
  const double pi = 3.14159265358979323846;
  dReal M1 = (dReal)(pi * radius * radius * length * density);
  dReal M2 = (dReal)((4.0f / 3.0f) * pi * radius * radius * radius * density);
  m->mass = M1 + M2;
  dReal Ia = M1 * (0.25f * radius * radius + (1.0f / 12.0f) * length * length) +
             M2 * (0.4f * radius * radius + 0.375f * radius * length + 0.25f * length * length);
  dReal Ib = (M1 * 0.5f + M2 * 0.4f) * radius * radius;
  m->I[0] = m->I[5] = m->I[10] = Ia;
  m->I[5 * (direction - 1)] = Ib;
}
void dMassSetCapsuleTotal(dMass *m, dReal total, int direction, dReal radius, dReal length) {
  dMassSetCapsule(m, 1.0f, direction, radius, length);
  dMassAdjust(m, total);
}
// mass.cpp:173-198: a solid cylinder, axis along `direction` (1..3)
void dMassSetCylinderTotal(dMass *m, dReal total, int direction, dReal radius, dReal length) {
  if (direction < 1 || direction > 3) {
    report_debug("bad direction number");
    return;
  }
  dMassSetZero(m);
  dReal r2 = radius * radius;
  m->mass = total;
  dReal I = total * (0.25f * r2 + (1.0f / 12.0f) * length * length);
  m->I[0] = m->I[5] = m->I[10] = I;
  m->I[5 * (direction - 1)] = total * 0.5f * r2;
}
void dMassSetCylinder(dMass *m, dReal density, int direction, dReal radius, dReal length) {
  dMassSetCylinderTotal(m, (dReal)(3.14159265358979323846 * radius * radius * length * density), direction, radius,
                        length);
}
void dMassAdjust(dMass *m, dReal newmass) {  // mass.cpp:441-452
  dReal s = newmass / m->mass;
  m->mass = newmass;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m->I[4 * i + j] *= s;
}

// ---------------------------------------------------------------- joints
dJointGroupID dJointGroupCreate(int) { return new dxJointGroup(); }
void dJointGroupEmpty(dJointGroupID g) {
  // newest first (ode.cpp:1305-1332)
  for (auto it = g->joints.rbegin(); it != g->joints.rend(); ++it) destroy_joint(*it);
  g->joints.clear();
}
void dJointGroupDestroy(dJointGroupID g) {
  dJointGroupEmpty(g);
  delete g;
}
static dxJoint *make_joint(dWorldID w, dJointGroupID g, int type) {
  dxJoint *j = new dxJoint();
  j->type = type;
  j->world = w;
  j->group = g;
  int jt = type == dJointTypeBall ? JT_BALL : (type == dJointTypeHinge ? JT_HINGE : (type == dJointTypeHinge2 ? JT_HINGE2 : JT_CONTACT));
  h_joint_defaults(j->dev, jt, w->erp, w->cfm);
  w->joints.push_back(j);
  if (g) g->joints.push_back(j);
  if (type != dJointTypeContact) mark_struct(w);
  return j;
}
dJointID dJointCreateBall(dWorldID w, dJointGroupID g) { return make_joint(w, g, dJointTypeBall); }
dJointID dJointCreateHinge(dWorldID w, dJointGroupID g) { return make_joint(w, g, dJointTypeHinge); }
dJointID dJointCreateHinge2(dWorldID w, dJointGroupID g) { return make_joint(w, g, dJointTypeHinge2); }
dJointID dJointCreateContact(dWorldID w, dJointGroupID g, const dContact *c) {
  dxJoint *j = make_joint(w, g, dJointTypeContact);
  j->contact = *c;
  return j;
}
void dJointDestroy(dJointID j) {
  if (!j || j->group) return;   // grouped joints are destroyed by their group (ode.cpp:1276-1287)
  destroy_joint(j);
}
// dJointAttach (ode.cpp:1348-1399)
void dJointAttach(dJointID j, dBodyID b1, dBodyID b2) {
  if (b1 && b1 == b2) {   // dUASSERT in the reference (ode.cpp:1352)
    report_debug("can't have body1==body2");
    return;
  }
  if (j->node[0] || j->node[1]) detach(j);
  if (!b1) { b1 = b2; b2 = nullptr; j->flags |= JF_REVERSE; }
  else j->flags &= ~JF_REVERSE;
  j->node[0] = b1;
  j->node[1] = b2;
  if (b1) b1->jlist.insert(b1->jlist.begin(), j);
  if (b2) b2->jlist.insert(b2->jlist.begin(), j);
  j->seq = j->world->attach_seq++;
  if (is_permanent(j)) {
    // relative values are (re)computed by the anchor/axis setters
    mark_struct(j->world);
  }
}
int dJointGetType(dJointID j) { return j->type; }
dBodyID dJointGetBody(dJointID j, int i) {
  if (i < 0 || i > 1) return nullptr;
  if (j->flags & JF_REVERSE) return j->node[1 - i];
  return j->node[i];
}
void dJointSetFeedback(dJointID j, dJointFeedback *fb) { j->fb = fb; }
dJointFeedback *dJointGetFeedback(dJointID j) { return j->fb; }

void dJointSetBallAnchor(dJointID j, dReal x, dReal y, dReal z) {
  pull_state(j->world);
  HPose a, b;
  if (j->node[0]) a = pose_of(j->node[0]);
  if (j->node[1]) b = pose_of(j->node[1]);
  h_set_anchors(j->dev, j->node[0] ? &a : nullptr, j->node[1] ? &b : nullptr, x, y, z);
  joint_changed(j);
}
void dJointSetBallParam(dJointID j, int p, dReal v) {
  if (p == dParamCFM) j->dev.cfm = v;
  else if (p == dParamERP) j->dev.erp = v;
  joint_changed(j);
}
void dJointSetHingeAnchor(dJointID j, dReal x, dReal y, dReal z) {
  pull_state(j->world);
  HPose a, b;
  if (j->node[0]) a = pose_of(j->node[0]);
  if (j->node[1]) b = pose_of(j->node[1]);
  h_set_anchors(j->dev, j->node[0] ? &a : nullptr, j->node[1] ? &b : nullptr, x, y, z);
  h_hinge_qrel(j->dev, j->node[0] ? &a : nullptr, j->node[1] ? &b : nullptr);
  joint_changed(j);
}
void dJointSetHingeAxis(dJointID j, dReal x, dReal y, dReal z) {
  pull_state(j->world);
  HPose a, b;
  if (j->node[0]) a = pose_of(j->node[0]);
  if (j->node[1]) b = pose_of(j->node[1]);
  h_set_axes(j->dev, j->node[0] ? &a : nullptr, j->node[1] ? &b : nullptr, x, y, z, true, true);
  h_hinge_qrel(j->dev, j->node[0] ? &a : nullptr, j->node[1] ? &b : nullptr);
  joint_changed(j);
}
void dJointSetHingeParam(dJointID j, int p, dReal v) {
  h_limot_set(j->dev.lm1, p, v);
  joint_changed(j);
}
void dJointSetHinge2Anchor(dJointID j, dReal x, dReal y, dReal z) {
  pull_state(j->world);
  HPose a, b;
  if (j->node[0]) a = pose_of(j->node[0]);
  if (j->node[1]) b = pose_of(j->node[1]);
  h_set_anchors(j->dev, j->node[0] ? &a : nullptr, j->node[1] ? &b : nullptr, x, y, z);
  if (j->node[0] && j->node[1]) h_hinge2_v1v2(j->dev, &a, &b);
  joint_changed(j);
}
void dJointSetHinge2Axis1(dJointID j, dReal x, dReal y, dReal z) {
  pull_state(j->world);
  if (!j->node[0]) return;
  HPose a = pose_of(j->node[0]), b;
  if (j->node[1]) b = pose_of(j->node[1]);
  h_set_axes(j->dev, &a, j->node[1] ? &b : nullptr, x, y, z, true, false);
  if (j->node[1]) {
    h_hinge2_axis_info(j->dev, &a, &b, j->dev.s0, j->dev.c0);
    h_hinge2_v1v2(j->dev, &a, &b);
  }
  joint_changed(j);
}
void dJointSetHinge2Axis2(dJointID j, dReal x, dReal y, dReal z) {
  pull_state(j->world);
  if (!j->node[0] || !j->node[1]) return;
  HPose a = pose_of(j->node[0]), b = pose_of(j->node[1]);
  h_set_axes(j->dev, &a, &b, x, y, z, false, true);
  h_hinge2_axis_info(j->dev, &a, &b, j->dev.s0, j->dev.c0);
  h_hinge2_v1v2(j->dev, &a, &b);
  joint_changed(j);
}
void dJointSetHinge2Param(dJointID j, int p, dReal v) {  // hinge2.cpp:259-274
  if ((p & 0xff00) == 0x100) h_limot_set(j->dev.lm2, p & 0xff, v);
  else if (p == dParamSuspensionERP) j->dev.susp_erp = v;
  else if (p == dParamSuspensionCFM) j->dev.susp_cfm = v;
  else h_limot_set(j->dev.lm1, p, v);
  joint_changed(j);
}
void dJointAddHinge2Torques(dJointID j, dReal t1, dReal t2) {  // hinge2.cpp:393-410
  if (!j->node[0] || !j->node[1]) return;
  pull_state(j->world);
  float a1[3], a2[3], t[3];
  h_mv(a1, j->node[0]->R, j->dev.axis1);
  h_mv(a2, j->node[1]->R, j->dev.axis2);
  for (int k = 0; k < 3; k++) t[k] = a1[k] * t1 + a2[k] * t2;
  dBodyAddTorque(j->node[0], t[0], t[1], t[2]);
  dBodyAddTorque(j->node[1], -t[0], -t[1], -t[2]);
}
// ode.cpp:1518-1536
int dAreConnected(dBodyID a, dBodyID b) {
  for (dxJoint *j : a->jlist) {
    dxBody *other = (j->node[0] == a) ? j->node[1] : j->node[0];
    if (other == b) return 1;
  }
  return 0;
}
int dAreConnectedExcluding(dBodyID a, dBodyID b, int joint_type) {
  for (dxJoint *j : a->jlist) {
    dxBody *other = (j->node[0] == a) ? j->node[1] : j->node[0];
    if (j->type != joint_type && other == b) return 1;
  }
  return 0;
}

// ---------------------------------------------------------------- spaces
static dxSpace *make_space(dSpaceID parent, int cls) {
  dxSpace *s = new dxSpace();
  s->cls = cls;
  std::memset(s->R, 0, sizeof(s->R));
  s->R[0] = s->R[5] = s->R[10] = 1;
  g_spaces.push_back(s);
  if (parent) dSpaceAdd(parent, s);   // collision_space.cpp:85-93
  return s;
}
dSpaceID dSimpleSpaceCreate(dSpaceID parent) { return make_space(parent, dSimpleSpaceClass); }
dSpaceID dHashSpaceCreate(dSpaceID parent) { return make_space(parent, dHashSpaceClass); }
void dHashSpaceSetLevels(dSpaceID s, int lo, int hi) {
  s->minlevel = lo;
  s->maxlevel = hi;
}
void dSpaceSetCleanup(dSpaceID s, int mode) { s->cleanup = mode ? 1 : 0; }
int dSpaceGetCleanup(dSpaceID s) { return s->cleanup; }
void dSpaceSetSublevel(dSpaceID s, int level) { s->sublevel = level; }
int dSpaceGetSublevel(dSpaceID s) { return s->sublevel; }
int dGeomIsSpace(dGeomID g) { return is_space(g) ? 1 : 0; }
dSpaceID dGeomGetSpace(dGeomID g) { return g->space; }
void dSpaceDestroy(dSpaceID s) {
  // the space owns its geoms unless its cleanup mode is off
  // (collision_space.cpp:96-118); a space leaves the space holding it
  std::vector<dxGeom *> gl = s->list;
  for (dxGeom *g : gl) {
    if (s->cleanup) dGeomDestroy(g);
    else dSpaceRemove(s, g);
  }
  if (s->space) dSpaceRemove(s->space, s);
  for (dxWorld *w : std::vector<dxWorld *>{s->world})
    if (w && w->mir_space == s) { w->mir_space = nullptr; mark_struct(w); }
  auto it = std::find(g_spaces.begin(), g_spaces.end(), s);
  if (it != g_spaces.end()) g_spaces.erase(it);
  delete s;
}
void dSpaceAdd(dSpaceID s, dGeomID g) {
  if (g->space || g == s) return;
  g->space = s;
  s->list.insert(s->list.begin(), g);
  g->dirty = true;
  geom_moved(s);   // the space, and the spaces above it, become dirty (collision_space.cpp:160-178)
  dxWorld *w = space_world(s);
  if (w) mark_struct(w);
}
void dSpaceRemove(dSpaceID s, dGeomID g) {
  auto it = std::find(s->list.begin(), s->list.end(), g);
  if (it != s->list.end()) s->list.erase(it);
  g->space = nullptr;
  geom_moved(s);
  if (s->world) mark_struct(s->world);
}
int dSpaceQuery(dSpaceID s, dGeomID g) { return g && g->space == s ? 1 : 0; }
dGeomID dSpaceGetGeom(dSpaceID s, int i) {
  if (i < 0 || i >= (int)s->list.size()) return nullptr;
  return s->list[i];
}
int dSpaceGetNumGeoms(dSpaceID s) { return (int)s->list.size(); }

// ---------------------------------------------------------------- geoms
static dxGeom *make_geom(dSpaceID s, int cls) {
  dxGeom *g = new dxGeom();
  g->cls = cls;
  std::memset(g->R, 0, sizeof(g->R));
  g->R[0] = g->R[5] = g->R[10] = 1;
  if (s) dSpaceAdd(s, g);
  return g;
}
dGeomID dCreateSphere(dSpaceID s, dReal r) {
  dxGeom *g = make_geom(s, dSphereClass);
  g->param[0] = r;
  return g;
}
dGeomID dCreateBox(dSpaceID s, dReal lx, dReal ly, dReal lz) {
  dxGeom *g = make_geom(s, dBoxClass);
  g->param[0] = lx; g->param[1] = ly; g->param[2] = lz;
  return g;
}
dGeomID dCreatePlane(dSpaceID s, dReal a, dReal b, dReal c, dReal d) {
  dxGeom *g = make_geom(s, dPlaneClass);
  float l = a * a + b * b + c * c;   // plane.cpp:49-76
  if (l > 0) {
    l = 1.0f / std::sqrt(l);
    g->param[0] = a * l; g->param[1] = b * l; g->param[2] = c * l; g->param[3] = d * l;
  } else {
    g->param[0] = 1;
  }
  return g;
}
dGeomID dCreateCapsule(dSpaceID s, dReal radius, dReal length) {   // capsule.cpp:77-80
  dxGeom *g = make_geom(s, dCapsuleClass);
  g->param[0] = radius;
  g->param[1] = length;
  return g;
}
void dGeomCapsuleSetParams(dGeomID g, dReal radius, dReal length) {
  if (g->cls != dCapsuleClass) {
    report_debug("argument not a ccylinder");
    return;
  }
  g->param[0] = radius;
  g->param[1] = length;
  geom_moved(g);
  dxWorld *w = g->body ? g->body->world : (g->space ? space_world(g->space) : nullptr);
  if (w) mark_struct(w);
}
void dGeomCapsuleGetParams(dGeomID g, dReal *radius, dReal *length) {
  *radius = g->param[0];
  *length = g->param[1];
}
dGeomID dCreateCylinder(dSpaceID s, dReal radius, dReal length) {   // cylinder.cpp:47-56, 80-83
  dxGeom *g = make_geom(s, dCylinderClass);
  g->param[0] = radius;
  g->param[1] = length;
  return g;
}
void dGeomCylinderSetParams(dGeomID g, dReal radius, dReal length) {
  if (g->cls != dCylinderClass) {
    report_debug("argument not a ccylinder");
    return;
  }
  g->param[0] = radius;
  g->param[1] = length;
  geom_moved(g);
  dxWorld *w = g->body ? g->body->world : (g->space ? space_world(g->space) : nullptr);
  if (w) mark_struct(w);
}
void dGeomCylinderGetParams(dGeomID g, dReal *radius, dReal *length) {
  *radius = g->param[0];
  *length = g->param[1];
}

// ---------------------------------------------------------------- rays
// ray.cpp:47-176; the length is param[0], the direction column 2 of R
dGeomID dCreateRay(dSpaceID s, dReal length) {
  dxGeom *g = make_geom(s, dRayClass);
  g->param[0] = length;
  return g;
}
static bool is_ray(dGeomID g) {
  if (g && g->cls == dRayClass) return true;
  report_debug("argument not a ray");
  return false;
}
static dxWorld *geom_world(dxGeom *g) {
  return g->body ? g->body->world : (g->space ? space_world(g->space) : nullptr);
}
void dGeomRaySetLength(dGeomID g, dReal length) {
  if (!is_ray(g)) return;
  g->param[0] = length;
  geom_moved(g);
  if (dxWorld *w = geom_world(g)) mark_struct(w);
}
dReal dGeomRayGetLength(dGeomID g) { return is_ray(g) ? g->param[0] : 0; }
void dGeomRaySet(dGeomID g, dReal px, dReal py, dReal pz, dReal dx, dReal dy, dReal dz) {
  if (!is_ray(g)) return;
  float n[3] = {dx, dy, dz};
  // dNormalize3 (odemath.h:383): scaled by the largest component first
  float a0 = std::fabs(n[0]), a1 = std::fabs(n[1]), a2 = std::fabs(n[2]);
  int k = a1 > a0 ? (a2 > a1 ? 2 : 1) : (a2 > a0 ? 2 : 0);
  float amax = std::fabs(n[k]);
  if (amax > 0) {
    for (int i = 0; i < 3; i++) n[i] /= amax;
    float l = 1.0f / std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int i = 0; i < 3; i++) n[i] *= l;
  } else {
    n[0] = 1; n[1] = 0; n[2] = 0;
  }
  if (g->body) {
    // the reference writes the body's own pose here (its position and the
    // third column of its rotation)
    const float *R = g->body->R;
    if (n[0] != R[2] || n[1] != R[6] || n[2] != R[10]) {
      report_error("dGeomRaySet: the ray is attached to a body; its direction is the body's z axis here");
      return;
    }
    dBodySetPosition(g->body, px, py, pz);
    return;
  }
  g->pos[0] = px; g->pos[1] = py; g->pos[2] = pz;
  g->R[2] = n[0]; g->R[6] = n[1]; g->R[10] = n[2];
  geom_moved(g);
  if (dxWorld *w = geom_world(g)) mark_struct(w);
}
void dGeomRayGet(dGeomID g, dVector3 start, dVector3 dir) {
  if (!is_ray(g)) return;
  const float *p = dGeomGetPosition(g), *R = dGeomGetRotation(g);
  for (int i = 0; i < 3; i++) {
    start[i] = p[i];
    dir[i] = R[4 * i + 2];
  }
}
enum { RAY_FIRST_CONTACT = 1, RAY_BACKFACE_CULL = 2, RAY_CLOSEST_HIT = 4 };
void dGeomRaySetParams(dGeomID g, int first_contact, int backface_cull) {
  if (!is_ray(g)) return;
  g->ray_flags = (g->ray_flags & RAY_CLOSEST_HIT) | (first_contact ? RAY_FIRST_CONTACT : 0) |
                 (backface_cull ? RAY_BACKFACE_CULL : 0);
}
void dGeomRayGetParams(dGeomID g, int *first_contact, int *backface_cull) {
  if (!is_ray(g)) return;
  *first_contact = (g->ray_flags & RAY_FIRST_CONTACT) != 0;
  *backface_cull = (g->ray_flags & RAY_BACKFACE_CULL) != 0;
}
void dGeomRaySetClosestHit(dGeomID g, int closest_hit) {
  if (!is_ray(g)) return;
  if (closest_hit) g->ray_flags |= RAY_CLOSEST_HIT;
  else g->ray_flags &= ~RAY_CLOSEST_HIT;
}
int dGeomRayGetClosestHit(dGeomID g) { return is_ray(g) ? (g->ray_flags & RAY_CLOSEST_HIT) != 0 : 0; }

void dGeomDestroy(dGeomID g) {
  if (is_space(g)) {
    dSpaceDestroy(static_cast<dxSpace *>(g));
    return;
  }
  if (g->space) dSpaceRemove(g->space, g);
  if (g->body) {
    auto &L = g->body->geoms;
    L.erase(std::find(L.begin(), L.end(), g));
    mark_struct(g->body->world);
  }
  delete g;
}
void dGeomSetData(dGeomID g, void *d) { g->data = d; }
void *dGeomGetData(dGeomID g) { return g->data; }
void dGeomSetBody(dGeomID g, dBodyID b) {  // collision_kernel.cpp:507-556
  if (b && g->cls == dPlaneClass) {   // planes are not placeable (collision_kernel.cpp:510)
    report_debug("geom must be placeable");
    return;
  }
  if (b) {
    if (g->body != b) {
      if (g->body) {
        auto &L = g->body->geoms;
        L.erase(std::find(L.begin(), L.end(), g));
      }
      g->body = b;
      b->geoms.insert(b->geoms.begin(), g);
    }
    geom_moved(g);
    if (g->space && !g->space->world) g->space->world = b->world;
    mark_struct(b->world);
  } else if (g->body) {
    geom_detach_body(g);
  }
}
dBodyID dGeomGetBody(dGeomID g) { return g->body; }
void dGeomSetPosition(dGeomID g, dReal x, dReal y, dReal z) {
  if (g->body) { dBodySetPosition(g->body, x, y, z); return; }
  g->pos[0] = x; g->pos[1] = y; g->pos[2] = z;
  geom_moved(g);
  if (g->space && g->space->world) mark_struct(g->space->world);
}
void dGeomSetRotation(dGeomID g, const dMatrix3 R) {
  if (g->body) { dBodySetRotation(g->body, R); return; }
  std::memcpy(g->R, R, sizeof(g->R));
  geom_moved(g);
  if (g->space && g->space->world) mark_struct(g->space->world);
}
void dGeomSetQuaternion(dGeomID g, const dQuaternion q) {
  if (g->body) { dBodySetQuaternion(g->body, q); return; }
  h_quat_to_rot(q, g->R);
  geom_moved(g);
  if (g->space && g->space->world) mark_struct(g->space->world);
}
const dReal *dGeomGetPosition(dGeomID g) {
  if (g->body) return dBodyGetPosition(g->body);
  return g->pos;
}
const dReal *dGeomGetRotation(dGeomID g) {
  if (g->body) return dBodyGetRotation(g->body);
  return g->R;
}
int dGeomGetClass(dGeomID g) { return g->cls; }
void dGeomSetCategoryBits(dGeomID g, unsigned long bits) {
  g->cat = bits;
  if (g->space && g->space->world) mark_struct(g->space->world);
}
void dGeomSetCollideBits(dGeomID g, unsigned long bits) {
  g->col = bits;
  if (g->space && g->space->world) mark_struct(g->space->world);
}
dReal dGeomSphereGetRadius(dGeomID g) { return g->param[0]; }
void dGeomBoxGetLengths(dGeomID g, dVector3 r) { for (int k = 0; k < 3; k++) r[k] = g->param[k]; }
void dGeomPlaneGetParams(dGeomID g, dVector4 r) { for (int k = 0; k < 4; k++) r[k] = g->param[k]; }
void dGeomGetAABB(dGeomID g, dReal aabb[6]) {
  if (is_space(g)) {
    // union of the geoms' boxes, zero for an empty space (collision_space.cpp:120-145)
    const dxSpace *sp = static_cast<const dxSpace *>(g);
    if (sp->list.empty()) {
      for (int k = 0; k < 6; k++) aabb[k] = 0;
      return;
    }
    for (int k = 0; k < 6; k++) aabb[k] = (k & 1) ? -INFINITY : INFINITY;
    for (dxGeom *m : sp->list) {
      float b[6];
      dGeomGetAABB(m, b);
      for (int k = 0; k < 6; k += 2) if (b[k] < aabb[k]) aabb[k] = b[k];
      for (int k = 1; k < 6; k += 2) if (b[k] > aabb[k]) aabb[k] = b[k];
    }
    return;
  }
  // box.cpp:60-77, sphere.cpp:59-67, plane.cpp:80-107 on the host copies
  const float *p = dGeomGetPosition(g);
  const float *R = dGeomGetRotation(g);
  if (g->cls == dSphereClass) {
    for (int k = 0; k < 3; k++) { aabb[2 * k] = p[k] - g->param[0]; aabb[2 * k + 1] = p[k] + g->param[0]; }
  } else if (g->cls == dBoxClass) {
    for (int k = 0; k < 3; k++) {
      float e = 0.5f * (std::fabs(R[4 * k] * g->param[0]) + std::fabs(R[4 * k + 1] * g->param[1]) +
                        std::fabs(R[4 * k + 2] * g->param[2]));
      aabb[2 * k] = p[k] - e;
      aabb[2 * k + 1] = p[k] + e;
    }
  } else if (g->cls == dCapsuleClass) {   // capsule.cpp:60-74
    for (int k = 0; k < 3; k++) {
      float e = std::fabs(R[4 * k + 2] * g->param[1]) * 0.5f + g->param[0];
      aabb[2 * k] = p[k] - e;
      aabb[2 * k + 1] = p[k] + e;
    }
  } else if (g->cls == dCylinderClass) {  // cylinder.cpp:58-74
    for (int k = 0; k < 3; k++) {
      float e = std::fabs(R[4 * k] * g->param[0]) + std::fabs(R[4 * k + 1] * g->param[0]) +
                0.5f * std::fabs(R[4 * k + 2] * g->param[1]);
      aabb[2 * k] = p[k] - e;
      aabb[2 * k + 1] = p[k] + e;
    }
  } else if (g->cls == dRayClass) {       // ray.cpp:56-88
    for (int k = 0; k < 3; k++) {
      float e = p[k] + R[4 * k + 2] * g->param[0];
      aabb[2 * k] = p[k] < e ? p[k] : e;
      aabb[2 * k + 1] = p[k] < e ? e : p[k];
    }
  } else {
    for (int k = 0; k < 6; k++) aabb[k] = (k & 1) ? INFINITY : -INFINITY;
    const float *n = g->param;
    int ax = (n[1] == 0 && n[2] == 0) ? 0 : ((n[0] == 0 && n[2] == 0) ? 1 : ((n[0] == 0 && n[1] == 0) ? 2 : -1));
    if (ax >= 0) {
      aabb[2 * ax] = n[ax] > 0 ? -INFINITY : -n[3];
      aabb[2 * ax + 1] = n[ax] > 0 ? n[3] : INFINITY;
    }
  }
}

// dSpaceCollide (collision.h:825): GPU broadphase + speculative narrowphase,
// then the user's callback for every pair in the reference's report order.
void dSpaceCollide(dSpaceID s, void *data, dNearCallback *cb) {
  if (!s || !cb || s->list.size() < 2) {
    if (s) for (dxGeom *g : s->list) g->dirty = false;
    return;
  }
  for (dxGeom *g : s->list)
    if (is_space(g)) {
      report_error("dSpaceCollide: the space holds a space; collide the spaces with dSpaceCollide2 in this build");
      return;
    }
  dxWorld *w = space_world(s);
  if (!w) {
    report_error("dSpaceCollide: the space holds no geom attached to a body");
    return;
  }
  if (!sync_mirror(w, s)) return;
  BatchImpl *m = w->mir;
  if (!world_params_upload(w, 0, s->cached_maxc, 0)) return;
  refresh_dev(m);
  DevBatch D = m->dev;
  D.pairs_out = w->pairs_out.as<PairRec>();
  D.pcount_out = w->pcount_out.as<int>();
  D.contacts_out = w->contacts_out.as<ContactRec>();
  D.ctrl = nullptr;
  if (launch_world_kernel(KM_API_COLLIDE, D, 0, 1, 0, m->wg, s->cached_maxc, m->stream) != hipSuccess ||
      hipStreamSynchronize(m->stream) != hipSuccess) {
    report_error("dSpaceCollide: collide kernel failed");
    return;
  }
  long long cnt[5];
  batch_counters(m, cnt);
  if (cnt[4]) report_error("dSpaceCollide: device capacity exceeded (flags %lld)", cnt[4]);
  int np = (int)cnt[0], nk = (int)cnt[1];
  std::vector<PairRec> pr(np);
  std::vector<int> pc(np);
  std::vector<ContactRec> cr(nk);
  if (np) {
    (void)hipMemcpy(pr.data(), w->pairs_out.p, sizeof(PairRec) * np, hipMemcpyDeviceToHost);
    (void)hipMemcpy(pc.data(), w->pcount_out.p, sizeof(int) * np, hipMemcpyDeviceToHost);
  }
  if (nk) (void)hipMemcpy(cr.data(), w->contacts_out.p, sizeof(ContactRec) * nk, hipMemcpyDeviceToHost);
  for (dxGeom *g : s->list) g->dirty = false;
  std::vector<dxGeom *> mg = w->mgeoms;
  int off = 0;
  for (int i = 0; i < np; i++) {
    CollideCtx saved = g_ctx;
    g_ctx.active = true;
    g_ctx.space = s;
    g_ctx.g1 = mg[pr[i].g1];
    g_ctx.g2 = mg[pr[i].g2];
    g_ctx.con = cr.data() + off;
    g_ctx.n = pc[i];
    g_ctx.maxc = s->cached_maxc;
    cb(data, g_ctx.g1, g_ctx.g2);
    g_ctx = saved;
    off += pc[i];
  }
}

// pose and shape of a geom for the device colliders, from the host copies
static GeomView geom_view(dxGeom *g) {
  GeomView v;
  v.type = device_type(g->cls);
  const float *p = g->body ? g->body->pos : g->pos;
  const float *R = g->body ? g->body->R : g->R;
  v.p = make_float3(p[0], p[1], p[2]);
  v.R.r0 = make_float3(R[0], R[1], R[2]);
  v.R.r1 = make_float3(R[4], R[5], R[6]);
  v.R.r2 = make_float3(R[8], R[9], R[10]);
  v.prm = make_float4(g->param[0], g->param[1], g->param[2], g->param[3]);
  return v;
}

// The checks and the cached contacts of dCollide (collision_kernel.cpp:
// 292-339): returns 1 with *n set when the call is answered without the
// device (no contacts requested, the same geom or body, or the pair the
// current dSpaceCollide callback was called for), else 0.
static int collide_host_answer(dGeomID o1, dGeomID o2, int maxc, dContactGeom *contact, int skip, int *n) {
  *n = 0;
  if (maxc <= 0 || o1 == o2) return 1;
  if (device_type(o1->cls) < 0 || device_type(o2->cls) < 0) {
    report_error("dCollide: geom classes %d and %d have no collider in this build (spaces are collided with "
                 "dSpaceCollide2)", o1->cls, o2->cls);
    return 1;
  }
  if (o1->body && o1->body == o2->body) return 1;
  if (g_ctx.active && g_ctx.g1 == o1 && g_ctx.g2 == o2 && g_ctx.space) {
    int want = maxc > MAX_SLOTS ? MAX_SLOTS : maxc;
    int have = g_ctx.maxc > MAX_SLOTS ? MAX_SLOTS : g_ctx.maxc;
    g_ctx.space->cached_maxc = maxc;
    if (want == have) {
      for (int i = 0; i < g_ctx.n; i++) {
        const ContactRec &c = g_ctx.con[i];
        write_contact((dContactGeom *)((char *)contact + (size_t)skip * i), c.pos, c.normal, c.depth, o1, o2);
      }
      *n = g_ctx.n;
      return 1;
    }
  }
  return 0;
}

static int write_pair_contacts(dGeomID o1, dGeomID o2, int maxc, const CPoint *pts, int n, dContactGeom *contact,
                               int skip) {
  if (n > maxc) n = maxc;
  for (int i = 0; i < n; i++) {
    float p[3] = {pts[i].pos.x, pts[i].pos.y, pts[i].pos.z}, nn[3] = {pts[i].n.x, pts[i].n.y, pts[i].n.z};
    write_contact((dContactGeom *)((char *)contact + (size_t)skip * i), p, nn, pts[i].depth, o1, o2);
  }
  return n;
}

int dCollide(dGeomID o1, dGeomID o2, int flags, dContactGeom *contact, int skip) {
  int maxc = flags & 0xffff, n = 0;
  if (collide_host_answer(o1, o2, maxc, contact, skip, &n)) return n;
  // one pair on the device with the current poses
  dxWorld *w = o1->body ? o1->body->world : (o2->body ? o2->body->world : nullptr);
  if (w) pull_state(w);
  static DevBuf out, nout;
  static hipStream_t st = nullptr;
  if (!st && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return 0;
  if (!out.alloc(sizeof(CPoint) * MAX_SLOTS) || !nout.alloc(16)) return 0;
  if (launch_collide_one(geom_view(o1), geom_view(o2), maxc, out.as<CPoint>(), nout.as<int>(), st) != hipSuccess)
    return 0;
  CPoint pts[MAX_SLOTS];
  if (hipMemcpyAsync(&n, nout.p, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(pts, out.p, sizeof(pts), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return 0;
  return write_pair_contacts(o1, o2, maxc, pts, n, contact, skip);
}

// dCollide without device allocation: the one-pair collide runs in the
// library's per-GPU scratch array (set up by dInitODE2 or on the GPU's first
// call) on the calling thread's per-thread stream.  At most MAX_SLOTS (8)
// contacts may be requested, the most any collider of this library returns.
int dCollideNoAlloc(dGeomID o1, dGeomID o2, int flags, dContactGeom *contact, int skip) {
  int maxc = flags & 0xffff, n = 0;
  if (maxc > MAX_SLOTS) {
    report_error("dCollideNoAlloc: %d contacts requested; at most %d per call", maxc, MAX_SLOTS);
    return 0;
  }
  if (collide_host_answer(o1, o2, maxc, contact, skip, &n)) return n;
  dxWorld *w = o1->body ? o1->body->world : (o2->body ? o2->body->world : nullptr);
  if (w) pull_state(w);
  hipError_t err = hipSuccess;
  void *addr = noalloc_scratch(&err);
  if (!addr) {
    report_error("dCollideNoAlloc: the scratch array is not available on this device (%s)", hipGetErrorString(err));
    return 0;
  }
  CPoint pts[MAX_SLOTS];
  err = collide_one_noalloc(addr, geom_view(o1), geom_view(o2), maxc, pts, &n, hipStreamPerThread);
  if (err != hipSuccess) {
    report_error("dCollideNoAlloc: %s", hipGetErrorString(err));
    return 0;
  }
  return write_pair_contacts(o1, o2, maxc, pts, n, contact, skip);
}

// ---------------------------------------------------------------- dSpaceCollide2

namespace {

// a pair dSpaceCollide2 reports, in callback argument order; `space` is the
// space whose geoms were tested (it keeps the contact count dCollide asks for)
struct Pair2 {
  dxGeom *a, *b;
  dxSpace *space;
};

// collideAABBs (collision_space_internal.h:48-82) without the callback
bool aabbs_may_touch(const dxGeom *g1, const dxGeom *g2, const float *b1, const float *b2) {
  if (g1->body == g2->body && g1->body) return false;
  if (((g1->cat & g2->col) || (g2->cat & g1->col)) == 0) return false;
  return !(b1[0] > b2[1] || b1[1] < b2[0] || b1[2] > b2[3] || b1[3] < b2[2] || b1[4] > b2[5] || b1[5] < b2[4]);
}

// cleanGeoms (collision_space.cpp:239-251): the space's geoms, and those of
// the spaces it holds, are clean once tested
void clean_space(dxSpace *s) {
  for (dxGeom *g : s->list) {
    if (is_space(g)) clean_space(static_cast<dxSpace *>(g));
    g->dirty = false;
  }
}

// dxHashSpace::collide2 (collision_space.cpp:585-603): every geom of S, in
// list order, against G; `swap` gives the callback (G, g) instead of (g, G)
void collide2_space_geom(dxSpace *S, dxGeom *G, bool swap, std::vector<Pair2> &out) {
  clean_space(S);
  float bg[6];
  dGeomGetAABB(G, bg);
  for (dxGeom *g : S->list) {
    float b[6];
    dGeomGetAABB(g, b);
    if (aabbs_may_touch(g, G, b, bg)) out.push_back(swap ? Pair2{G, g, S} : Pair2{g, G, S});
  }
}

void pull_geom_world(dxGeom *g) {
  if (g->body) pull_state(g->body->world);
  if (is_space(g))
    for (dxGeom *m : static_cast<dxSpace *>(g)->list) pull_geom_world(m);
}

}  // namespace

// dSpaceCollide2 (collision.h:845, collision_space.cpp:768-837): the pairs
// of geoms of g1 and g2 (either may be a space; between two spaces the one
// with fewer geoms is iterated outside, the callback keeping the (g1-side,
// g2-side) argument order) whose boxes touch, reported in the reference's
// order; a space against itself is dSpaceCollide.  The geom-geom pairs'
// narrowphase runs in one device launch first, so that dCollide inside the
// callback returns the pair's contacts; a pair holding a space goes to the
// callback as it is (the callback descends with dSpaceCollide2).
void dSpaceCollide2(dGeomID g1, dGeomID g2, void *data, dNearCallback *cb) {
  if (!g1 || !g2 || !cb) return;
  dxSpace *s1 = is_space(g1) ? static_cast<dxSpace *>(g1) : nullptr;
  dxSpace *s2 = is_space(g2) ? static_cast<dxSpace *>(g2) : nullptr;
  if (s1 && s2 && s1->sublevel != s2->sublevel) {
    if (s1->sublevel > s2->sublevel) s2 = nullptr;
    else s1 = nullptr;
  }
  if (s1 && s2 && s1 == s2) {
    dSpaceCollide(s1, data, cb);
    return;
  }
  pull_geom_world(g1);
  pull_geom_world(g2);
  std::vector<Pair2> pairs;
  if (s1 && s2) {
    if (s1->list.size() < s2->list.size()) {
      std::vector<dxGeom *> outer = s1->list;
      for (dxGeom *g : outer) collide2_space_geom(s2, g, true, pairs);
    } else {
      std::vector<dxGeom *> outer = s2->list;
      for (dxGeom *g : outer) collide2_space_geom(s1, g, false, pairs);
    }
  } else if (s1) {
    collide2_space_geom(s1, g2, false, pairs);
  } else if (s2) {
    collide2_space_geom(s2, g1, true, pairs);
  } else {
    float b1[6], b2[6];
    dGeomGetAABB(g1, b1);
    dGeomGetAABB(g2, b2);
    if (aabbs_may_touch(g1, g2, b1, b2)) pairs.push_back(Pair2{g1, g2, nullptr});
  }
  if (pairs.empty()) return;
  // speculative narrowphase of the geom-geom pairs
  std::vector<int> slot(pairs.size(), -1);
  std::vector<GeomView> va, vb;
  int maxc = 1;
  for (size_t i = 0; i < pairs.size(); i++) {
    const Pair2 &p = pairs[i];
    if (is_space(p.a) || is_space(p.b)) continue;
    slot[i] = (int)va.size();
    va.push_back(geom_view(p.a));
    vb.push_back(geom_view(p.b));
    const int mc = p.space ? p.space->cached_maxc : 4;
    if (mc > maxc) maxc = mc;
  }
  if (maxc > MAX_SLOTS) maxc = MAX_SLOTS;
  const int np = (int)va.size();
  std::vector<int> cnt(np, 0);
  std::vector<CPoint> pts((size_t)np * MAX_SLOTS);
  if (np > 0) {
    static DevBuf dva, dvb, dout, dn;
    static hipStream_t st = nullptr;
    if (!st && hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return;
    if (!dva.alloc(sizeof(GeomView) * np) || !dvb.alloc(sizeof(GeomView) * np) ||
        !dout.alloc(sizeof(CPoint) * MAX_SLOTS * (size_t)np) || !dn.alloc(sizeof(int) * np)) {
      report_error("dSpaceCollide2: device allocation failed");
      return;
    }
    if (hipMemcpyAsync(dva.p, va.data(), sizeof(GeomView) * np, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(dvb.p, vb.data(), sizeof(GeomView) * np, hipMemcpyHostToDevice, st) != hipSuccess ||
        launch_collide_list(dva.as<GeomView>(), dvb.as<GeomView>(), np, maxc, dout.as<CPoint>(), dn.as<int>(), st) !=
            hipSuccess ||
        hipMemcpyAsync(cnt.data(), dn.p, sizeof(int) * np, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(pts.data(), dout.p, sizeof(CPoint) * pts.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
      report_error("dSpaceCollide2: narrowphase kernel failed");
      return;
    }
  }
  for (size_t i = 0; i < pairs.size(); i++) {
    const Pair2 &p = pairs[i];
    CollideCtx saved = g_ctx;
    std::vector<ContactRec> cr;
    if (slot[i] >= 0) {
      const int k = slot[i];
      for (int j = 0; j < cnt[k]; j++) {
        const CPoint &c = pts[(size_t)MAX_SLOTS * k + j];
        ContactRec r;
        std::memset(&r, 0, sizeof(r));
        r.pos[0] = c.pos.x; r.pos[1] = c.pos.y; r.pos[2] = c.pos.z; r.depth = c.depth;
        r.normal[0] = c.n.x; r.normal[1] = c.n.y; r.normal[2] = c.n.z;
        cr.push_back(r);
      }
      g_ctx.active = p.space != nullptr;
      g_ctx.space = p.space;
      g_ctx.g1 = p.a;
      g_ctx.g2 = p.b;
      g_ctx.con = cr.data();
      g_ctx.n = (int)cr.size();
      g_ctx.maxc = maxc;
    } else {
      g_ctx.active = false;
    }
    cb(data, p.a, p.b);
    g_ctx = saved;
  }
}

// ---------------------------------------------------------------- adjoint (ODE API form)
int dWorldSetTape(dWorldID w, int on) {
  if (!w) return 0;
  w->tape = on != 0;
  if (!w->tape) {
    w->tape_ok = false;
    std::snprintf(w->tape_why, sizeof(w->tape_why), "tape recording is off");
  }
  return 1;
}

// Reverse step of the last recorded dWorldQuickStep (adj_step.hpp on the
// world's device mirror): the LCP differentiated implicitly at the step's
// impulses on their active set (exact when the solve converged: raise
// dWorldSetQuickStepNumIterations until it does), the colliders and the
// integrator analytically.  grad_out / grad_in: 13 floats per body in
// creation order (pos 3, q 4, lvel 3, avel 3).
int dWorldQuickStepBackward(dWorldID w, const dReal *grad_out, dReal *grad_in) {
  if (!w || !grad_out || !grad_in) return 0;
  if (!w->tape_ok || !w->mir) {
    report_error("dWorldQuickStepBackward: %s", w->tape_why);
    return 0;
  }
  BatchImpl *m = w->mir;
  const int nb = m->nb, kcap = m->wi[0].kcap;
  // the contact policy of the recorded step, for the adjoint kernel
  WorldParams P0, P;
  OD_HIP(hipMemcpy(&P0, m->d.wp.p, sizeof(P0), hipMemcpyDeviceToHost));
  P = P0;
  P.surf_mode = w->tape_mode;
  P.mu = w->tape_mu;
  P.bounce = w->tape_bounce;
  P.bounce_vel = w->tape_bounce_vel;
  P.max_contacts = w->tape_maxc;
  OD_HIP(hipMemcpy(m->d.wp.p, &P, sizeof(P), hipMemcpyHostToDevice));
  DevBuf gbuf, gsc, stats, scr, off;
  const long long nscr = adj::adj_scratch_elems<float>(nb, kcap, m->nj);
  bool ok = gbuf.alloc(8 * sizeof(float4) * (size_t)(nb + 1)) && gsc.alloc(2 * sizeof(float)) &&
            stats.alloc(2 * sizeof(int)) && scr.alloc(sizeof(float) * (size_t)nscr) &&
            off.alloc(sizeof(long long));
  if (!ok) {
    OD_HIP(hipMemcpy(m->d.wp.p, &P0, sizeof(P0), hipMemcpyHostToDevice));
    report_error("dWorldQuickStepBackward: device allocation failed");
    return 0;
  }
  float4 *G = gbuf.as<float4>();
  const float4 *gin[4];
  float4 *gout[4];
  for (int k = 0; k < 4; k++) {
    gin[k] = G + (size_t)k * nb;
    gout[k] = G + (size_t)(4 + k) * nb;
  }
  std::vector<float4> h(4 * (size_t)nb, make_float4(0, 0, 0, 0));
  for (size_t bi = 0; bi < w->bodies.size(); bi++) {
    const int i = w->bodies[bi]->mindex;
    if (i < 0 || i >= nb) continue;
    const float *g = grad_out + 13 * bi;
    h[i] = make_float4(g[0], g[1], g[2], 0);
    h[nb + i] = make_float4(g[3], g[4], g[5], g[6]);
    h[2 * (size_t)nb + i] = make_float4(g[7], g[8], g[9], 0);
    h[3 * (size_t)nb + i] = make_float4(g[10], g[11], g[12], 0);
  }
  OD_HIP(hipMemcpy(G, h.data(), sizeof(float4) * h.size(), hipMemcpyHostToDevice));
  OD_HIP(hipMemset(gsc.p, 0, gsc.n));
  OD_HIP(hipMemset(stats.p, 0, stats.n));
  OD_HIP(hipMemset(off.p, 0, off.n));
  const float4 *st[4];
  const float4 *S = w->tape_st.as<float4>();
  for (int k = 0; k < 4; k++) st[k] = S + (size_t)k * nb;
  refresh_dev(m);
  hipError_t e = launch_adjoint_step(m->dev, st, w->tape_con.as<adj::TapeContact<float>>(), off.as<long long>(),
                                     w->tape_nk.as<int>(), gin, gout, gsc.as<float>(), gsc.as<float>() + 1,
                                     stats.as<int>(), scr.as<float>(), off.as<long long>(), 1e-6f, 1e-6f, 500, 256,
                                     m->nj ? w->tape_jnt.as<adj::TapeJoint<float>>() : nullptr, m->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(m->stream);
  OD_HIP(hipMemcpy(m->d.wp.p, &P0, sizeof(P0), hipMemcpyHostToDevice));
  if (e != hipSuccess) {
    report_error("dWorldQuickStepBackward: %s", hipGetErrorString(e));
    return 0;
  }
  OD_HIP(hipMemcpy(h.data(), gout[0], sizeof(float4) * h.size(), hipMemcpyDeviceToHost));
  int st2[2];
  OD_HIP(hipMemcpy(st2, stats.p, sizeof(st2), hipMemcpyDeviceToHost));
  if (st2[1] != 0) {
    report_error("dWorldQuickStepBackward: the colliders returned a different contact set on the recorded poses");
    return 0;
  }
  for (size_t bi = 0; bi < w->bodies.size(); bi++) {
    const int i = w->bodies[bi]->mindex;
    float *g = grad_in + 13 * bi;
    if (i < 0 || i >= nb) { for (int k = 0; k < 13; k++) g[k] = 0; continue; }
    const float4 a = h[i], q = h[nb + i], v = h[2 * (size_t)nb + i], r = h[3 * (size_t)nb + i];
    g[0] = a.x; g[1] = a.y; g[2] = a.z;
    g[3] = q.x; g[4] = q.y; g[5] = q.z; g[6] = q.w;
    g[7] = v.x; g[8] = v.y; g[9] = v.z;
    g[10] = r.x; g[11] = r.y; g[12] = r.z;
  }
  return 1;
}

}  // extern "C"
