// batch.cpp -- the dBatch* entry points: device state for many independent
// worlds, one persistent launch per step replayed from a HIP graph, the
// one-world-per-launch plain path, and the launch-shape autotuner.
#include <algorithm>
#include <chrono>
#include <cstdarg>
#include <numeric>

#include "host_common.hpp"
#include "odemi.h"
#include "odemi_internal.hpp"

namespace odemi {

namespace {
inline bool isnan_d(double v) { return std::isnan(v); }
}  // namespace

void destroy_graph(BatchImpl *b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

void refresh_dev(BatchImpl *b) {
  DevBatch &D = b->dev;
  D.nworlds = b->nworlds;
  // exact LCP workspaces (dBatchSetExactLCP; the ODE API's dWorldStep sets
  // its own per call)
  D.lcp_ws = b->lcp_rows > 0 ? b->lcp_ws.as<float>() : nullptr;
  D.lcp_ws_stride = b->lcp_rows > 0 ? lcp_ws_floats(b->lcp_rows) : 0;
  D.lcp_ws_rows = b->lcp_rows;
  D.order_mode = b->order_mode;
  D.wi = b->d.wi.as<WorldInfo>();
  D.wp = b->d.wp.as<WorldParams>();
  D.pos = b->d.pos.as<float4>(); D.q = b->d.q.as<float4>(); D.lvel = b->d.lvel.as<float4>();
  D.avel = b->d.avel.as<float4>(); D.facc = b->d.facc.as<float4>(); D.tacc = b->d.tacc.as<float4>();
  D.R = b->d.R.as<float4>(); D.I = b->d.I.as<float4>(); D.invI = b->d.invI.as<float4>();
  D.mass = b->d.mass.as<float>(); D.invMass = b->d.invMass.as<float>();
  D.gtype = b->d.gtype.as<int>(); D.gbody = b->d.gbody.as<int>(); D.gparam = b->d.gparam.as<float4>();
  D.gpos = b->d.gpos.as<float4>(); D.gR = b->d.gR.as<float4>(); D.grank = b->d.grank.as<int>();
  D.gcat = b->d.gcat.as<unsigned>(); D.gcol = b->d.gcol.as<unsigned>();
  D.joints = b->d.joints.as<JointDev>();
  D.adj_start = b->d.adj_start.as<int>(); D.adj_other = b->d.adj_other.as<int>();
  D.bgeom_start = b->d.bgeom_start.as<int>(); D.bgeom = b->d.bgeom.as<int>();
  D.cin = nullptr; D.cin_start = nullptr; D.cin_count = nullptr;
  D.ctrl = b->ctrl_set ? b->d.ctrl.as<float4>() : nullptr;
  D.pairs_out = nullptr; D.contacts_out = nullptr; D.pcount_out = nullptr;
  D.feedback = nullptr;
  D.counters = b->d.counters.as<int>();
  D.residual = b->d.residual.as<float>();
  D.work_counter = b->d.work_counter.as<int>();
  D.world_order = b->d.world_order.as<int>();
  D.scratch = b->d.scratch.as<char>();
  D.warm = b->d.warm.as<char>();
  D.warm_n = b->d.warm_n.as<int>();
  D.tape = nullptr;
  D.tape_off = nullptr;
  D.prof = b->prof_on ? b->d.prof.as<unsigned long long>() : nullptr;
  D.lds_bytes = b->lds;
  D.sor_lds_bytes = b->sor_lds;
}

// LDS limits of the current device and the default pool: half a CU's LDS
// less the kernel's static block, so that two workgroups share a CU.
static void init_lds(BatchImpl *b, int dev) {
  int per_cu = 0, per_block = 0, khz = 0;
  if (hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) != hipSuccess ||
      per_cu <= 0)
    per_cu = 65536;
  if (hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
      per_block <= 0)
    per_block = 65536;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess && khz > 0) b->wall_khz = khz;
  int st = world_kernel_static_lds();
  b->lds_cu = per_cu;
  b->lds_max = ((per_block < per_cu ? per_block : per_cu) - st) & ~255;
  if (b->lds_max < 0) b->lds_max = 0;
  if (b->lds_max > 0 && allow_world_kernel_lds(b->lds_max) != hipSuccess) {
    (void)hipGetLastError();
    b->lds_max = ((per_block < 65536 ? per_block : 65536) - st) & ~255;
  }
  int half = (per_cu / 2 - st - 256) & ~255;
  b->lds = half < b->lds_max ? half : b->lds_max;
  if (b->lds < 0) b->lds = 0;
}

template <class T>
static int upload(DevBuf &buf, const std::vector<T> &v) {
  if (!buf.alloc(sizeof(T) * (v.size() ? v.size() : 1))) {
    report_error("device allocation of %zu bytes failed", sizeof(T) * v.size());
    return 0;
  }
  if (!v.empty()) OD_HIP(hipMemcpy(buf.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
  return 1;
}

// Builds the joint records of one world from the descriptor with the
// semantics of the reference's setters (ball.cpp:90-97, hinge.cpp:153-210,
// hinge2.cpp:214-300, joint.cpp:472-509); bodies are in their initial pose.
static void build_joint(JointDev &j, const dBatchDesc *D, const dBatchWorldDesc &W, int jg, int b0, int seq,
                        const std::vector<HPose> &poses) {
  int type = D->joint_type[jg];
  h_joint_defaults(j, type, (float)W.erp, (float)W.cfm);
  int b1 = D->joint_b1[jg], b2 = D->joint_b2[jg];
  if (b1 < 0) { b1 = b2; b2 = -1; j.flags |= JF_REVERSE; }
  j.b1 = b1;
  j.b2 = b2;
  j.seq = seq;
  const HPose *p1 = b1 >= 0 ? &poses[b0 + b1] : nullptr;
  const HPose *p2 = b2 >= 0 ? &poses[b0 + b2] : nullptr;
  const double *a = D->joint_anchor + 3 * (size_t)jg;
  const double *x1 = D->joint_axis1 + 3 * (size_t)jg;
  const double *x2 = D->joint_axis2 + 3 * (size_t)jg;
  const double *prm = D->joint_param + ODEMI_NJP * (size_t)jg;
  if (type == JT_BALL) {
    h_set_anchors(j, p1, p2, (float)a[0], (float)a[1], (float)a[2]);
    if (!isnan_d(prm[13])) j.erp = (float)prm[13];
    if (!isnan_d(prm[6])) j.cfm = (float)prm[6];
  } else if (type == JT_HINGE) {
    h_set_anchors(j, p1, p2, (float)a[0], (float)a[1], (float)a[2]);
    h_set_axes(j, p1, p2, (float)x1[0], (float)x1[1], (float)x1[2], true, true);
    h_hinge_qrel(j, p1, p2);
    for (int k = 0; k < 9; k++)
      if (!isnan_d(prm[k])) h_limot_set(j.lm1, k, (float)prm[k]);
  } else if (type == JT_HINGE2) {
    h_set_anchors(j, p1, p2, (float)a[0], (float)a[1], (float)a[2]);
    if (p1) {
      h_set_axes(j, p1, p2, (float)x1[0], (float)x1[1], (float)x1[2], true, false);
      if (p2) {
        h_set_axes(j, p1, p2, (float)x2[0], (float)x2[1], (float)x2[2], false, true);
        h_hinge2_axis_info(j, p1, p2, j.s0, j.c0);
        h_hinge2_v1v2(j, p1, p2);
      }
    }
    for (int k = 0; k < 9; k++)
      if (!isnan_d(prm[k])) h_limot_set(j.lm1, k, (float)prm[k]);
    if (!isnan_d(prm[9])) j.susp_erp = (float)prm[9];
    if (!isnan_d(prm[10])) j.susp_cfm = (float)prm[10];
    if (!isnan_d(prm[11])) h_limot_set(j.lm2, 2, (float)prm[11]);
    if (!isnan_d(prm[12])) h_limot_set(j.lm2, 3, (float)prm[12]);
  }
}

static BatchArrays arrays_from_desc(const dBatchDesc *D, double cap_scale, float cell_scale) {
  BatchArrays A;
  int nb = 0, ng = 0, nj = 0;
  for (int w = 0; w < D->nworlds; w++) {
    nb += D->worlds[w].nb;
    ng += D->worlds[w].ng;
    nj += D->worlds[w].nj;
  }
  A.pos.resize(nb); A.q.resize(nb); A.lvel.resize(nb); A.avel.resize(nb);
  A.facc.assign(nb, make_float4(0, 0, 0, 0)); A.tacc.assign(nb, make_float4(0, 0, 0, 0));
  A.I.resize(3 * (size_t)nb); A.invI.resize(3 * (size_t)nb);
  A.mass.resize(nb); A.invMass.resize(nb);
  std::vector<HPose> poses(nb);
  for (int i = 0; i < nb; i++) {
    float qq[4];
    for (int k = 0; k < 4; k++) qq[k] = (float)D->body_q[4 * (size_t)i + k];
    h_normalize4(qq);  // dBodySetQuaternion normalises (ode.cpp:380-396)
    HPose &P = poses[i];
    for (int k = 0; k < 3; k++) P.pos[k] = (float)D->body_pos[3 * (size_t)i + k];
    P.pos[3] = 0;
    std::memcpy(P.q, qq, sizeof(qq));
    h_quat_to_rot(qq, P.R);
    A.pos[i] = make_float4(P.pos[0], P.pos[1], P.pos[2], 0);
    A.q[i] = make_float4(qq[0], qq[1], qq[2], qq[3]);
    A.lvel[i] = make_float4((float)D->body_lvel[3 * (size_t)i], (float)D->body_lvel[3 * (size_t)i + 1],
                            (float)D->body_lvel[3 * (size_t)i + 2], 0);
    A.avel[i] = make_float4((float)D->body_avel[3 * (size_t)i], (float)D->body_avel[3 * (size_t)i + 1],
                            (float)D->body_avel[3 * (size_t)i + 2], 0);
    float Im[12] = {0}, Iv[12];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) Im[4 * r + c] = (float)D->body_I[9 * (size_t)i + 3 * r + c];
    if (!h_invert_pd3(Im, Iv)) {  // dBodySetMass (ode.cpp:492-509)
      std::memset(Iv, 0, sizeof(Iv));
      Iv[0] = Iv[5] = Iv[10] = 1;
    }
    for (int r = 0; r < 3; r++) {
      A.I[3 * (size_t)i + r] = make_float4(Im[4 * r], Im[4 * r + 1], Im[4 * r + 2], 0);
      A.invI[3 * (size_t)i + r] = make_float4(Iv[4 * r], Iv[4 * r + 1], Iv[4 * r + 2], 0);
    }
    A.mass[i] = (float)D->body_mass[i];
    A.invMass[i] = 1.0f / A.mass[i];
  }
  A.gtype.resize(ng); A.gbody.resize(ng); A.grank.resize(ng);
  A.gparam.resize(ng); A.gpos.resize(ng); A.gR.resize(3 * (size_t)ng);
  A.gcat.assign(ng, ~0u); A.gcol.assign(ng, ~0u);
  A.joints.resize(nj);
  A.wi.resize(D->nworlds);
  A.wp.resize(D->nworlds);
  int b0 = 0, g0 = 0, j0 = 0;
  for (int w = 0; w < D->nworlds; w++) {
    const dBatchWorldDesc &W = D->worlds[w];
    WorldInfo &wi = A.wi[w];
    wi.b0 = b0; wi.nb = W.nb; wi.g0 = g0; wi.ng = W.ng; wi.j0 = j0; wi.nj = W.nj;
    world_capacities(W.nb, W.ng, W.nj, cap_scale, wi.pcap, wi.kcap, wi.rcap);
    wi.adj0 = 0;
    WorldParams &P = A.wp[w];
    std::memset(&P, 0, sizeof(P));
    for (int k = 0; k < 3; k++) P.gravity[k] = (float)W.gravity[k];
    P.erp = (float)W.erp; P.cfm = (float)W.cfm; P.iterations = W.iterations; P.sor_w = (float)W.sor_w;
    P.max_vel = (float)W.max_vel; P.min_depth = (float)W.min_depth; P.dt = (float)W.dt;
    P.surf_mode = W.surf_mode; P.mu = (float)W.mu; P.mu2 = (float)W.mu2; P.bounce = (float)W.bounce;
    P.bounce_vel = (float)W.bounce_vel; P.soft_erp = (float)W.soft_erp; P.soft_cfm = (float)W.soft_cfm;
    P.max_contacts = W.max_contacts; P.skip_connected = W.skip_connected;
    P.hash_min = W.hash_minlevel; P.hash_max = W.hash_maxlevel;
    P.seed = (unsigned)W.seed;
    P.cell_scale = cell_scale;
    for (int g = 0; g < W.ng; g++) {
      int gg = g0 + g;
      A.gtype[gg] = D->geom_type[gg];
      A.gbody[gg] = D->geom_body[gg];
      A.grank[gg] = W.ng - 1 - g;   // geoms are prepended to the space list
      float prm[4];
      for (int k = 0; k < 4; k++) prm[k] = (float)D->geom_param[4 * (size_t)gg + k];
      if (A.gtype[gg] == GEOM_PLANE) {  // plane.cpp:49-76
        float l = prm[0] * prm[0] + prm[1] * prm[1] + prm[2] * prm[2];
        if (l > 0) {
          l = 1.f / std::sqrt(l);
          for (int k = 0; k < 4; k++) prm[k] *= l;
        } else {
          prm[0] = 1; prm[1] = prm[2] = prm[3] = 0;
        }
      }
      A.gparam[gg] = make_float4(prm[0], prm[1], prm[2], prm[3]);
      float gq[4], R[12];
      for (int k = 0; k < 4; k++) gq[k] = (float)D->geom_q[4 * (size_t)gg + k];
      h_quat_to_rot(gq, R);   // dGeomSetQuaternion (collision_kernel.cpp:599-613)
      A.gpos[gg] = make_float4((float)D->geom_pos[3 * (size_t)gg], (float)D->geom_pos[3 * (size_t)gg + 1],
                               (float)D->geom_pos[3 * (size_t)gg + 2], 0);
      for (int r = 0; r < 3; r++) A.gR[3 * (size_t)gg + r] = make_float4(R[4 * r], R[4 * r + 1], R[4 * r + 2], 0);
    }
    for (int j = 0; j < W.nj; j++) {
      int jg = j0 + j;
      build_joint(A.joints[jg], D, W, jg, b0, j, poses);
    }
    b0 += W.nb; g0 += W.ng; j0 += W.nj;
  }
  return A;
}

// per-body adjacency over permanent joints and per-body geom lists
static void derive_lists(BatchArrays &A) {
  int nb = (int)A.pos.size();
  std::vector<std::vector<int>> adj(nb), geoms(nb);
  for (size_t w = 0; w < A.wi.size(); w++) {
    const WorldInfo &W = A.wi[w];
    for (int g = 0; g < W.ng; g++)
      if (A.gbody[W.g0 + g] >= 0) geoms[W.b0 + A.gbody[W.g0 + g]].push_back(g);
    for (int j = 0; j < W.nj; j++) {
      const JointDev &J = A.joints[W.j0 + j];
      if (J.b1 >= 0 && J.b2 >= 0) {
        adj[W.b0 + J.b1].push_back(J.b2);
        adj[W.b0 + J.b2].push_back(J.b1);
      }
    }
  }
  A.adj_start.assign(nb + 1, 0);
  A.bgeom_start.assign(nb + 1, 0);
  A.adj_other.clear();
  A.bgeom.clear();
  for (int i = 0; i < nb; i++) {
    A.adj_start[i + 1] = A.adj_start[i] + (int)adj[i].size();
    A.adj_other.insert(A.adj_other.end(), adj[i].begin(), adj[i].end());
    // newest geom first (bodyAdd prepends, collision_kernel.h:173-177)
    std::sort(geoms[i].begin(), geoms[i].end(), std::greater<int>());
    A.bgeom_start[i + 1] = A.bgeom_start[i] + (int)geoms[i].size();
    A.bgeom.insert(A.bgeom.end(), geoms[i].begin(), geoms[i].end());
  }
}

BatchImpl *batch_create_arrays(BatchArrays &A, int order_mode) {
  if (A.wi.empty()) return nullptr;
  BatchImpl *b = new BatchImpl();
  b->nworlds = (int)A.wi.size();
  b->order_mode = order_mode == ODEMI_ORDER_REFERENCE ? ORDER_REFERENCE : ORDER_COLORED;
  int dev = 0;
  (void)hipGetDevice(&dev);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess) b->cu_count = prop.multiProcessorCount;
  init_lds(b, dev);
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    report_error("hipStreamCreate failed");
    delete b;
    return nullptr;
  }
  b->nb = (int)A.pos.size();
  b->ng = (int)A.gtype.size();
  b->h_gtype = A.gtype;
  b->h_joints = A.joints;
  b->nj = (int)A.joints.size();
  if (!A.wp.empty()) b->cell_scale = A.wp[0].cell_scale;
  derive_lists(A);
  long long scratch = 0, warm = 0;
  for (auto &wi : A.wi) {
    wi.scratch = scratch;
    scratch += scratch_layout(wi.nb, wi.ng, wi.nj, wi.pcap, wi.kcap, wi.rcap).total;
    wi.warm = warm;
    warm += warm_bytes(wi.kcap, wi.nj);
  }
  b->wi = A.wi;
  b->wp = A.wp;
  // largest worlds first for the persistent queue
  std::vector<int> order(b->nworlds);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    return b->wi[a].nb + b->wi[a].ng > b->wi[c].nb + b->wi[c].ng;
  });
  BatchBufs &d = b->d;
  int nb = b->nb, nj = b->nj;
  bool ok = upload(d.wi, b->wi) && upload(d.wp, b->wp) && upload(d.pos, A.pos) && upload(d.q, A.q) &&
            upload(d.lvel, A.lvel) && upload(d.avel, A.avel) && upload(d.facc, A.facc) && upload(d.tacc, A.tacc) &&
            upload(d.I, A.I) && upload(d.invI, A.invI) && upload(d.mass, A.mass) && upload(d.invMass, A.invMass) &&
            upload(d.gtype, A.gtype) && upload(d.gbody, A.gbody) && upload(d.gparam, A.gparam) &&
            upload(d.gpos, A.gpos) && upload(d.gR, A.gR) && upload(d.grank, A.grank) && upload(d.gcat, A.gcat) &&
            upload(d.gcol, A.gcol) && upload(d.joints, A.joints) && upload(d.adj_start, A.adj_start) &&
            upload(d.adj_other, A.adj_other) && upload(d.bgeom_start, A.bgeom_start) && upload(d.bgeom, A.bgeom) &&
            upload(d.world_order, order) && upload(d.world_rank, order) &&
            upload(d.cost_est, std::vector<float>(b->nworlds, 0.f)) && d.R.alloc(sizeof(float4) * 3 * (size_t)(nb ? nb : 1)) &&
            d.counters.alloc(sizeof(int) * CNT_N * (size_t)b->nworlds) &&
            d.residual.alloc(sizeof(float) * (size_t)b->nworlds) && d.work_counter.alloc(64 * WORK_QUEUES) &&
            d.scratch.alloc((size_t)scratch) && d.warm.alloc((size_t)(warm > 0 ? warm : 256)) &&
            d.warm_n.alloc(sizeof(int) * (size_t)b->nworlds) && d.ctrl.alloc(sizeof(float4) * (size_t)(nj ? nj : 1)) &&
            d.prof.alloc(sizeof(unsigned long long) * PH_N * (size_t)b->nworlds);
  if (!ok) {
    report_error("batch setup failed (%lld scratch bytes)", scratch);
    batch_destroy(b);
    return nullptr;
  }
  (void)hipMemset(d.counters.p, 0, d.counters.n);
  (void)hipMemset(d.residual.p, 0, d.residual.n);
  (void)hipMemset(d.prof.p, 0, d.prof.n);
  (void)hipMemset(d.warm_n.p, 0xff, d.warm_n.n);   // no warm-start records yet
  refresh_dev(b);
  if (launch_init_rot(b->dev, nb, b->stream) != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess) {
    report_error("batch setup: rotation init failed");
    batch_destroy(b);
    return nullptr;
  }
  return b;
}

// The descriptor checks the ODE setters make by assertion (sphere radius and
// box sides >= 0, sphere.cpp:52, box.cpp:51; only placeable geoms take a
// body, collision_kernel.cpp:510; dJointAttach refuses body1 == body2,
// ode.cpp:1352; dMassCheck wants mass > 0 and a positive definite inertia,
// mass.cpp:46-53; dBodySetQuaternion normalises, so a zero quaternion has no
// meaning), plus index ranges and finite poses: a descriptor that fails one
// is refused before any device work, with the reason reported.
static bool validate_desc(const dBatchDesc *D) {
  if (!D || D->nworlds <= 0 || !D->worlds) {
    report_error("dBatchCreate: no worlds");
    return false;
  }
  long long nb = 0, ng = 0, nj = 0;
  for (int w = 0; w < D->nworlds; w++) {
    const dBatchWorldDesc &W = D->worlds[w];
    if (W.nb < 0 || W.ng < 0 || W.nj < 0) {
      report_error("dBatchCreate: world %d has a negative body, geom or joint count", w);
      return false;
    }
    if (!(W.dt > 0 && std::isfinite(W.dt)) || W.iterations < 0 || W.hash_minlevel > W.hash_maxlevel) {
      report_error("dBatchCreate: world %d: step size, iteration count or hash levels out of range", w);
      return false;
    }
    nb += W.nb;
    ng += W.ng;
    nj += W.nj;
  }
  if (nb > INT32_MAX / 4 || ng > INT32_MAX / 8 || nj > INT32_MAX / 8) {
    report_error("dBatchCreate: batch too large");
    return false;
  }
  if ((nb && (!D->body_pos || !D->body_q || !D->body_lvel || !D->body_avel || !D->body_mass || !D->body_I)) ||
      (ng && (!D->geom_type || !D->geom_body || !D->geom_param || !D->geom_pos || !D->geom_q)) ||
      (nj && (!D->joint_type || !D->joint_b1 || !D->joint_b2 || !D->joint_anchor || !D->joint_axis1 ||
              !D->joint_axis2 || !D->joint_param))) {
    report_error("dBatchCreate: a body, geom or joint array is missing");
    return false;
  }
  long long b0 = 0, g0 = 0, j0 = 0;
  for (int w = 0; w < D->nworlds; w++) {
    const dBatchWorldDesc &W = D->worlds[w];
    for (int i = 0; i < W.nb; i++) {
      const long long bi = b0 + i;
      const double m = D->body_mass[bi];
      if (!(m > 0 && std::isfinite(m))) {
        report_error("dBatchCreate: world %d body %d: mass must be positive and finite", w, i);
        return false;
      }
      // inertia positive definite (Cholesky of the symmetric part)
      const double *I = D->body_I + 9 * bi;
      const double a = I[0], b = 0.5 * (I[1] + I[3]), c = 0.5 * (I[2] + I[6]);
      const double e = I[4], f = 0.5 * (I[5] + I[7]), k = I[8];
      bool pd = a > 0;
      if (pd) {
        const double l11 = std::sqrt(a), l21 = b / l11, l31 = c / l11;
        const double d2 = e - l21 * l21;
        pd = d2 > 0;
        if (pd) {
          const double l22 = std::sqrt(d2), l32 = (f - l31 * l21) / l22;
          pd = k - l31 * l31 - l32 * l32 > 0;
        }
      }
      if (!pd) {
        report_error("dBatchCreate: world %d body %d: inertia must be positive definite", w, i);
        return false;
      }
      const double *q = D->body_q + 4 * bi;
      double qq = 0;
      bool fin = true;
      for (int t = 0; t < 4; t++) { qq += q[t] * q[t]; fin &= std::isfinite(q[t]); }
      for (int t = 0; t < 3; t++)
        fin &= std::isfinite(D->body_pos[3 * bi + t]) && std::isfinite(D->body_lvel[3 * bi + t]) &&
               std::isfinite(D->body_avel[3 * bi + t]);
      if (!fin || !(qq > 0)) {
        report_error("dBatchCreate: world %d body %d: pose or velocity not finite, or zero quaternion", w, i);
        return false;
      }
    }
    for (int g = 0; g < W.ng; g++) {
      const long long gg = g0 + g;
      const int t = D->geom_type[gg], body = D->geom_body[gg];
      const double *p = D->geom_param + 4 * gg;
      if (t == GEOM_RAY) {
        report_error("dBatchCreate: world %d geom %d: a ray (its contacts are reported through dSpaceCollide / "
                     "dCollide; every contact of a batch world becomes a joint)", w, g);
        return false;
      }
      if (t != GEOM_SPHERE && t != GEOM_BOX && t != GEOM_PLANE && t != GEOM_CAPSULE && t != GEOM_CYLINDER) {
        report_error("dBatchCreate: world %d geom %d: unknown type %d", w, g, t);
        return false;
      }
      if (body < -1 || body >= W.nb || (t == GEOM_PLANE && body != -1)) {
        report_error("dBatchCreate: world %d geom %d: body %d out of range (planes take none)", w, g, body);
        return false;
      }
      const int nsz = t == GEOM_SPHERE ? 1 : (t == GEOM_BOX ? 3 : ((t == GEOM_CAPSULE || t == GEOM_CYLINDER) ? 2 : 0));
      for (int k = 0; k < nsz; k++)
        if (!(p[k] >= 0 && std::isfinite(p[k]))) {
          report_error("dBatchCreate: world %d geom %d: size must be >= 0 and finite", w, g);
          return false;
        }
    }
    for (int j = 0; j < W.nj; j++) {
      const long long jg = j0 + j;
      const int t = D->joint_type[jg], a = D->joint_b1[jg], b = D->joint_b2[jg];
      if (t != JT_BALL && t != JT_HINGE && t != JT_HINGE2) {
        report_error("dBatchCreate: world %d joint %d: unknown type %d", w, j, t);
        return false;
      }
      if (a < -1 || a >= W.nb || b < -1 || b >= W.nb || (a >= 0 && a == b)) {
        report_error("dBatchCreate: world %d joint %d: bodies %d, %d out of range or equal", w, j, a, b);
        return false;
      }
    }
    b0 += W.nb;
    g0 += W.ng;
    j0 += W.nj;
  }
  return true;
}

BatchImpl *batch_create(const dBatchDesc *D, int order_mode, double cap_scale) {
  if (!validate_desc(D)) return nullptr;
  BatchArrays A = arrays_from_desc(D, cap_scale, 1.0f);
  return batch_create_arrays(A, order_mode);
}

void batch_destroy(BatchImpl *b) {
  if (!b) return;
  destroy_graph(b);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

static int persistent_grid(BatchImpl *b) {
  if (b->grid > 0) return b->grid;
  return 0;
}

// one step of every world: one fused kernel, or collide and step kernels, or
// (coloured order at a tolerance) collide + set-up, the sweeps, integration
static hipError_t launch_step(BatchImpl *b, int first_world, int nworlds, int grid) {
  if (b->cost_order && grid > 0 && first_world == 0 && nworlds == b->nworlds && b->nworlds >= 2 &&
      b->nworlds <= 8192) {
    // the queue order by estimated solver work, rewritten every step
    hipError_t e = launch_order_worlds(b->dev, b->d.world_rank.as<int>(), b->d.cost_est.as<float>(), b->stream);
    if (e != hipSuccess) return e;
    b->order_rewritten = true;
  }
  if (b->tol_on && b->sor_split && b->order_mode == ORDER_COLORED) {
    hipError_t e = launch_world_kernel(KM_FUSED_PRE, b->dev, first_world, nworlds, grid, b->wg, 0, b->stream, 0);
    if (e != hipSuccess) return e;
    DevBatch d2 = b->dev;
    d2.lds_bytes = b->sor_lds;
    const int g2 = grid > 0 ? b->sor_grid : 0;
    e = launch_world_kernel(KM_SOR, d2, first_world, nworlds, g2, b->sor_wg, 0, b->stream, 1);
    if (e != hipSuccess) return e;
    DevBatch d3 = b->dev;
    d3.lds_bytes = 0;
    return launch_world_kernel(KM_POST, d3, first_world, nworlds, grid, b->wg, 0, b->stream, 2);
  }
  if (!b->split)
    return launch_world_kernel(b->tol_on ? KM_FUSED_TOL : KM_FUSED, b->dev, first_world, nworlds, grid, b->wg, 0,
                               b->stream, 0);
  hipError_t e = launch_world_kernel(KM_BATCH_COLLIDE, b->dev, first_world, nworlds, grid, b->wg, 0, b->stream, 0);
  if (e != hipSuccess) return e;
  return launch_world_kernel(b->tol_on ? KM_BATCH_STEP_TOL : KM_BATCH_STEP, b->dev, first_world, nworlds, grid,
                             b->wg, 0, b->stream, 1);
}

static int build_graph(BatchImpl *b) {
  destroy_graph(b);
  refresh_dev(b);
  OD_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = launch_step(b, 0, b->nworlds, persistent_grid(b));
  hipGraph_t g = nullptr;
  hipError_t e2 = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess || e2 != hipSuccess || !g) {
    report_error("graph capture failed: %s / %s", hipGetErrorString(e), hipGetErrorString(e2));
    if (g) (void)hipGraphDestroy(g);
    return 0;
  }
  b->graph = g;
  OD_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
  return 1;
}

int batch_step(BatchImpl *b, int nsteps, bool sync) {
  if (nsteps <= 0) return 1;
  if (b->tape.recording) {
    for (int s = 0; s < nsteps; s++)
      if (!tape_step(b, s == nsteps - 1 && sync)) return 0;
    return 1;
  }
  if (!b->gexec && !build_graph(b)) return 0;
  for (int s = 0; s < nsteps; s++) OD_HIP(hipGraphLaunch(b->gexec, b->stream));
  if (sync) OD_HIP(hipStreamSynchronize(b->stream));
  return 1;
}

// nsteps graph replays bracketed by HIP events on the batch stream; returns
// the elapsed device time in ms.
int batch_step_timed(BatchImpl *b, int nsteps, float *ms) {
  if (!b->gexec && !build_graph(b)) return 0;
  hipEvent_t e0, e1;
  OD_HIP(hipEventCreate(&e0));
  OD_HIP(hipEventCreate(&e1));
  OD_HIP(hipEventRecord(e0, b->stream));
  for (int s = 0; s < nsteps; s++) OD_HIP(hipGraphLaunch(b->gexec, b->stream));
  OD_HIP(hipEventRecord(e1, b->stream));
  OD_HIP(hipEventSynchronize(e1));
  float t = 0;
  OD_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (ms) *ms = t;
  return 1;
}

int batch_step_plain(BatchImpl *b, int nsteps, bool sync) {
  refresh_dev(b);
  for (int s = 0; s < nsteps; s++)
    for (int w = 0; w < b->nworlds; w++)
      OD_HIP(launch_step(b, w, 1, 0));
  if (sync) OD_HIP(hipStreamSynchronize(b->stream));
  return 1;
}

// one fused step launched directly (used by the tape recorder and autotuner)
int batch_launch_once(BatchImpl *b) {
  refresh_dev(b);
  OD_HIP(launch_step(b, 0, b->nworlds, persistent_grid(b)));
  return 1;
}

int batch_get_state(BatchImpl *b, float *out) {
  OD_HIP(hipStreamSynchronize(b->stream));
  std::vector<float4> p(b->nb), q(b->nb), v(b->nb), w(b->nb);
  if (b->nb) {
    OD_HIP(hipMemcpy(p.data(), b->d.pos.p, sizeof(float4) * b->nb, hipMemcpyDeviceToHost));
    OD_HIP(hipMemcpy(q.data(), b->d.q.p, sizeof(float4) * b->nb, hipMemcpyDeviceToHost));
    OD_HIP(hipMemcpy(v.data(), b->d.lvel.p, sizeof(float4) * b->nb, hipMemcpyDeviceToHost));
    OD_HIP(hipMemcpy(w.data(), b->d.avel.p, sizeof(float4) * b->nb, hipMemcpyDeviceToHost));
  }
  for (int i = 0; i < b->nb; i++) {
    float *r = out + 13 * (size_t)i;
    r[0] = p[i].x; r[1] = p[i].y; r[2] = p[i].z;
    r[3] = q[i].x; r[4] = q[i].y; r[5] = q[i].z; r[6] = q[i].w;
    r[7] = v[i].x; r[8] = v[i].y; r[9] = v[i].z;
    r[10] = w[i].x; r[11] = w[i].y; r[12] = w[i].z;
  }
  return 1;
}

template <class F>
static int patch_params(BatchImpl *b, F f);

int batch_set_state(BatchImpl *b, const float *in) {
  OD_HIP(hipStreamSynchronize(b->stream));
  std::vector<float4> p(b->nb), q(b->nb), v(b->nb), w(b->nb);
  for (int i = 0; i < b->nb; i++) {
    const float *r = in + 13 * (size_t)i;
    p[i] = make_float4(r[0], r[1], r[2], 0);
    q[i] = make_float4(r[3], r[4], r[5], r[6]);
    v[i] = make_float4(r[7], r[8], r[9], 0);
    w[i] = make_float4(r[10], r[11], r[12], 0);
  }
  if (b->nb) {
    OD_HIP(hipMemcpy(b->d.pos.p, p.data(), sizeof(float4) * b->nb, hipMemcpyHostToDevice));
    OD_HIP(hipMemcpy(b->d.q.p, q.data(), sizeof(float4) * b->nb, hipMemcpyHostToDevice));
    OD_HIP(hipMemcpy(b->d.lvel.p, v.data(), sizeof(float4) * b->nb, hipMemcpyHostToDevice));
    OD_HIP(hipMemcpy(b->d.avel.p, w.data(), sizeof(float4) * b->nb, hipMemcpyHostToDevice));
  }
  // the previous impulses belong to the old state: the next solve starts from zero
  OD_HIP(hipMemset(b->d.warm_n.p, 0xff, b->d.warm_n.n));
  refresh_dev(b);
  if (launch_init_rot(b->dev, b->nb, b->stream) != hipSuccess) return 0;
  OD_HIP(hipStreamSynchronize(b->stream));
  // a new state revives frozen worlds
  return patch_params(b, [](WorldParams &p) { p.dead = 0; });
}

int batch_counters(BatchImpl *b, long long *out5) {
  OD_HIP(hipStreamSynchronize(b->stream));
  std::vector<int> c((size_t)CNT_N * b->nworlds);
  OD_HIP(hipMemcpy(c.data(), b->d.counters.p, sizeof(int) * c.size(), hipMemcpyDeviceToHost));
  for (int k = 0; k < 5; k++) out5[k] = 0;
  for (int w = 0; w < b->nworlds; w++) {
    out5[0] += c[CNT_N * w + CNT_PAIRS];
    out5[1] += c[CNT_N * w + CNT_CONTACTS];
    out5[2] += c[CNT_N * w + CNT_ROWS];
    out5[3] += c[CNT_N * w + CNT_GROUPS];
    out5[4] |= c[CNT_N * w + CNT_OVERFLOW];
  }
  return 1;
}

int batch_set_controls(BatchImpl *b, const double *ctrl) {
  if (!ctrl) {
    if (b->ctrl_set) { b->ctrl_set = false; destroy_graph(b); }
    return 1;
  }
  std::vector<float4> c(b->nj ? b->nj : 1);
  for (int j = 0; j < b->nj; j++)
    c[j] = make_float4((float)ctrl[4 * (size_t)j], (float)ctrl[4 * (size_t)j + 1], (float)ctrl[4 * (size_t)j + 2],
                       (float)ctrl[4 * (size_t)j + 3]);
  OD_HIP(hipMemcpyAsync(b->d.ctrl.p, c.data(), sizeof(float4) * c.size(), hipMemcpyHostToDevice, b->stream));
  OD_HIP(hipStreamSynchronize(b->stream));
  if (!b->ctrl_set) { b->ctrl_set = true; destroy_graph(b); }
  return 1;
}

// Diagnostics of the last step of world w: the candidate pairs in report
// order, and the contacts with the impulses of their rows.
int batch_get_pairs(BatchImpl *b, int w, int maxn, int *out) {
  if (w < 0 || w >= b->nworlds) return -1;
  OD_HIP(hipStreamSynchronize(b->stream));
  const WorldInfo &W = b->wi[w];
  ScratchLayout L = scratch_layout(W.nb, W.ng, W.nj, W.pcap, W.kcap, W.rcap);
  int cnt[CNT_N];
  OD_HIP(hipMemcpy(cnt, b->d.counters.as<int>() + CNT_N * w, sizeof(cnt), hipMemcpyDeviceToHost));
  int n = cnt[CNT_PAIRS] < maxn ? cnt[CNT_PAIRS] : maxn;
  if (n > 0) {
    std::vector<PairRec> pr(n);
    OD_HIP(hipMemcpy(pr.data(), b->d.scratch.as<char>() + W.scratch + L.pairs, sizeof(PairRec) * n,
                     hipMemcpyDeviceToHost));
    for (int i = 0; i < n; i++) { out[2 * i] = pr[i].g1; out[2 * i + 1] = pr[i].g2; }
  }
  return cnt[CNT_PAIRS];
}

int batch_get_contacts(BatchImpl *b, int w, int maxn, float *out) {
  if (w < 0 || w >= b->nworlds) return -1;
  OD_HIP(hipStreamSynchronize(b->stream));
  const WorldInfo &W = b->wi[w];
  ScratchLayout L = scratch_layout(W.nb, W.ng, W.nj, W.pcap, W.kcap, W.rcap);
  int cnt[CNT_N];
  OD_HIP(hipMemcpy(cnt, b->d.counters.as<int>() + CNT_N * w, sizeof(cnt), hipMemcpyDeviceToHost));
  int nk = cnt[CNT_CONTACTS] < maxn ? cnt[CNT_CONTACTS] : maxn;
  if (nk <= 0) return cnt[CNT_CONTACTS];
  const char *S = b->d.scratch.as<char>() + W.scratch;
  std::vector<ContactRec> cr(nk);
  std::vector<Cons> cons(nk);
  OD_HIP(hipMemcpy(cr.data(), S + L.contacts, sizeof(ContactRec) * nk, hipMemcpyDeviceToHost));
  OD_HIP(hipMemcpy(cons.data(), S + L.cons, sizeof(Cons) * nk, hipMemcpyDeviceToHost));
  int nrows = cnt[CNT_ROWS];
  std::vector<Row> rows(nrows > 0 ? nrows : 1);
  if (nrows > 0) OD_HIP(hipMemcpy(rows.data(), S + L.rows, sizeof(Row) * nrows, hipMemcpyDeviceToHost));
  for (int i = 0; i < nk; i++) {
    float *o = out + 12 * (size_t)i;
    for (int k = 0; k < 3; k++) o[k] = cr[i].pos[k];
    for (int k = 0; k < 3; k++) o[3 + k] = cr[i].normal[k];
    o[6] = cr[i].depth;
    o[7] = (float)cr[i].g1;
    o[8] = (float)cr[i].g2;
    for (int k = 0; k < 3; k++) {
      int r = cons[i].row0 + k;
      o[9 + k] = (k < cons[i].m && r < nrows) ? rows[r].lambda : 0.f;
    }
  }
  return cnt[CNT_CONTACTS];
}

int snapshot(BatchImpl *b, Snapshot &s, bool save) {
  struct P { DevBuf *src, *dst; };
  P items[] = {{&b->d.pos, &s.pos}, {&b->d.q, &s.q}, {&b->d.lvel, &s.lvel}, {&b->d.avel, &s.avel},
               {&b->d.facc, &s.facc}, {&b->d.tacc, &s.tacc}, {&b->d.R, &s.R}, {&b->d.grank, &s.grank},
               {&b->d.joints, &s.joints}, {&b->d.wp, &s.wp}, {&b->d.warm, &s.warm},
               {&b->d.warm_n, &s.warm_n}};
  for (auto &it : items) {
    if (save) {
      if (!it.dst->alloc(it.src->n)) return 0;
      OD_HIP(hipMemcpyAsync(it.dst->p, it.src->p, it.src->n, hipMemcpyDeviceToDevice, b->stream));
    } else {
      OD_HIP(hipMemcpyAsync(it.src->p, it.dst->p, it.src->n, hipMemcpyDeviceToDevice, b->stream));
    }
  }
  OD_HIP(hipStreamSynchronize(b->stream));
  return 1;
}

// applies f to the world parameters on the host and on the device (the
// device copy holds live RNG seeds, so it is read back and patched)
template <class F>
static int patch_params(BatchImpl *b, F f) {
  for (auto &p : b->wp) f(p);
  std::vector<WorldParams> cur(b->nworlds);
  OD_HIP(hipStreamSynchronize(b->stream));
  OD_HIP(hipMemcpy(cur.data(), b->d.wp.p, sizeof(WorldParams) * b->nworlds, hipMemcpyDeviceToHost));
  for (auto &p : cur) f(p);
  OD_HIP(hipMemcpy(b->d.wp.p, cur.data(), sizeof(WorldParams) * b->nworlds, hipMemcpyHostToDevice));
  return 1;
}

int batch_set_tolerance(BatchImpl *b, float tol, int max_iterations, int check_every, int accel) {
  if (!(tol >= 0.f) || (tol > 0.f && (max_iterations <= 0 || check_every <= 0))) return 0;
  if (!patch_params(b, [&](WorldParams &p) {
        p.tol = tol;
        p.max_iterations = max_iterations;
        p.check_every = check_every;
        p.accel = accel ? 1 : 0;
      }))
    return 0;
  b->tol_on = tol > 0.f;
  destroy_graph(b);
  return 1;
}

int batch_set_warm_start(BatchImpl *b, float warm) {
  if (!(warm >= 0.f && warm <= 1.f)) return 0;
  if (!patch_params(b, [&](WorldParams &p) { p.warm = warm; })) return 0;
  OD_HIP(hipMemset(b->d.warm_n.p, 0xff, b->d.warm_n.n));
  return 1;
}

static int set_cell_scale(BatchImpl *b, float s) {
  b->cell_scale = s;
  for (auto &p : b->wp) p.cell_scale = s;
  // only the cell_scale field changes; the seeds on the device stay live
  std::vector<WorldParams> cur(b->nworlds);
  OD_HIP(hipMemcpy(cur.data(), b->d.wp.p, sizeof(WorldParams) * b->nworlds, hipMemcpyDeviceToHost));
  for (auto &p : cur) p.cell_scale = s;
  OD_HIP(hipMemcpy(b->d.wp.p, cur.data(), sizeof(WorldParams) * b->nworlds, hipMemcpyHostToDevice));
  return 1;
}

int batch_set_launch(BatchImpl *b, int wg, float cell_scale, int grid) {
  if (wg != 64 && wg != 128 && wg != 192 && wg != 256) return 0;
  if (!(cell_scale >= 1.0f && cell_scale <= 8.0f)) return 0;
  b->wg = wg;
  b->grid = grid;
  if (!set_cell_scale(b, cell_scale)) return 0;
  destroy_graph(b);
  return 1;
}

int batch_set_lds(BatchImpl *b, int bytes) {
  if (bytes < 0 || bytes > b->lds_max) return 0;
  b->lds = bytes & ~255;
  destroy_graph(b);
  return 1;
}

int batch_set_sor_launch(BatchImpl *b, int on, int wg, int lds, int grid) {
  if (on) {
    // switched off: a timing run of a 4096-world batch with the split sweeps
    // did not finish; the cause is not found yet
    report_error("dBatchSetSorLaunch: the split sweep kernel is switched off");
    return 0;
  }
  if (wg != 64 && wg != 128 && wg != 192 && wg != 256) return 0;
  if (lds < 0 || lds > b->lds_max || grid < 0) return 0;
  b->sor_split = on != 0;
  b->sor_wg = wg;
  b->sor_lds = lds & ~255;
  b->sor_grid = grid;
  destroy_graph(b);
  return 1;
}

int batch_set_profiling(BatchImpl *b, bool on) {
  OD_HIP(hipStreamSynchronize(b->stream));
  OD_HIP(hipMemset(b->d.prof.p, 0, b->d.prof.n));
  if (b->prof_on != on) {
    b->prof_on = on;
    destroy_graph(b);
  }
  return 1;
}

// per-phase workgroup time summed over worlds since the last call (ms)
int batch_phase_times(BatchImpl *b, double *ms) {
  OD_HIP(hipStreamSynchronize(b->stream));
  std::vector<unsigned long long> t((size_t)PH_N * b->nworlds);
  OD_HIP(hipMemcpy(t.data(), b->d.prof.p, sizeof(unsigned long long) * t.size(), hipMemcpyDeviceToHost));
  OD_HIP(hipMemset(b->d.prof.p, 0, b->d.prof.n));
  for (int k = 0; k < PH_N; k++) ms[k] = 0;
  for (int w = 0; w < b->nworlds; w++)
    for (int k = 0; k < PH_N; k++) ms[k] += (double)t[(size_t)PH_N * w + k];
  for (int k = 0; k < PH_N; k++) ms[k] /= (double)b->wall_khz;
  return 1;
}

// Launch-shape search on the live batch (state saved and restored around
// the trials): broadphase cell scale first, then workgroup size x LDS pool x
// grid (one workgroup per world, or a persistent grid filling the CUs at the
// kernel's occupancy for that shape).
int batch_autotune(BatchImpl *b, int trial_steps, int flags) {
  if (trial_steps <= 0) trial_steps = 3;
  Snapshot snap;
  if (!snapshot(b, snap, true)) return 0;
  auto trial = [&](double &t) -> int {
    if (!batch_step(b, 1, true)) return 0;   // warm-up and graph build
    auto t0 = std::chrono::steady_clock::now();
    if (!batch_step(b, trial_steps, true)) return 0;
    t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return snapshot(b, snap, false);
  };
  double best = 1e30;
  b->sor_split = false;
  int bw = b->wg, bg = b->grid, bl = b->lds;
  float bs = b->cell_scale;
  const float scales[] = {1.0f, 1.5f, 2.0f};
  if (!(flags & ODEMI_TUNE_KEEP_CELL))
    for (float sc : scales) {
      double t;
      if (!batch_set_launch(b, bw, sc, 0) || !trial(t)) continue;
      t *= 1.02;   // grid 0, as below
      if (t < best) { best = t; bs = sc; bg = 0; }
    }
  const int wgs[] = {64, 128, 256};
  const int lds_div[] = {1, 2, 3, 4, 8};   // pool = LDS per CU / k
  int st = world_kernel_static_lds();
  // the fused kernel only: the two-kernel step (split) is switched off
  b->split = false;
  for (int wg : wgs)
    for (int k : lds_div) {
      int lds = (b->lds_cu / k - st - 256) & ~255;
      if (lds < 0) lds = 0;
      if (lds > b->lds_max) lds = b->lds_max;
      int occ = world_kernel_occupancy(KM_FUSED, wg, lds);
      if (occ <= 0) continue;
      const int grids[] = {0, occ * b->cu_count};
      for (int grid : grids) {
        double t;
        if (!batch_set_launch(b, wg, bs, grid) || !batch_set_lds(b, lds) || !trial(t)) continue;
        // one launch per world (grid 0) must beat the persistent, size-ordered
        // queue by 2 %: a single trial's noise is about that, and the queue's
        // largest-first order is what bounds the tail of uneven batches
        if (grid == 0) t *= 1.02;
        if (t < best) { best = t; bw = wg; bg = grid; bl = lds; }
      }
    }
  batch_set_launch(b, bw, bs, bg);
  batch_set_lds(b, bl);
  // coloured order at a tolerance: the sweeps in their own kernel, with
  // their own workgroup size, LDS pool and grid
  bool bsor = false;
  int sw = b->sor_wg, sl = b->sor_lds, sg = b->sor_grid;
  if (false && b->tol_on && b->order_mode == ORDER_COLORED && !(flags & ODEMI_TUNE_NO_SOR_SPLIT)) {
    const int sdiv[] = {2, 3, 4, 6, 8};
    for (int wg : wgs)
      for (int k : sdiv) {
        int lds = (b->lds_cu / k - st - 256) & ~255;
        if (lds < 0) lds = 0;
        if (lds > b->lds_max) lds = b->lds_max;
        int occ = world_kernel_occupancy(KM_SOR, wg, lds);
        if (occ <= 0) continue;
        const int grids[] = {0, occ * b->cu_count};
        for (int grid : grids) {
          double t;
          if (!batch_set_sor_launch(b, 1, wg, lds, grid) || !trial(t)) continue;
          if (t < best) { best = t; bsor = true; sw = wg; sl = lds; sg = grid; }
        }
      }
  }
  batch_set_sor_launch(b, bsor ? 1 : 0, sw, sl, sg);
  if (!snapshot(b, snap, false)) return 0;
  return bw;
}

}  // namespace odemi

using namespace odemi;

struct dxBatch {
  BatchImpl *impl;
};

extern "C" {

dBatchID dBatchCreate(const dBatchDesc *desc, int order_mode) {
  const char *env = std::getenv("ODEMI_CAPACITY_SCALE");
  double scale = env ? std::atof(env) : 1.0;
  if (!(scale > 0)) scale = 1.0;
  BatchImpl *impl = batch_create(desc, order_mode, scale);
  if (!impl) return nullptr;
  dxBatch *b = new dxBatch;
  b->impl = impl;
  return b;
}

void dBatchDestroy(dBatchID b) {
  if (!b) return;
  batch_destroy(b->impl);
  delete b;
}

int dBatchStep(dBatchID b, int nsteps) { return b ? batch_step(b->impl, nsteps, true) : 0; }
int dBatchStepTimed(dBatchID b, int nsteps, float *ms) { return b ? batch_step_timed(b->impl, nsteps, ms) : 0; }
int dBatchStepPlain(dBatchID b, int nsteps) { return b ? batch_step_plain(b->impl, nsteps, true) : 0; }
int dBatchSetControls(dBatchID b, const double *ctrl) { return b ? batch_set_controls(b->impl, ctrl) : 0; }
int dBatchGetState(dBatchID b, float *out) { return (b && out) ? batch_get_state(b->impl, out) : 0; }
int dBatchSetState(dBatchID b, const float *in) { return (b && in) ? batch_set_state(b->impl, in) : 0; }
int dBatchGetCounters(dBatchID b, long long *out5) { return (b && out5) ? batch_counters(b->impl, out5) : 0; }

int dBatchSetResidualTracking(dBatchID b, int on) {
  if (!b) return 0;
  b->impl->dev.track_residual = on ? 1 : 0;
  destroy_graph(b->impl);
  return 1;
}

int dBatchGetResidual(dBatchID b, float *out) {
  if (!b || !out) return 0;
  BatchImpl *i = b->impl;
  if (hipStreamSynchronize(i->stream) != hipSuccess) return 0;
  return hipMemcpy(out, i->d.residual.p, sizeof(float) * i->nworlds, hipMemcpyDeviceToHost) == hipSuccess;
}

int dBatchAutotune(dBatchID b, int trial_steps) { return b ? batch_autotune(b->impl, trial_steps, 0) : 0; }
int dBatchAutotuneEx(dBatchID b, int trial_steps, int flags) {
  return b ? batch_autotune(b->impl, trial_steps, flags) : 0;
}

int dBatchGetLaunchConfig(dBatchID b, int *wg, float *cell_scale) {
  if (!b) return 0;
  if (wg) *wg = b->impl->wg;
  if (cell_scale) *cell_scale = b->impl->cell_scale;
  return b->impl->grid;
}

int dBatchSetLaunchConfig(dBatchID b, int wg, float cell_scale) {
  return b ? batch_set_launch(b->impl, wg, cell_scale, b->impl->grid) : 0;
}

void *dBatchDeviceState(dBatchID b) { return b ? b->impl->d.pos.p : nullptr; }
void *dBatchStream(dBatchID b) { return b ? (void *)b->impl->stream : nullptr; }
int dBatchNumBodies(dBatchID b) { return b ? b->impl->nb : 0; }

int dBatchGetLds(dBatchID b) { return b ? b->impl->lds : -1; }
int dBatchSetSolverTolerance(dBatchID b, float tol, int max_iterations, int check_every) {
  return b ? batch_set_tolerance(b->impl, tol, max_iterations, check_every, 0) : 0;
}
int dBatchSetSolverToleranceAccel(dBatchID b, float tol, int max_iterations, int check_every, int accel) {
  return b ? batch_set_tolerance(b->impl, tol, max_iterations, check_every, accel) : 0;
}
int dBatchSetWarmStart(dBatchID b, float warm) { return b ? batch_set_warm_start(b->impl, warm) : 0; }
int dBatchSetExactLCP(dBatchID b, int max_island_rows) {
  if (!b || max_island_rows < 0) return 0;
  BatchImpl *i = b->impl;
  if (max_island_rows > 0 && i->order_mode != ORDER_REFERENCE) {
    report_error("dBatchSetExactLCP: the exact solve runs per island, in the reference order");
    return 0;
  }
  if (hipStreamSynchronize(i->stream) != hipSuccess) return 0;
  if (max_island_rows > 0) {
    const size_t bytes = sizeof(float) * (size_t)lcp_ws_floats(max_island_rows) * (size_t)i->nworlds;
    if (!i->lcp_ws.alloc(bytes)) {
      report_error("dBatchSetExactLCP: device allocation of %zu bytes failed", bytes);
      return 0;
    }
  } else {
    i->lcp_ws.release();
  }
  i->lcp_rows = max_island_rows;
  const int on = max_island_rows > 0 ? 1 : 0;
  if (!patch_params(i, [&](WorldParams &p) { p.exact = on; })) return 0;
  destroy_graph(i);
  refresh_dev(i);
  return 1;
}
int dBatchGetWorldCounters(dBatchID b, int *out) {
  if (!b || !out) return 0;
  BatchImpl *i = b->impl;
  if (hipStreamSynchronize(i->stream) != hipSuccess) return 0;
  return hipMemcpy(out, i->d.counters.p, sizeof(int) * CNT_N * (size_t)i->nworlds, hipMemcpyDeviceToHost) ==
         hipSuccess;
}
int dBatchResetSolveStats(dBatchID b) {
  if (!b) return 0;
  BatchImpl *i = b->impl;
  if (hipStreamSynchronize(i->stream) != hipSuccess) return 0;
  std::vector<int> c((size_t)CNT_N * i->nworlds);
  OD_HIP(hipMemcpy(c.data(), i->d.counters.p, sizeof(int) * c.size(), hipMemcpyDeviceToHost));
  for (int w = 0; w < i->nworlds; w++) c[(size_t)CNT_N * w + CNT_UNCONV] = c[(size_t)CNT_N * w + CNT_RESMAX] = 0;
  OD_HIP(hipMemcpy(i->d.counters.p, c.data(), sizeof(int) * c.size(), hipMemcpyHostToDevice));
  return 1;
}
int dBatchSetPersistentGrid(dBatchID b, int grid) {
  if (!b || grid < 0) return 0;
  b->impl->grid = grid;
  destroy_graph(b->impl);
  return 1;
}
int dBatchGetSplit(dBatchID b) { return b ? (b->impl->split ? 1 : 0) : -1; }
int dBatchSetQueueOrder(dBatchID b, int by_cost) {
  if (!b) return 0;
  BatchImpl *m = b->impl;
  m->cost_order = by_cost != 0;
  destroy_graph(m);
  if (!m->cost_order && m->order_rewritten) {
    // back to the static (largest first) order
    OD_HIP(hipMemcpyAsync(m->d.world_order.p, m->d.world_rank.p, sizeof(int) * m->nworlds, hipMemcpyDeviceToDevice,
                          m->stream));
    OD_HIP(hipStreamSynchronize(m->stream));
    m->order_rewritten = false;
  }
  return 1;
}
int dBatchGetQueueOrder(dBatchID b) { return b ? (b->impl->cost_order ? 1 : 0) : -1; }
int dBatchSetSorLaunch(dBatchID b, int on, int wg, int lds, int grid) {
  return b ? batch_set_sor_launch(b->impl, on, wg, lds, grid) : 0;
}
int dBatchGetSorLaunch(dBatchID b, int *on, int *wg, int *lds, int *grid) {
  if (!b) return 0;
  if (on) *on = b->impl->sor_split ? 1 : 0;
  if (wg) *wg = b->impl->sor_wg;
  if (lds) *lds = b->impl->sor_lds;
  if (grid) *grid = b->impl->sor_grid;
  return 1;
}
int dBatchSetSplit(dBatchID b, int on) {
  if (!b) return 0;
  if (on) {
    // switched off: a 4096-world timing run with collision and quickstep in
    // two persistent kernels did not finish; the cause is not found yet
    report_error("dBatchSetSplit: the two-kernel step is switched off");
    return 0;
  }
  b->impl->split = false;
  destroy_graph(b->impl);
  return 1;
}
int dBatchSetLds(dBatchID b, int bytes) { return b ? batch_set_lds(b->impl, bytes) : 0; }
int dBatchSetProfiling(dBatchID b, int on) { return b ? batch_set_profiling(b->impl, on != 0) : 0; }
int dBatchGetPhaseTimes(dBatchID b, double *ms) { return (b && ms) ? batch_phase_times(b->impl, ms) : 0; }
int dBatchGetWorldPhaseTimes(dBatchID b, double *ms) {
  if (!b || !ms) return 0;
  BatchImpl *i = b->impl;
  OD_HIP(hipStreamSynchronize(i->stream));
  std::vector<unsigned long long> t((size_t)PH_N * i->nworlds);
  OD_HIP(hipMemcpy(t.data(), i->d.prof.p, sizeof(unsigned long long) * t.size(), hipMemcpyDeviceToHost));
  OD_HIP(hipMemset(i->d.prof.p, 0, i->d.prof.n));
  for (size_t k = 0; k < t.size(); k++) ms[k] = (double)t[k] / (double)i->wall_khz;
  return 1;
}

int dBatchGetPairs(dBatchID b, int world, int maxn, int *out) {
  return b ? batch_get_pairs(b->impl, world, maxn, out) : -1;
}
int dBatchGetContacts(dBatchID b, int world, int maxn, float *out) {
  return b ? batch_get_contacts(b->impl, world, maxn, out) : -1;
}

int dBatchTapeBegin(dBatchID b, int max_steps) { return b ? tape_begin(b->impl, max_steps) : 0; }
int dBatchTapeSetOptions(dBatchID b, int checkpoint_interval, float act_tol, float lin_tol, int lin_maxit) {
  if (!b || checkpoint_interval <= 0 || !(act_tol >= 0.f) || !(lin_tol > 0.f) || lin_maxit <= 0) return 0;
  Tape &t = b->impl->tape;
  if (t.recording) return 0;
  t.interval = checkpoint_interval;
  t.act_tol = act_tol;
  t.lin_tol = lin_tol;
  t.lin_maxit = lin_maxit;
  return 1;
}
int dBatchTapeKeepSteps(dBatchID b, int on) { return b ? tape_keep_steps(b->impl, on != 0) : 0; }
int dBatchTapeStats(dBatchID b, long long *out2) {
  if (!b || !out2) return 0;
  out2[0] = b->impl->tape.iters;
  out2[1] = b->impl->tape.mismatches;
  return 1;
}
int dBatchTapeEnd(dBatchID b) { return b ? tape_end(b->impl) : 0; }
int dBatchStepBackward(dBatchID b, int nsteps, const float *grad_out, float *grad_in, float *grad_mu,
                       float *grad_bounce) {
  return b ? tape_backward(b->impl, nsteps, grad_out, grad_in, grad_mu, grad_bounce) : 0;
}

}  // extern "C"
