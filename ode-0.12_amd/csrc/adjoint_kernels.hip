// adjoint_kernels.hip -- device side of the step adjoint: capture of a
// replayed step's contacts and row impulses into the segment tape, and the
// reverse step (adj_step.hpp) with one workgroup per world.
#include <hip/hip_runtime.h>

#include "adj_step.hpp"
#include "layout.hpp"

namespace odemi {

using adj::TapeContact;
using adj::TapeJoint;

struct AdjShared {
  float wsum[16];
};

__device__ inline float adj_block_sum(float v, AdjShared &sh) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) sh.wsum[wid] = v;
  __syncthreads();
  float s = 0.f;
  int nw = (blockDim.x + 63) >> 6;
  for (int k = 0; k < nw; k++) s += sh.wsum[k];
  __syncthreads();
  return s;
}

// workgroup context of adj::adjoint_step
struct ParWG {
  AdjShared *sh;
  __device__ int tid() const { return threadIdx.x; }
  __device__ int nt() const { return blockDim.x; }
  __device__ void sync() { __syncthreads(); }
  __device__ float sum(float v) { return adj_block_sum(v, *sh); }
  __device__ void add(float *p, float v) { atomicAdd(p, v); }
};

// settings of the reverse step
struct AdjSettings {
  float act_tol;    // impulses at most this count as at their bound
  float lin_tol;    // relative residual of the adjoint linear solve
  int lin_maxit;
};

// Every permanent joint of world w after its step (a joint's constraint
// follows the step's nk contacts): frame data, Info2 erp, each row's bounds
// and cfm / h and impulse, the limit / motor rows and the hinge2 torques of
// the step's controls.  ovf |= 2 when a joint has rows the adjoint does not
// cover.
__device__ void capture_joints(const DevBatch &B, int w, const WorldInfo &W, const char *S, const ScratchLayout &L,
                               const Cons *cons, const Row *rows, int nk, TapeJoint<float> *jt, int *ovf) {
  const float *erp_of = B.order_mode == ORDER_REFERENCE ? (const float *)((const int *)(S + L.color) + W.nb) : nullptr;
  for (int j = threadIdx.x; j < W.nj; j += blockDim.x) {
    const JointDev &J = B.joints[W.j0 + j];
    const Cons &q = cons[nk + j];
    TapeJoint<float> t;
    t.type = J.type; t.b1 = J.b1; t.b2 = J.b2;
    t.m = q.m;
    bool covered = q.m <= 6 && q.row0 + q.m <= W.rcap;
    for (int r = 0; r < 6; r++) {
      const bool in = covered && r < q.m;
      const Row *R = in ? rows + q.row0 + r : nullptr;
      t.lam[r] = in ? R->lambda : 0.f;
      t.lo[r] = in ? R->lo : 0.f;
      t.hi[r] = in ? R->hi : 0.f;
      t.rcfm[r] = in ? R->cfm : 0.f;   // divided by h in the step's rhs pass
    }
    for (int k = 0; k < 3; k++) {
      t.anc1[k] = J.anchor1[k]; t.anc2[k] = J.anchor2[k];
      t.ax1[k] = J.axis1[k]; t.ax2[k] = J.axis2[k];
      t.v1[k] = J.v1[k]; t.v2[k] = J.v2[k];
    }
    // Info2 erp: a ball joint's own; otherwise the world's, or in the
    // reference order the running value the step used
    t.erp = J.type == JT_BALL ? J.erp : (erp_of ? erp_of[nk + j] : B.wp[w].erp);
    t.susp_erp = J.susp_erp;
    t.c0 = J.c0;
    t.s0 = J.s0;
    // limit / motor rows after the joint's own rows (joint.cpp:556-733)
    int row = J.type == JT_HINGE2 ? 4 : 5;
    for (int l = 0; l < 2; l++) {
      t.lrow[l] = -1;
      t.lkind[l] = 0;
      t.lerp[l] = t.lstop[l] = t.lc[l] = 0.f;
      const Limot &lm = l == 0 ? J.lm1 : J.lm2;
      if (J.type == JT_BALL || (l == 1 && J.type != JT_HINGE2)) continue;
      bool powered = lm.fmax > 0.f;
      if (!(powered || lm.limit)) continue;
      t.lrow[l] = row;
      if (lm.limit && lm.lostop == lm.histop) powered = false;
      if (powered && !lm.limit) {
        t.lkind[l] = 2;
        t.lc[l] = lm.vel;
      } else {
        // a motor pushing against a one-sided limit, a bouncy limit and the
        // hinge's angle are not covered
        if (powered || (lm.bounce > 0.f && lm.lostop != lm.histop) || J.type == JT_HINGE) covered = false;
        t.lkind[l] = 1;
        t.lerp[l] = lm.stop_erp;
        t.lstop[l] = lm.limit == 1 ? lm.lostop : lm.histop;
      }
      row++;
    }
    if (row != q.m && J.type != JT_BALL) covered = false;
    if (J.type == JT_BALL && q.m != 3) covered = false;
    t.torque[0] = t.torque[1] = 0.f;
    if (J.type == JT_HINGE2 && B.ctrl && J.b1 >= 0 && J.b2 >= 0) {
      const float4 u = B.ctrl[W.j0 + j];
      t.torque[0] = u.z;
      t.torque[1] = u.w;
    }
    if (!covered) atomicOr(ovf, 2);
    jt[W.j0 + j] = t;
  }
}

// After a replayed forward step: the step's contacts (in creation order)
// with the impulses of their rows, from the world's scratch; and (jt not
// null) every permanent joint's frame data, Info2 erp / cfm and row impulses
// (a joint's constraint follows the step's nk contacts).  ovf: 1 when a
// capacity overflowed, 2 when a joint had rows the adjoint does not cover.
__global__ void tape_capture_kernel(DevBatch B, TapeContact<float> *con, const long long *con_off, int *nk_out,
                                    int *ovf, TapeJoint<float> *jt) {
  const int w = blockIdx.x;
  if (w >= B.nworlds) return;
  const WorldInfo W = B.wi[w];
  const ScratchLayout L = scratch_layout(W.nb, W.ng, W.nj, W.pcap, W.kcap, W.rcap);
  const char *S = B.scratch + W.scratch;
  const ContactRec *cr = (const ContactRec *)(S + L.contacts);
  const Cons *cons = (const Cons *)(S + L.cons);
  const Row *rows = (const Row *)(S + L.rows);
  int nk = B.counters[w * CNT_N + CNT_CONTACTS];
  if (nk > W.kcap) nk = W.kcap;
  const int flags = B.counters[w * CNT_N + CNT_OVERFLOW];
  if (flags & OVF_ROWS) nk = 0;   // the step left its rows unbuilt: nothing valid to record
  TapeContact<float> *dst = con + con_off[w];
  for (int i = threadIdx.x; i < nk; i += blockDim.x) {
    const ContactRec &c = cr[i];
    TapeContact<float> t;
    for (int k = 0; k < 3; k++) { t.pos[k] = c.pos[k]; t.normal[k] = c.normal[k]; }
    t.depth = c.depth;
    t.g1 = c.g1; t.g2 = c.g2; t.b1 = c.b1; t.b2 = c.b2; t.flags = c.flags;
    int m = cons[i].m;
    t.m = m;
    for (int r = 0; r < 3; r++) {
      int ri = cons[i].row0 + r;
      t.lam[r] = (r < m && ri < W.rcap) ? rows[ri].lambda : 0.f;
    }
    dst[i] = t;
  }
  if (jt && !(flags & OVF_ROWS)) capture_joints(B, w, W, S, L, cons, rows, nk, jt, ovf);
  if (threadIdx.x == 0) {
    nk_out[w] = nk;
    if (flags & (OVF_PAIRS | OVF_CONTACTS | OVF_ROWS)) atomicOr(ovf, 1);
  }
}

// One reverse step for every world.  st_*: the step's pre-step state;
// gin_*: adjoint of the post-step state; gout_*: adjoint of the pre-step
// state (written).  Per-world parameter adjoints accumulate into gmu,
// gbounce; stats: per world [iterations, collider mismatches].
__global__ __launch_bounds__(256) void adjoint_step_kernel(
    DevBatch B, const float4 *st_x, const float4 *st_q, const float4 *st_v, const float4 *st_w,
    const TapeContact<float> *con, const long long *con_off, const int *nk, const float4 *gin_x,
    const float4 *gin_q, const float4 *gin_v, const float4 *gin_w, float4 *gout_x, float4 *gout_q, float4 *gout_v,
    float4 *gout_w, float *gmu, float *gbounce, int *stats, float *scr, const long long *scr_off, AdjSettings S,
    const TapeJoint<float> *jt) {
  __shared__ AdjShared sh;
  const int w = blockIdx.x;
  if (w >= B.nworlds) return;
  const WorldInfo W = B.wi[w];
  const WorldParams &P = B.wp[w];
  adj::StepData<float> D;
  D.nb = W.nb;
  D.nk = nk[w];
  D.ng = W.ng;
  D.x = (const float *)(st_x + W.b0);
  D.q = (const float *)(st_q + W.b0);
  D.v = (const float *)(st_v + W.b0);
  D.w = (const float *)(st_w + W.b0);
  D.mass = B.mass + W.b0;
  D.invMass = B.invMass + W.b0;
  D.I = (const float *)(B.I + 3 * (size_t)W.b0);
  D.invI = (const float *)(B.invI + 3 * (size_t)W.b0);
  D.gtype = B.gtype + W.g0;
  D.gbody = B.gbody + W.g0;
  D.gprm = (const float *)(B.gparam + W.g0);
  D.gpos = (const float *)(B.gpos + W.g0);
  D.gR = (const float *)(B.gR + 3 * (size_t)W.g0);
  D.con = con + con_off[w];
  if (jt) {
    D.nj = W.nj;
    D.jnt = jt + W.j0;
  }
  adj::StepParams<float> Q;
  Q.h = P.dt; Q.erp = P.erp; Q.cfm = P.cfm; Q.min_depth = P.min_depth; Q.max_vel = P.max_vel;
  for (int k = 0; k < 3; k++) Q.gravity[k] = P.gravity[k];
  Q.mu = P.mu; Q.bounce = P.bounce; Q.bounce_vel = P.bounce_vel;
  Q.surf_bounce = (P.surf_mode & SM_BOUNCE) ? 1 : 0;
  Q.maxc = P.max_contacts;
  Q.act_tol = S.act_tol; Q.lin_tol = S.lin_tol; Q.lin_maxit = S.lin_maxit;
  adj::StepAdj<float> G;
  G.gx_in = (const float *)(gin_x + W.b0);
  G.gq_in = (const float *)(gin_q + W.b0);
  G.gv_in = (const float *)(gin_v + W.b0);
  G.gw_in = (const float *)(gin_w + W.b0);
  G.gx = (float *)(gout_x + W.b0);
  G.gq = (float *)(gout_q + W.b0);
  G.gv = (float *)(gout_v + W.b0);
  G.gw = (float *)(gout_w + W.b0);
  G.gmu = gmu + w;
  G.gbounce = gbounce + w;
  G.stats = stats + 2 * w;
  ParWG par{&sh};
  adj::adjoint_step<float>(par, D, Q, G, scr + scr_off[w]);
}

// After a dWorldQuickStep of the ODE API path (one world, contacts from the
// host's contact joints): the step's contacts in creation order with the
// impulses of their rows, for dWorldQuickStepBackward.
__global__ void api_tape_capture_kernel(DevBatch B, const ContactIn *cin, int nk, TapeContact<float> *con,
                                        int *nk_out, TapeJoint<float> *jt) {
  const WorldInfo W = B.wi[0];
  const ScratchLayout L = scratch_layout(W.nb, W.ng, W.nj, W.pcap, W.kcap, W.rcap);
  const char *S = B.scratch + W.scratch;
  const Cons *cons = (const Cons *)(S + L.cons);
  const Row *rows = (const Row *)(S + L.rows);
  const int flags = B.counters[CNT_OVERFLOW];
  if (flags & OVF_ROWS) nk = 0;
  if (nk > W.kcap) nk = W.kcap;
  for (int i = threadIdx.x; i < nk; i += blockDim.x) {
    const ContactIn &c = cin[i];
    TapeContact<float> t;
    for (int k = 0; k < 3; k++) { t.pos[k] = c.pos[k]; t.normal[k] = c.normal[k]; }
    t.depth = c.depth;
    t.g1 = c.g1; t.g2 = c.g2; t.b1 = c.b1; t.b2 = c.b2; t.flags = c.flags;
    const int m = cons[i].m;
    t.m = m;
    for (int r = 0; r < 3; r++) {
      const int ri = cons[i].row0 + r;
      t.lam[r] = (r < m && ri < W.rcap) ? rows[ri].lambda : 0.f;
    }
    con[i] = t;
  }
  // nk_out[1]: 2 when a joint has rows the adjoint does not cover
  if (threadIdx.x == 0) nk_out[1] = 0;
  __syncthreads();
  if (jt && !(flags & OVF_ROWS)) capture_joints(B, 0, W, S, L, cons, rows, nk, jt, nk_out + 1);
  if (threadIdx.x == 0) *nk_out = nk;
}

hipError_t launch_api_tape_capture(const DevBatch &B, const ContactIn *cin, int nk, TapeContact<float> *con,
                                   int *nk_out, TapeJoint<float> *jt, hipStream_t st) {
  api_tape_capture_kernel<<<1, 128, 0, st>>>(B, cin, nk, con, nk_out, jt);
  return hipGetLastError();
}

hipError_t launch_tape_capture(const DevBatch &B, TapeContact<float> *con, const long long *con_off, int *nk,
                               int *ovf, TapeJoint<float> *jt, hipStream_t st) {
  if (B.nworlds <= 0) return hipSuccess;
  tape_capture_kernel<<<B.nworlds, 128, 0, st>>>(B, con, con_off, nk, ovf, jt);
  return hipGetLastError();
}

hipError_t launch_adjoint_step(const DevBatch &B, const float4 *const st[4], const TapeContact<float> *con,
                               const long long *con_off, const int *nk, const float4 *const gin[4],
                               float4 *const gout[4], float *gmu, float *gbounce, int *stats, float *scr,
                               const long long *scr_off, float act_tol, float lin_tol, int lin_maxit, int wg,
                               const TapeJoint<float> *jt, hipStream_t st_) {
  if (B.nworlds <= 0) return hipSuccess;
  if (wg < 64 || wg > 256 || (wg & 63)) return hipErrorInvalidValue;
  AdjSettings S{act_tol, lin_tol, lin_maxit};
  adjoint_step_kernel<<<B.nworlds, wg, 0, st_>>>(B, st[0], st[1], st[2], st[3], con, con_off, nk, gin[0], gin[1],
                                                  gin[2], gin[3], gout[0], gout[1], gout[2], gout[3], gmu, gbounce,
                                                  stats, scr, scr_off, S, jt);
  return hipGetLastError();
}

}  // namespace odemi
