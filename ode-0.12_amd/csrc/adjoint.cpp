// adjoint.cpp -- host side of dBatchStepBackward: the checkpoint tape, the
// segment replay and the reverse sweep over the recorded steps.
#include <algorithm>

#include "adj_step.hpp"
#include "odemi_internal.hpp"

namespace odemi {

hipError_t launch_tape_capture(const DevBatch &B, adj::TapeContact<float> *con, const long long *con_off, int *nk,
                               int *ovf, adj::TapeJoint<float> *jt, hipStream_t st);
hipError_t launch_adjoint_step(const DevBatch &B, const float4 *const st[4], const adj::TapeContact<float> *con,
                               const long long *con_off, const int *nk, const float4 *const gin[4],
                               float4 *const gout[4], float *gmu, float *gbounce, int *stats, float *scr,
                               const long long *scr_off, float act_tol, float lin_tol, int lin_maxit, int wg,
                               const adj::TapeJoint<float> *jt, hipStream_t st_);

namespace {

// A world the adjoint supports: contact joints, ball joints, hinges without
// limits, hinge2 joints, dContactApprox1 friction (or mu = 0), optional
// bounce, no other surface modes.  Rows a recorded step has that the adjoint
// does not cover (a hinge limit set by the controls, a motor against a
// one-sided limit, a bouncy limit) are reported by dBatchStepBackward.
bool world_supported(const BatchImpl *b, int w, char *why, size_t n) {
  const WorldInfo &W = b->wi[w];
  const WorldParams &P = b->wp[w];
  const float PI = 3.14159265358979323846f;
  for (int j = 0; j < W.nj; j++) {
    const JointDev &J = b->h_joints[W.j0 + j];
    if (J.type == JT_HINGE && (J.lm1.lostop >= -PI || J.lm1.histop <= PI) && J.lm1.lostop <= J.lm1.histop) {
      snprintf(why, n, "world %d has a hinge with a limit (the adjoint covers hinge motors, not hinge limits)", w);
      return false;
    }
  }
  for (int g = 0; g < W.ng; g++)
    if (b->h_gtype[W.g0 + g] != GEOM_SPHERE && b->h_gtype[W.g0 + g] != GEOM_BOX &&
        b->h_gtype[W.g0 + g] != GEOM_PLANE) {
      snprintf(why, n, "world %d has a capsule, cylinder or ray geom (the adjoint colliders are sphere, box and plane)", w);
      return false;
    }
  int other = P.surf_mode & ~(SM_BOUNCE | SM_APPROX1_1 | SM_APPROX1_2);
  if (other) {
    snprintf(why, n, "world %d uses contact surface mode bits 0x%x", w, other);
    return false;
  }
  if (P.mu > 0.f && (P.surf_mode & (SM_APPROX1_1 | SM_APPROX1_2)) != (SM_APPROX1_1 | SM_APPROX1_2)) {
    snprintf(why, n, "world %d: friction without dContactApprox1", w);
    return false;
  }
  return true;
}

}  // namespace

int tape_begin(BatchImpl *b, int max_steps) {
  if (max_steps <= 0) {
    report_error("dBatchTapeBegin: max_steps must be positive");
    return 0;
  }
  char why[160];
  for (int w = 0; w < b->nworlds; w++)
    if (!world_supported(b, w, why, sizeof(why))) {
      report_error("dBatchTapeBegin: %s", why);
      return 0;
    }
  Tape &t = b->tape;
  t.ckpt.clear();
  t.nsteps = 0;
  t.max_steps = 0;
  t.kept = false;
  if (t.keep) {
    const int nb = b->nb, nw = b->nworlds;
    std::vector<long long> con_off(nw);
    long long ncon = 0;
    for (int w = 0; w < nw; w++) {
      con_off[w] = ncon;
      ncon += b->wi[w].kcap;
    }
    t.kncon = ncon;
    const size_t stbytes = sizeof(float4) * (size_t)nb;
    if (!(t.kst.alloc(4 * stbytes * (size_t)max_steps) &&
          t.kcon.alloc(sizeof(adj::TapeContact<float>) * (size_t)ncon * max_steps) &&
          t.knk.alloc(sizeof(int) * (size_t)nw * max_steps) && t.kovf.alloc(sizeof(int)) &&
          t.kcon_off.alloc(sizeof(long long) * nw) &&
          (b->nj == 0 || t.kjnt.alloc(sizeof(adj::TapeJoint<float>) * (size_t)b->nj * max_steps)))) {
      report_error("dBatchTapeBegin: device allocation for %d kept steps failed", max_steps);
      return 0;
    }
    OD_HIP(hipMemcpy(t.kcon_off.p, con_off.data(), sizeof(long long) * nw, hipMemcpyHostToDevice));
    OD_HIP(hipMemset(t.kovf.p, 0, sizeof(int)));
  }
  if (b->nj > 0 && !t.kctrl.alloc(sizeof(float4) * (size_t)b->nj * max_steps)) {
    report_error("dBatchTapeBegin: device allocation for the controls of %d steps failed", max_steps);
    return 0;
  }
  t.ctrl_on.assign(max_steps, 0);
  t.ckpt.emplace_back(new Snapshot());
  if (!snapshot(b, *t.ckpt.back(), true)) return 0;
  // the mode is latched only once every buffer it needs exists; the backward
  // pass follows the latched mode, whatever dBatchTapeKeepSteps says later
  t.kept = t.keep;
  t.max_steps = max_steps;
  t.recording = true;
  destroy_graph(b);
  return 1;
}

int tape_keep_steps(BatchImpl *b, bool on) {
  if (b->tape.recording) {
    report_error("dBatchTapeKeepSteps: set before dBatchTapeBegin");
    return 0;
  }
  Tape &t = b->tape;
  t.keep = on;
  if (!on) {
    // the buffers go with the tape recorded in keep mode: that tape can no
    // longer be differentiated
    t.kst.release();
    t.kcon.release();
    t.knk.release();
    t.kovf.release();
    t.kcon_off.release();
    t.kjnt.release();
    if (t.kept) {
      t.kept = false;
      t.nsteps = 0;
      t.max_steps = 0;
      t.ckpt.clear();
    }
  }
  return 1;
}

int tape_end(BatchImpl *b) {
  b->tape.recording = false;
  return 1;
}

// one recorded step: launched directly (not from the graph); a checkpoint
// every `interval` steps
int tape_step(BatchImpl *b, bool sync) {
  Tape &t = b->tape;
  if (t.nsteps >= t.max_steps) {
    report_error("dBatchStep: the tape holds at most %d steps", t.max_steps);
    return 0;
  }
  if (b->nj > 0) {
    // the controls this step runs with, for the replays
    t.ctrl_on[t.nsteps] = b->ctrl_set ? 1 : 0;
    if (b->ctrl_set)
      OD_HIP(hipMemcpyAsync(t.kctrl.as<float4>() + (size_t)t.nsteps * b->nj, b->d.ctrl.p, sizeof(float4) * b->nj,
                            hipMemcpyDeviceToDevice, b->stream));
  }
  if (t.kept) {
    // this step's pre-state, then (after the step) its contacts
    const int nb = b->nb, nw = b->nworlds;
    const size_t stbytes = sizeof(float4) * (size_t)nb;
    float4 *slot = t.kst.as<float4>() + (size_t)4 * t.nsteps * nb;
    OD_HIP(hipMemcpyAsync(slot, b->d.pos.p, stbytes, hipMemcpyDeviceToDevice, b->stream));
    OD_HIP(hipMemcpyAsync(slot + nb, b->d.q.p, stbytes, hipMemcpyDeviceToDevice, b->stream));
    OD_HIP(hipMemcpyAsync(slot + 2 * (size_t)nb, b->d.lvel.p, stbytes, hipMemcpyDeviceToDevice, b->stream));
    OD_HIP(hipMemcpyAsync(slot + 3 * (size_t)nb, b->d.avel.p, stbytes, hipMemcpyDeviceToDevice, b->stream));
    if (!batch_launch_once(b)) return 0;
    OD_HIP(launch_tape_capture(b->dev, t.kcon.as<adj::TapeContact<float>>() + (size_t)t.nsteps * t.kncon,
                               t.kcon_off.as<long long>(), t.knk.as<int>() + (size_t)t.nsteps * nw,
                               t.kovf.as<int>(),
                               b->nj ? t.kjnt.as<adj::TapeJoint<float>>() + (size_t)t.nsteps * b->nj : nullptr,
                               b->stream));
  } else if (!batch_launch_once(b)) {
    return 0;
  }
  t.nsteps++;
  if (t.nsteps % t.interval == 0 && t.nsteps < t.max_steps) {
    t.ckpt.emplace_back(new Snapshot());
    if (!snapshot(b, *t.ckpt.back(), true)) return 0;
  }
  if (sync) OD_HIP(hipStreamSynchronize(b->stream));
  return 1;
}

int tape_backward(BatchImpl *b, int nsteps, const float *grad_out, float *grad_in, float *grad_mu,
                  float *grad_bounce) {
  Tape &t = b->tape;
  if (nsteps <= 0 || nsteps > t.nsteps || t.ckpt.empty()) {
    report_error("dBatchStepBackward: %d steps requested, %d recorded", nsteps, t.nsteps);
    return 0;
  }
  if (!grad_out || !grad_in) {
    report_error("dBatchStepBackward: null gradient buffer");
    return 0;
  }
  const int nb = b->nb, nw = b->nworlds, K = t.interval;
  // per-world offsets of contacts and adjoint scratch
  std::vector<long long> con_off(nw), scr_off(nw);
  long long ncon = 0, nscr = 0;
  for (int w = 0; w < nw; w++) {
    con_off[w] = ncon;
    scr_off[w] = nscr;
    ncon += b->wi[w].kcap;
    nscr += adj::adj_scratch_elems<float>(b->wi[w].nb, b->wi[w].kcap, b->wi[w].nj);
    nscr = (nscr + 63) & ~63LL;
  }
  DevBuf d_con_off, d_scr_off, seg_st, seg_con, seg_nk, seg_jnt, gbuf, scr, gmu, gbn, stats, ctrl_save;
  const size_t stbytes = sizeof(float4) * (size_t)nb;
  bool ok = d_con_off.alloc(sizeof(long long) * nw) && d_scr_off.alloc(sizeof(long long) * nw) &&
            seg_st.alloc(4 * stbytes * K) && seg_con.alloc(sizeof(adj::TapeContact<float>) * (size_t)ncon * K) &&
            seg_nk.alloc(sizeof(int) * (size_t)nw * K) && gbuf.alloc(8 * stbytes) &&
            scr.alloc(sizeof(float) * (size_t)nscr) && gmu.alloc(sizeof(float) * nw) &&
            gbn.alloc(sizeof(float) * nw) && stats.alloc(sizeof(int) * (2 * (size_t)nw + 1)) &&
            (b->nj == 0 || t.kept ||
             (seg_jnt.alloc(sizeof(adj::TapeJoint<float>) * (size_t)b->nj * K) &&
              ctrl_save.alloc(sizeof(float4) * (size_t)b->nj)));
  if (!ok) {
    report_error("dBatchStepBackward: device allocation failed (%lld contacts, %lld scratch floats per step)", ncon,
                 nscr);
    return 0;
  }
  OD_HIP(hipMemcpy(d_con_off.p, con_off.data(), sizeof(long long) * nw, hipMemcpyHostToDevice));
  OD_HIP(hipMemcpy(d_scr_off.p, scr_off.data(), sizeof(long long) * nw, hipMemcpyHostToDevice));
  OD_HIP(hipMemset(gmu.p, 0, gmu.n));
  OD_HIP(hipMemset(gbn.p, 0, gbn.n));
  OD_HIP(hipMemset(stats.p, 0, stats.n));
  // adjoint buffers: two sets of (x, q, v, w) float4 arrays
  float4 *G[2][4];
  for (int s = 0; s < 2; s++)
    for (int k = 0; k < 4; k++) G[s][k] = gbuf.as<float4>() + (size_t)(4 * s + k) * nb;
  {
    std::vector<float4> h[4];
    for (int k = 0; k < 4; k++) h[k].assign(nb, make_float4(0, 0, 0, 0));
    for (int i = 0; i < nb; i++) {
      const float *g = grad_out + 13 * (size_t)i;
      h[0][i] = make_float4(g[0], g[1], g[2], 0);
      h[1][i] = make_float4(g[3], g[4], g[5], g[6]);
      h[2][i] = make_float4(g[7], g[8], g[9], 0);
      h[3][i] = make_float4(g[10], g[11], g[12], 0);
    }
    for (int k = 0; k < 4; k++) OD_HIP(hipMemcpy(G[0][k], h[k].data(), stbytes, hipMemcpyHostToDevice));
  }
  int cur = 0;
  // the state at the end of the tape, restored afterwards
  Snapshot endsnap;
  if (!snapshot(b, endsnap, true)) return 0;
  const int T = t.nsteps, first = T - nsteps;
  bool fail = false;
  if (t.kept) {
    // the forward pass kept every step: reverse sweep only
    int ovf = 0;
    OD_HIP(hipMemcpy(&ovf, t.kovf.p, sizeof(int), hipMemcpyDeviceToHost));
    if (ovf) {
      report_error(ovf & 1 ? "dBatchStepBackward: a recorded step overflowed its pair/contact/row capacities"
                           : "dBatchStepBackward: a recorded step had joint rows the adjoint does not cover");
      fail = true;
    }
    for (int s = T - 1; s >= first && !fail; s--) {
      const float4 *slot = t.kst.as<float4>() + (size_t)4 * s * nb;
      const float4 *st4[4] = {slot, slot + nb, slot + 2 * (size_t)nb, slot + 3 * (size_t)nb};
      const float4 *gin[4] = {G[cur][0], G[cur][1], G[cur][2], G[cur][3]};
      float4 *gout[4] = {G[1 - cur][0], G[1 - cur][1], G[1 - cur][2], G[1 - cur][3]};
      OD_HIP(launch_adjoint_step(b->dev, st4, t.kcon.as<adj::TapeContact<float>>() + (size_t)s * t.kncon,
                                 t.kcon_off.as<long long>(), t.knk.as<int>() + (size_t)s * nw, gin, gout,
                                 gmu.as<float>(), gbn.as<float>(), stats.as<int>(), scr.as<float>(),
                                 d_scr_off.as<long long>(), t.act_tol, t.lin_tol, t.lin_maxit, 256,
                                 b->nj ? t.kjnt.as<adj::TapeJoint<float>>() + (size_t)s * b->nj : nullptr,
                                 b->stream));
      cur = 1 - cur;
    }
  }
  // the replays re-apply each step's controls; the current ones are restored
  // when this function returns, on every path
  struct CtrlRestore {
    BatchImpl *b;
    DevBuf *save;
    bool set0, armed;
    ~CtrlRestore() {
      if (!armed) return;
      b->ctrl_set = set0;
      (void)hipMemcpyAsync(b->d.ctrl.p, save->p, sizeof(float4) * b->nj, hipMemcpyDeviceToDevice, b->stream);
    }
  } ctrl_restore{b, &ctrl_save, b->ctrl_set, false};
  if (!t.kept && b->nj > 0) {
    OD_HIP(hipMemcpyAsync(ctrl_save.p, b->d.ctrl.p, sizeof(float4) * b->nj, hipMemcpyDeviceToDevice, b->stream));
    ctrl_restore.armed = true;
  }
  int seg0 = ((T - 1) / K) * K;
  const int seg_first = (first / K) * K;
  for (; !t.kept && seg0 >= seg_first && !fail; seg0 -= K) {
    const int seg1 = std::min(seg0 + K, T);
    if (!snapshot(b, *t.ckpt[seg0 / K], false)) return 0;
    // replay the segment, recording each step's pre-state and contacts
    for (int s = seg0; s < seg1; s++) {
      int j = s - seg0;
      float4 *slot = seg_st.as<float4>() + (size_t)4 * j * nb;
      OD_HIP(hipMemcpyAsync(slot, b->d.pos.p, stbytes, hipMemcpyDeviceToDevice, b->stream));
      OD_HIP(hipMemcpyAsync(slot + nb, b->d.q.p, stbytes, hipMemcpyDeviceToDevice, b->stream));
      OD_HIP(hipMemcpyAsync(slot + 2 * (size_t)nb, b->d.lvel.p, stbytes, hipMemcpyDeviceToDevice, b->stream));
      OD_HIP(hipMemcpyAsync(slot + 3 * (size_t)nb, b->d.avel.p, stbytes, hipMemcpyDeviceToDevice, b->stream));
      if (b->nj > 0) {
        b->ctrl_set = t.ctrl_on[s] != 0;
        if (b->ctrl_set)
          OD_HIP(hipMemcpyAsync(b->d.ctrl.p, t.kctrl.as<float4>() + (size_t)s * b->nj, sizeof(float4) * b->nj,
                                hipMemcpyDeviceToDevice, b->stream));
      }
      if (!batch_launch_once(b)) return 0;
      OD_HIP(launch_tape_capture(b->dev, seg_con.as<adj::TapeContact<float>>() + (size_t)j * ncon,
                                 d_con_off.as<long long>(), seg_nk.as<int>() + (size_t)j * nw,
                                 stats.as<int>() + 2 * (size_t)nw,
                                 b->nj ? seg_jnt.as<adj::TapeJoint<float>>() + (size_t)j * b->nj : nullptr,
                                 b->stream));
    }
    {
      int ovf = 0;
      OD_HIP(hipMemcpy(&ovf, stats.as<int>() + 2 * (size_t)nw, sizeof(int), hipMemcpyDeviceToHost));
      if (ovf) {
        report_error(ovf & 1 ? "dBatchStepBackward: a replayed step overflowed its pair/contact/row capacities"
                             : "dBatchStepBackward: a replayed step had joint rows the adjoint does not cover");
        fail = true;
        break;
      }
    }
    // reverse sweep over the segment
    for (int s = seg1 - 1; s >= std::max(seg0, first); s--) {
      int j = s - seg0;
      const float4 *slot = seg_st.as<float4>() + (size_t)4 * j * nb;
      const float4 *st4[4] = {slot, slot + nb, slot + 2 * (size_t)nb, slot + 3 * (size_t)nb};
      const float4 *gin[4] = {G[cur][0], G[cur][1], G[cur][2], G[cur][3]};
      float4 *gout[4] = {G[1 - cur][0], G[1 - cur][1], G[1 - cur][2], G[1 - cur][3]};
      OD_HIP(launch_adjoint_step(b->dev, st4, seg_con.as<adj::TapeContact<float>>() + (size_t)j * ncon,
                                 d_con_off.as<long long>(), seg_nk.as<int>() + (size_t)j * nw, gin, gout,
                                 gmu.as<float>(), gbn.as<float>(), stats.as<int>(), scr.as<float>(),
                                 d_scr_off.as<long long>(), t.act_tol, t.lin_tol, t.lin_maxit, 256,
                                 b->nj ? seg_jnt.as<adj::TapeJoint<float>>() + (size_t)j * b->nj : nullptr,
                                 b->stream));
      cur = 1 - cur;
    }
  }
  if (ctrl_restore.armed) {
    b->ctrl_set = ctrl_restore.set0;
    OD_HIP(hipMemcpyAsync(b->d.ctrl.p, ctrl_save.p, sizeof(float4) * b->nj, hipMemcpyDeviceToDevice, b->stream));
    ctrl_restore.armed = false;
  }
  if (!snapshot(b, endsnap, false)) return 0;
  if (fail) return 0;
  OD_HIP(hipStreamSynchronize(b->stream));
  {
    std::vector<float4> h[4];
    for (int k = 0; k < 4; k++) {
      h[k].resize(nb);
      OD_HIP(hipMemcpy(h[k].data(), G[cur][k], stbytes, hipMemcpyDeviceToHost));
    }
    for (int i = 0; i < nb; i++) {
      float *g = grad_in + 13 * (size_t)i;
      g[0] = h[0][i].x; g[1] = h[0][i].y; g[2] = h[0][i].z;
      g[3] = h[1][i].x; g[4] = h[1][i].y; g[5] = h[1][i].z; g[6] = h[1][i].w;
      g[7] = h[2][i].x; g[8] = h[2][i].y; g[9] = h[2][i].z;
      g[10] = h[3][i].x; g[11] = h[3][i].y; g[12] = h[3][i].z;
    }
  }
  if (grad_mu) OD_HIP(hipMemcpy(grad_mu, gmu.p, sizeof(float) * nw, hipMemcpyDeviceToHost));
  if (grad_bounce) OD_HIP(hipMemcpy(grad_bounce, gbn.p, sizeof(float) * nw, hipMemcpyDeviceToHost));
  std::vector<int> st(2 * (size_t)nw);
  OD_HIP(hipMemcpy(st.data(), stats.p, sizeof(int) * st.size(), hipMemcpyDeviceToHost));
  t.iters = 0;
  t.mismatches = 0;
  for (int w = 0; w < nw; w++) {
    t.iters += st[2 * w];
    t.mismatches += st[2 * w + 1];
  }
  return 1;
}

}  // namespace odemi
