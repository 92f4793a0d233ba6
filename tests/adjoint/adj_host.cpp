// adj_host.cpp -- host build (double) of the step adjoint in
// ode-0.12_amd/csrc/adj_step.hpp, driven by a recorded run of the CPU
// restatement: the adjoint's mathematics is checked against central finite
// differences in double precision before the same code runs on the GPU in
// float.  Test harness only.
#include <cstring>
#include <vector>

#include "adj_step.hpp"

using namespace odemi;
using namespace odemi::adj;

namespace {
struct ParSerial {
  int tid() const { return 0; }
  int nt() const { return 1; }
  void sync() {}
  double sum(double v) { return v; }
  void add(double *p, double v) { *p += v; }
};
}  // namespace

extern "C" {

// bodies: mass[nb], I9[nb*9] (body frame); geoms: type[ng], body[ng],
// prm[ng*4], pos[ng*3], R9[ng*9] (static geoms); params[16]: h, erp, cfm,
// min_depth, max_vel, gx, gy, gz, mu, bounce, bounce_vel, surf_bounce, maxc,
// act_tol, lin_tol, lin_maxit.  states: (nsteps+1)*nb*13; nk[nsteps];
// contacts: 16 doubles each (see orc_tape_run).  Permanent joints: nj and
// jrec[nsteps*nj*64], each step's joint records of orc_tape_run.  grad_out:
// nb*13 (dL/d final state); grad_in: nb*13 out; gmu, gbounce out; stats[2]
// out.  Returns 0 when a recorded joint has rows the adjoint does not cover.
int adj_host_backward(int nb, int ng, const double *mass, const double *I9, const int *gtype, const int *gbody,
                      const double *gprm, const double *gpos, const double *gR9, const double *params, int nsteps,
                      const double *states, const int *nk, const double *contacts, int nj, const double *jrec,
                      const double *grad_out, double *grad_in, double *gmu, double *gbounce, int *stats) {
  StepParams<double> P;
  P.h = params[0]; P.erp = params[1]; P.cfm = params[2]; P.min_depth = params[3]; P.max_vel = params[4];
  P.gravity[0] = params[5]; P.gravity[1] = params[6]; P.gravity[2] = params[7];
  P.mu = params[8]; P.bounce = params[9]; P.bounce_vel = params[10]; P.surf_bounce = (int)params[11];
  P.maxc = (int)params[12]; P.act_tol = params[13]; P.lin_tol = params[14]; P.lin_maxit = (int)params[15];
  std::vector<double> m(nb), im(nb), I(12 * nb, 0.0), iI(12 * nb, 0.0);
  for (int b = 0; b < nb; b++) {
    m[b] = mass[b];
    im[b] = 1.0 / mass[b];
    const double *a = I9 + 9 * b;
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) I[12 * b + 4 * r + c] = a[3 * r + c];
    // inverse of the symmetric positive definite 3x3
    double det = a[0] * (a[4] * a[8] - a[5] * a[7]) - a[1] * (a[3] * a[8] - a[5] * a[6]) +
                 a[2] * (a[3] * a[7] - a[4] * a[6]);
    double inv[9] = {(a[4] * a[8] - a[5] * a[7]) / det, (a[2] * a[7] - a[1] * a[8]) / det,
                     (a[1] * a[5] - a[2] * a[4]) / det, (a[5] * a[6] - a[3] * a[8]) / det,
                     (a[0] * a[8] - a[2] * a[6]) / det, (a[2] * a[3] - a[0] * a[5]) / det,
                     (a[3] * a[7] - a[4] * a[6]) / det, (a[1] * a[6] - a[0] * a[7]) / det,
                     (a[0] * a[4] - a[1] * a[3]) / det};
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) iI[12 * b + 4 * r + c] = inv[3 * r + c];
  }
  std::vector<double> gp(4 * ng, 0.0), gR(12 * ng, 0.0);
  for (int g = 0; g < ng; g++) {
    for (int k = 0; k < 3; k++) gp[4 * g + k] = gpos[3 * g + k];
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) gR[12 * g + 4 * r + c] = gR9[9 * g + 3 * r + c];
  }
  std::vector<long> off(nsteps + 1, 0);
  for (int s = 0; s < nsteps; s++) off[s + 1] = off[s] + nk[s];
  // adjoint of the current (post-step) state, stride 4
  std::vector<double> ax(4 * nb, 0), aq(4 * nb, 0), av(4 * nb, 0), aw(4 * nb, 0);
  for (int b = 0; b < nb; b++)
    for (int k = 0; k < 3; k++) {
      ax[4 * b + k] = grad_out[13 * b + k];
      av[4 * b + k] = grad_out[13 * b + 7 + k];
      aw[4 * b + k] = grad_out[13 * b + 10 + k];
    }
  for (int b = 0; b < nb; b++)
    for (int k = 0; k < 4; k++) aq[4 * b + k] = grad_out[13 * b + 3 + k];
  *gmu = 0;
  *gbounce = 0;
  stats[0] = stats[1] = 0;
  for (int s = nsteps - 1; s >= 0; s--) {
    const double *st = states + (size_t)s * nb * 13;
    std::vector<double> x(4 * nb, 0), q(4 * nb, 0), v(4 * nb, 0), w(4 * nb, 0);
    for (int b = 0; b < nb; b++) {
      for (int k = 0; k < 3; k++) {
        x[4 * b + k] = st[13 * b + k];
        v[4 * b + k] = st[13 * b + 7 + k];
        w[4 * b + k] = st[13 * b + 10 + k];
      }
      for (int k = 0; k < 4; k++) q[4 * b + k] = st[13 * b + 3 + k];
    }
    std::vector<TapeContact<double>> con(nk[s]);
    for (int i = 0; i < nk[s]; i++) {
      const double *c = contacts + 16 * (size_t)(off[s] + i);
      TapeContact<double> &t = con[i];
      for (int k = 0; k < 3; k++) { t.pos[k] = c[k]; t.normal[k] = c[3 + k]; t.lam[k] = c[7 + k]; }
      t.depth = c[6];
      t.g1 = (int)c[10]; t.g2 = (int)c[11]; t.b1 = (int)c[12]; t.b2 = (int)c[13];
      t.flags = (int)c[14]; t.m = (int)c[15];
    }
    std::vector<TapeJoint<double>> jt(nj);
    for (int j = 0; j < nj; j++) {
      const double *o = jrec + 64 * ((size_t)s * nj + j);
      TapeJoint<double> &t = jt[j];
      if (o[62] == 0) return 0;
      for (int k = 0; k < 6; k++) {
        t.lam[k] = o[k]; t.lo[k] = o[6 + k]; t.hi[k] = o[12 + k]; t.rcfm[k] = o[18 + k];
      }
      for (int k = 0; k < 3; k++) {
        t.anc1[k] = o[24 + k]; t.anc2[k] = o[27 + k]; t.ax1[k] = o[30 + k]; t.ax2[k] = o[33 + k];
        t.v1[k] = o[40 + k]; t.v2[k] = o[43 + k];
      }
      t.erp = o[36]; t.susp_erp = o[37]; t.c0 = o[38]; t.s0 = o[39];
      for (int l = 0; l < 2; l++) {
        t.lerp[l] = o[46 + l]; t.lstop[l] = o[48 + l]; t.lc[l] = o[50 + l]; t.torque[l] = o[52 + l];
        t.lrow[l] = (int)o[58 + l]; t.lkind[l] = (int)o[60 + l];
      }
      t.type = (int)o[54]; t.b1 = (int)o[55]; t.b2 = (int)o[56]; t.m = (int)o[57];
    }
    StepData<double> D;
    D.nb = nb; D.nk = nk[s]; D.ng = ng;
    D.nj = nj; D.jnt = jt.data();
    D.x = x.data(); D.q = q.data(); D.v = v.data(); D.w = w.data();
    D.mass = m.data(); D.invMass = im.data(); D.I = I.data(); D.invI = iI.data();
    D.gtype = gtype; D.gbody = gbody; D.gprm = gprm; D.gpos = gp.data(); D.gR = gR.data();
    D.con = con.data();
    std::vector<double> ox(4 * nb, 0), oq(4 * nb, 0), ov(4 * nb, 0), ow(4 * nb, 0);
    StepAdj<double> G;
    G.gx_in = ax.data(); G.gq_in = aq.data(); G.gv_in = av.data(); G.gw_in = aw.data();
    G.gx = ox.data(); G.gq = oq.data(); G.gv = ov.data(); G.gw = ow.data();
    G.gmu = gmu; G.gbounce = gbounce; G.stats = stats;
    std::vector<double> scr(adj_scratch_elems<double>(nb, nk[s], nj), 0.0);
    ParSerial par;
    adjoint_step<double>(par, D, P, G, scr.data());
    ax.swap(ox); aq.swap(oq); av.swap(ov); aw.swap(ow);
  }
  for (int b = 0; b < nb; b++) {
    for (int k = 0; k < 3; k++) {
      grad_in[13 * b + k] = ax[4 * b + k];
      grad_in[13 * b + 7 + k] = av[4 * b + k];
      grad_in[13 * b + 10 + k] = aw[4 * b + k];
    }
    for (int k = 0; k < 4; k++) grad_in[13 * b + 3 + k] = aq[4 * b + k];
  }
  return 1;
}

}  // extern "C"
