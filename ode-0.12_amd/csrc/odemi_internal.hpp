// odemi_internal.hpp -- declarations shared by the host-side translation
// units of the library (batch API, ODE API layer, adjoint).
#pragma once
#include <memory>

#include "host_common.hpp"
#include "odemi.h"

namespace odemi {

// device state that one step mutates; saved and restored around autotuning
// trials and as the adjoint's checkpoints
struct Snapshot {
  DevBuf pos, q, lvel, avel, facc, tacc, R, grank, joints, wp, warm, warm_n;
};

// Forward tape for the adjoint: checkpoints of the state every `interval`
// recorded steps; the backward pass replays each segment from its
// checkpoint and differentiates it step by step in reverse.
struct Tape {
  bool recording = false;
  int interval = 10;
  int max_steps = 0;
  int nsteps = 0;                                // steps recorded since dBatchTapeBegin
  std::vector<std::unique_ptr<Snapshot>> ckpt;   // state at tape steps 0, interval, 2*interval, ...
  float act_tol = 1e-6f;                         // active-set threshold on impulses
  float lin_tol = 1e-6f;                         // relative residual of the adjoint solve
  int lin_maxit = 500;
  long long iters = 0, mismatches = 0;           // statistics of the last backward pass
  // keep mode: every recorded step's pre-state (x, q, v, w) and contacts,
  // written during the forward pass, so that the backward pass needs no
  // replay (memory per step: 64 B per body plus the contact capacity)
  bool keep = false;                             // requested for the next dBatchTapeBegin
  bool kept = false;                             // mode the current tape was recorded in
  DevBuf kst, kcon, knk, kovf, kcon_off, kjnt;   // kjnt: every permanent joint's record per kept step
  long long kncon = 0;
  // the joint controls each recorded step ran with (replays re-apply them)
  DevBuf kctrl;
  std::vector<char> ctrl_on;
};

// device arrays of a batch
struct BatchBufs {
  DevBuf wi, wp, pos, q, lvel, avel, facc, tacc, R, I, invI, mass, invMass;
  DevBuf gtype, gbody, gparam, gpos, gR, grank, gcat, gcol;
  DevBuf joints, adj_start, adj_other, bgeom_start, bgeom;
  DevBuf ctrl, counters, residual, work_counter, world_order, world_rank, cost_est, scratch, feedback, prof, warm,
      warm_n;
};

struct BatchImpl {
  int nworlds = 0;
  int nb = 0, ng = 0, nj = 0;
  int order_mode = ORDER_COLORED;
  std::vector<WorldInfo> wi;
  std::vector<WorldParams> wp;
  std::vector<int> h_gtype;   // geom types (host copy)
  std::vector<JointDev> h_joints;   // permanent joints as created (host copy)
  BatchBufs d;
  DevBatch dev{};
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int wg = 256;
  int grid = 0;           // persistent grid size (0: one workgroup per world)
  float cell_scale = 1.0f;
  bool ctrl_set = false;
  int cu_count = 256;
  int lds = 0;            // dynamic LDS pool per workgroup (bytes)
  int lds_cu = 65536;     // LDS per CU
  int lds_max = 65536;    // largest pool one workgroup may request
  bool prof_on = false;   // per-phase wall-clock counters
  bool split = false;     // collide and quickstep as two kernels per step
  // persistent launches pull the worlds in order of their estimated solver
  // work (rows x sweeps) instead of the static largest-first order
  bool cost_order = false;
  bool order_rewritten = false;   // world_order differs from the static order
  bool tol_on = false;    // some world iterates to a residual tolerance
  // coloured order at a tolerance: the sweeps in a kernel of their own
  // (collide + set-up | sweeps | integration), with their own launch shape
  bool sor_split = false;
  int sor_wg = 128, sor_lds = 0, sor_grid = 0;
  int wall_khz = 100000;  // wall clock rate
  Tape tape;              // forward tape for the adjoint
  DevBuf lcp_ws;          // exact LCP (dWorldStep form): per-world workspaces
  int lcp_rows = 0;       // largest island (rows) they hold; 0: off
};

// host arrays of a batch before upload
struct BatchArrays {
  std::vector<WorldInfo> wi;
  std::vector<WorldParams> wp;
  std::vector<float4> pos, q, lvel, avel, facc, tacc, I, invI, gparam, gpos, gR;
  std::vector<float> mass, invMass;
  std::vector<int> gtype, gbody, grank, adj_start, adj_other, bgeom_start, bgeom;
  std::vector<unsigned> gcat, gcol;
  std::vector<JointDev> joints;
};

void destroy_graph(BatchImpl *b);
void refresh_dev(BatchImpl *b);
BatchImpl *batch_create_arrays(BatchArrays &A, int order_mode);
BatchImpl *batch_create(const dBatchDesc *D, int order_mode, double cap_scale);
void batch_destroy(BatchImpl *b);
int batch_step(BatchImpl *b, int nsteps, bool sync);
int batch_step_plain(BatchImpl *b, int nsteps, bool sync);
int batch_launch_once(BatchImpl *b);
int batch_get_state(BatchImpl *b, float *out);
int batch_set_state(BatchImpl *b, const float *in);
int batch_counters(BatchImpl *b, long long *out5);
int batch_set_controls(BatchImpl *b, const double *ctrl);
int batch_set_launch(BatchImpl *b, int wg, float cell_scale, int grid);
int batch_set_lds(BatchImpl *b, int bytes);
int batch_set_sor_launch(BatchImpl *b, int on, int wg, int lds, int grid);
int batch_set_tolerance(BatchImpl *b, float tol, int max_iterations, int check_every, int accel);
int batch_set_warm_start(BatchImpl *b, float warm);
int batch_set_profiling(BatchImpl *b, bool on);
int batch_phase_times(BatchImpl *b, double *ms);
int batch_autotune(BatchImpl *b, int trial_steps, int flags);

int snapshot(BatchImpl *b, Snapshot &s, bool save);
int tape_begin(BatchImpl *b, int max_steps);
int tape_keep_steps(BatchImpl *b, bool on);
int tape_end(BatchImpl *b);
int tape_step(BatchImpl *b, bool sync);
int tape_backward(BatchImpl *b, int nsteps, const float *grad_out, float *grad_in, float *grad_mu,
                  float *grad_bounce);

}  // namespace odemi
