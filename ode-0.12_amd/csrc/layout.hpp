// layout.hpp -- data layout shared by the host library and the HIP kernels.
//
// Device state is structure-of-arrays in HBM and stays resident across steps.
// Worlds are contiguous ranges of the global body / geom / joint arrays; all
// indices stored inside a world (geom->body, joint->bodies) are local to it.
// Every world also owns a private scratch region whose layout is computed by
// scratch_layout() identically on host and device.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define OD_HD __host__ __device__
#else
#define OD_HD
// host-only builds (the adjoint's CPU checks) see the device vector types as
// plain structs
struct float3 { float x, y, z; };
struct float4 { float x, y, z, w; };
#endif

namespace odemi {

enum { GEOM_SPHERE = 0, GEOM_BOX = 1, GEOM_PLANE = 2, GEOM_CAPSULE = 3, GEOM_RAY = 4, GEOM_CYLINDER = 5 };
enum { JT_CONTACT = 0, JT_BALL = 1, JT_HINGE = 2, JT_HINGE2 = 3 };
enum { ORDER_REFERENCE = 0, ORDER_COLORED = 1 };
enum { JF_REVERSE = 2 };
// contact surface mode bits (include/ode/contact.h:33-49)
enum {
  SM_MU2 = 0x001, SM_FDIR1 = 0x002, SM_BOUNCE = 0x004, SM_SOFT_ERP = 0x008, SM_SOFT_CFM = 0x010,
  SM_MOTION1 = 0x020, SM_MOTION2 = 0x040, SM_MOTIONN = 0x080, SM_SLIP1 = 0x100, SM_SLIP2 = 0x200,
  SM_APPROX1_1 = 0x1000, SM_APPROX1_2 = 0x2000
};
// per-world counters written by the step kernels
// per-world counters: the first six describe the last step; CNT_UNCONV and
// CNT_RESMAX accumulate over the steps since dBatchResetSolveStats (or the
// batch's creation): tolerance-mode solves that ended at the iteration cap
// above the tolerance, and the largest final residual of a tolerance-mode
// solve (float bits)
enum { CNT_PAIRS = 0, CNT_CONTACTS, CNT_ROWS, CNT_GROUPS, CNT_OVERFLOW, CNT_ITERS, CNT_UNCONV, CNT_RESMAX, CNT_N };
enum { OVF_CELLS = 1, OVF_PAIRS = 2, OVF_CONTACTS = 4, OVF_ROWS = 8, OVF_NONFINITE = 16, OVF_LCP = 32 };

static const int MAX_SLOTS = 8;   // most contacts a collider produces for one pair

struct WorldInfo {
  int b0, nb;        // bodies
  int g0, ng;        // geoms
  int j0, nj;        // permanent joints
  int pcap;          // pair capacity
  int kcap;          // contact capacity
  int rcap;          // row capacity
  int adj0;          // offset into the permanent-joint adjacency arrays
  long long scratch; // byte offset of the world's scratch region
  long long warm;    // byte offset of the world's warm-start records
};

struct WorldParams {
  float gravity[3];
  float erp, cfm;
  int iterations;
  float sor_w;
  float max_vel, min_depth;
  float dt;
  int surf_mode;
  float mu, mu2, bounce, bounce_vel, soft_erp, soft_cfm;
  float motion1, motion2, motionN, slip1, slip2;
  int max_contacts;
  int skip_connected;
  int hash_min, hash_max;
  unsigned int seed;       // SOR shuffle RNG state (advanced by the kernel)
  float cell_scale;        // uniform-grid cell = cell_scale * largest AABB extent
  float tol;               // coloured solver: stop once the LCP residual <= tol (0: fixed iterations)
  int max_iterations;      // iteration cap when tol > 0
  int check_every;         // iterations between residual checks when tol > 0
  int accel;               // with tol: nonsmooth nonlinear conjugate-gradient steps between sweeps
  int dead;                // set by the step when the world's state became non-finite; the world is frozen
  float warm;              // coloured solver: start from warm x the previous step's impulses (0: from zero)
  int exact;               // reference order: each island's LCP solved exactly (dWorldStep)
};

struct Limot {             // ode/src/joints/joint.h:172-190
  float vel, fmax, lostop, histop, fudge, normal_cfm, stop_erp, stop_cfm, bounce;
  int limit;
  float limit_err;
  float pad;
};

struct JointDev {          // a permanent joint
  int type, b1, b2, flags; // b1/b2: node order after dJointAttach (b2 may be -1)
  float anchor1[4], anchor2[4], axis1[4], axis2[4], qrel[4], v1[4], v2[4];
  float erp, cfm, c0, s0, susp_erp, susp_cfm;
  int seq;                 // attach sequence number (body joint lists are newest first)
  int fb;                  // feedback slot or -1
  Limot lm1, lm2;
};

struct ContactIn {         // a contact joint created on the host (ODE API path)
  float pos[4], normal[4], fdir1[4];
  float depth;
  int b1, b2, flags;       // node order after dJointAttach
  int seq, mode, fb, g1;   // g1, g2: mirror indices of the contact's geoms (-1: none)
  int g2, pad[3];
  float mu, mu2, bounce, bounce_vel, soft_erp, soft_cfm, motion1, motion2, motionN, slip1, slip2, pad2;
};

// a candidate pair; key orders pairs as the reference's hash space reports them
struct PairRec {
  int g1, g2;
  unsigned long long key;
};

// a contact (world-local geom/body indices)
struct ContactRec {
  float pos[3], depth;
  float normal[3];
  int g1, g2;
  int b1, b2;              // node order after attach (b1 >= 0 unless both static)
  int flags;               // JF_REVERSE; bits 8+: index of the contact within its pair
};

// Warm-start record of one contact: the contact's identity (geom pair and
// index within the pair) and the impulses of its rows after the step.
struct WarmRec {
  unsigned long long key;
  float lam[3];
  int pad;
};

// identity of a contact across steps: geom ids (ascending) and index within the pair
OD_HD inline unsigned long long contact_key(int g1, int g2, int slot) {
  unsigned long long lo = (unsigned)(g1 < g2 ? g1 : g2), hi = (unsigned)(g1 < g2 ? g2 : g1);
  return (lo << 36) | (hi << 12) | (unsigned long long)(slot & 0xfff);
}

// one constraint row, 144 bytes.  J and iMJ are unscaled; the solver
// applies ad = w / (J inv(M) J^T + cfm) (ode/src/quickstep.cpp:370-402
// scales J and b by it instead).
struct Row {
  float J[12];             // l1 a1 l2 a2
  float iMJ[12];           // inv(M) J^T
  float rhs;               // c until the rhs pass, then the LCP right hand side
  float cfm;               // divided by h in the rhs pass
  float lo, hi;
  float ad;
  float lambda;
  int findex;              // absolute row index of the normal row, or -1
  int b1, b2;              // world-local bodies, b2 may be -1
  int pad[3];
};

// one constraint: a contact (up to 3 rows) or a permanent joint (up to 6)
struct Cons {
  int kind;                // JT_*
  int src;                 // contact index or joint index
  int b1, b2;
  int row0, m;
  int seq;                 // attach sequence (body joint list order)
  int fb;                  // feedback slot or -1
};

// Compact SOR row (64 bytes) used by the coloured solver.  J_l2 = -J_l1
// for every row the supported joints produce, and inv(M) J^T is recomputed
// from the per-body inverse mass and world inverse inertia, so a row needs
// only J_l1, J_a1, J_a2 and its scalars.
struct alignas(16) SRow {
  float l[3], a1[3], a2[3];
  float rhs, cfm, ad, lo, hi;
  float pad;
  int findex;   // row (within its constraint) of the normal impulse, or -1
};

// per-body block of the coloured solver: fc (6), symmetric world inverse
// inertia (xx xy xz yy yz zz), inverse mass, pad -- 64 bytes
struct alignas(16) SBody {
  float fc[6];
  float iI[6];
  float im;
  float pad[3];
};

// per-body state of the accelerated coloured solver: fc before the sweep and
// the search direction's fc (M^-1 J^T p)
struct alignas(16) SAccBody {
  float fp[6];
  float fdir[6];
};

// per-constraint record of the coloured solver, constraints sorted by colour
// and their rows stored contiguously in that order
struct alignas(16) SCons {
  int row0, m, b1, b2;
};

constexpr int WORK_QUEUES = 3;   // persistent-queue counters per batch, 16 ints apart

// phases timed by the optional in-kernel profiler (wall-clock ticks)
enum { PH_AABB = 0, PH_BROAD, PH_NARROW, PH_BODIES, PH_ROWS, PH_ORDER, PH_SOR, PH_INTEGRATE, PH_N };

struct ScratchLayout {
  long long aabb, gbig, gidx, glevel, gdb, cells, ecount, pairs, pkeys, pcount, contacts, cons, rows, btmp,
      order, adjs, adj, island, color, misc, sor, total;
};

OD_HD inline long long od_align(long long x) { return (x + 255) & ~255LL; }
OD_HD inline long long od_pow2(long long n) {
  long long p = 1;
  while (p < n) p <<= 1;
  return p;
}

// warm-start records of one world: kcap contact records, then 8 floats per
// permanent joint (impulses of up to 6 rows, then the row count as int bits)
OD_HD inline long long warm_bytes(int kcap, int nj) { return od_align(24LL * kcap) + 32LL * nj; }

// floats of the exact-LCP workspace of one world for islands of up to m
// rows: the matrix and its factor (m*m each), then seven vectors of m
OD_HD inline long long lcp_ws_floats(int m) { return 2LL * m * m + 7LL * m + 64; }

// bytes of the coloured solver's working set held in LDS when its rows and
// all of the conjugate-gradient state stay in global memory: bodies,
// constraints in colour order, colour starts and row impulses
OD_HD inline long long sor_lds_bytes_rows_cg_global(int nb, int nrows, int ncons) {
  return (long long)sizeof(SBody) * nb + (long long)sizeof(SCons) * ncons +
         ((4LL * (2LL * ncons + 8) + 15) & ~15LL) + ((4LL * nrows + 15) & ~15LL);
}

// ... with the rows' conjugate-gradient state in LDS too (the rows and the
// bodies' conjugate-gradient state in global memory)
OD_HD inline long long sor_lds_bytes_rows_acc_global(int nb, int nrows, int ncons, bool accel = false) {
  return sor_lds_bytes_rows_cg_global(nb, nrows, ncons) + (accel ? (8LL * nrows + 15) & ~15LL : 0LL);
}

// ... with the bodies' conjugate-gradient state in LDS too (the rows in
// global memory)
OD_HD inline long long sor_lds_bytes_rows_global(int nb, int nrows, int ncons, bool accel = false) {
  return sor_lds_bytes_rows_acc_global(nb, nrows, ncons, accel) + (accel ? (long long)sizeof(SAccBody) * nb : 0LL);
}

// the whole working set (rows first); copied to the LDS pool for the sweeps
// when it fits
OD_HD inline long long sor_lds_bytes(int nb, int nrows, int ncons, bool accel = false) {
  return (long long)sizeof(SRow) * nrows + sor_lds_bytes_rows_global(nb, nrows, ncons, accel);
}

// the same working set in the world's scratch
OD_HD inline long long sor_bytes(int nb, int nrows, int ncons, bool accel = false) {
  return sor_lds_bytes(nb, nrows, ncons, accel);
}

// Per-world scratch layout.  nb/ng/nj: world sizes; pcap (a power of two),
// kcap, rcap: pair, contact and row capacities.  Every array the kernels
// index is sized here, including the power-of-two padding of the sorts.
OD_HD inline ScratchLayout scratch_layout(int nb, int ng, int nj, int pcap, int kcap, int rcap) {
  ScratchLayout L;
  long long o = 0;
  long long ncell = od_pow2(8LL * ng + 8);
  long long ncons = (long long)kcap + nj + 1;
  L.aabb = o;     o = od_align(o + 24LL * ng);
  L.gbig = o;     o = od_align(o + 4LL * ng);
  L.gidx = o;     o = od_align(o + 4LL * ng);
  L.glevel = o;   o = od_align(o + 4LL * ng);
  L.gdb = o;      o = od_align(o + 24LL * ng);
  L.cells = o;    o = od_align(o + 8LL * ncell);
  L.ecount = o;   o = od_align(o + 4LL * (ncell + 1));
  L.pairs = o;    o = od_align(o + (long long)sizeof(PairRec) * pcap);
  L.pkeys = o;    o = od_align(o + 8LL * pcap);
  L.pcount = o;   o = od_align(o + 4LL * (pcap + 1));
  L.contacts = o; o = od_align(o + (long long)sizeof(ContactRec) * kcap);
  L.cons = o;     o = od_align(o + (long long)sizeof(Cons) * ncons);
  L.rows = o;     o = od_align(o + (long long)sizeof(Row) * rcap);
  L.btmp = o;     o = od_align(o + 4LL * 32 * nb);
  L.order = o;    o = od_align(o + 4LL * (rcap + ncons + nb + 2LL * ng + 8));
  L.adjs = o;     o = od_align(o + 4LL * (nb + 1));
  L.adj = o;      o = od_align(o + 4LL * 2 * ncons);
  L.island = o;   o = od_align(o + 4LL * (6LL * nb + 2 * ncons + 3LL * ng + 16));
  L.color = o;    o = od_align(o + 4LL * (ncons + nb));
  L.misc = o;     o = od_align(o + 4LL * 64);
  L.sor = o;      o = od_align(o + sor_bytes(nb, rcap, (int)ncons, true));
  L.total = o;
  return L;
}

// All device pointers the kernels need.
struct DevBatch {
  int nworlds;
  int order_mode;
  int track_residual;
  const WorldInfo *wi;
  WorldParams *wp;
  // bodies: pos, q, lvel, avel as float4 each; R as 3 float4 rows
  float4 *pos, *q, *lvel, *avel, *facc, *tacc;
  float4 *R;
  float4 *I, *invI;        // body frame, 3 rows of float4
  float *mass, *invMass;
  // geoms
  int *gtype, *gbody;
  float4 *gparam, *gpos, *gR;
  int *grank;              // position of the geom in its space list
  unsigned int *gcat, *gcol;
  // permanent joints
  JointDev *joints;
  const int *adj_start, *adj_other;   // per-world body -> bodies joined by a permanent joint
  const int *bgeom_start, *bgeom;     // global body -> its geoms (world-local ids, newest first)
  // host-created contacts (ODE API path)
  const ContactIn *cin;
  const int *cin_start, *cin_count;   // per world
  // per-step joint controls (lostop, histop, torque1, torque2), may be null
  const float4 *ctrl;
  // collide-phase outputs for the host (ODE API path)
  PairRec *pairs_out;
  ContactRec *contacts_out;
  int *pcount_out;         // contacts per pair (pairs_out order)
  // feedback out: 12 floats per slot (f1 t1 f2 t2)
  float *feedback;
  // counters and diagnostics
  int *counters;           // CNT_N per world
  float *residual;         // per world
  int *work_counter;       // persistent-kernel world queue
  unsigned long long *prof; // optional: PH_N wall-clock tick counters per world
  int lds_bytes;           // dynamic LDS pool per workgroup
  int sor_lds_bytes;       // LDS pool of the separate sweep kernel (split coloured solver)
  const int *world_order;  // processing order of worlds (largest first)
  char *scratch;
  // warm start: per-world records (WorldInfo::warm) and the number of valid
  // contact records per world (-1: none, the next solve starts from zero)
  char *warm;
  int *warm_n;
  // tape for the adjoint (null when not recording)
  char *tape;
  const long long *tape_off;   // per world byte offset into the tape
  // exact LCP (dWorldStep): per world lcp_ws_stride floats of workspace,
  // for islands of at most lcp_ws_rows rows (null when not stepping exactly)
  float *lcp_ws;
  long long lcp_ws_stride;
  int lcp_ws_rows;
};

}  // namespace odemi
