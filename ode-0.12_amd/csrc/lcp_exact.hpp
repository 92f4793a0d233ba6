// lcp_exact.hpp -- the LCP of one island solved exactly on the device
// (dWorldStep): A = J M^-1 J^T + cfm/h, b = rhs, bounds lo/hi, friction rows
// bounded by their normal row.  The answer is the one dSolveLCP defines
// (lcp.cpp:784-1100): the LCP of the rows without a findex with the friction
// impulses at zero, then every friction row's bounds fixed at
// +-|hi * lambda_normal| (0 when the normal impulse is 0) and the LCP of all
// rows solved with those bounds.  With cfm > 0 each of the two
// box-constrained LCPs has one solution; it is found by block principal
// pivoting (Judice & Pires) with a single-index fallback (Murty), on a
// Cholesky factorisation of the free rows computed by the whole workgroup.
//
// Included by step_kernels.hip inside namespace odemi, after its workgroup
// helpers (Shared, block_sum, block_max).
#pragma once

// Workspace of one world (floats, lcp_ws_floats(M)): A[M*M], L[M*M], then
// x, r, lo, hi, b (M each) and the int arrays st, fr (M each).
struct LcpWs {
  float *A, *L, *x, *r, *lo, *hi, *b;
  int *st, *fr;
  __device__ LcpWs(float *base, int M) {
    A = base;
    L = A + (size_t)M * M;
    x = L + (size_t)M * M;
    r = x + M;
    lo = r + M;
    hi = lo + M;
    b = hi + M;
    st = (int *)(b + M);
    fr = st + M;
  }
};

// st: 0 free, 1 at lo, 2 at hi, 3 outside the problem (held at 0)
enum { LS_FREE = 0, LS_LO = 1, LS_HI = 2, LS_OUT = 3 };

// x_F solving A_FF x_F = b_F - A_F,fixed x_fixed for the n free rows listed
// in W.fr; false when A_FF is not positive definite.  All lanes call it.
__device__ inline bool lcp_free_solve(const LcpWs &W, const float *b, int m, int n, Shared &sh) {
  const int nt = blockDim.x, t = threadIdx.x;
  float *L = W.L;
  for (int a = t; a < n; a += nt) {
    const int i = W.fr[a];
    float v = b[i];
    for (int j = 0; j < m; j++)
      if (W.st[j] == LS_LO || W.st[j] == LS_HI) v -= W.A[(size_t)i * m + j] * W.x[j];
    W.r[a] = v;
  }
  for (int e = t; e < n * n; e += nt) {
    const int a = e / n, c = e - a * n;
    if (c <= a) L[(size_t)a * n + c] = W.A[(size_t)W.fr[a] * m + W.fr[c]];
  }
  __syncthreads();
  // right-looking Cholesky, one column per round
  for (int k = 0; k < n; k++) {
    if (t == 0) {
      const float d = L[(size_t)k * n + k];
      sh.ival[14] = d > 0.f ? 1 : 0;
      L[(size_t)k * n + k] = d > 0.f ? sqrtf(d) : 1.f;
    }
    __syncthreads();
    if (!sh.ival[14]) return false;
    const float dk = L[(size_t)k * n + k];
    for (int i = k + 1 + t; i < n; i += nt) L[(size_t)i * n + k] /= dk;
    __syncthreads();
    const int q = n - k - 1;   // trailing triangle rows k+1..n-1
    for (int e = t; e < q * q; e += nt) {
      const int i = k + 1 + e / q, j = k + 1 + e % q;
      if (j <= i) L[(size_t)i * n + j] -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
    }
    __syncthreads();
  }
  // L y = r, then L^T x = y
  for (int k = 0; k < n; k++) {
    if (t == 0) W.r[k] /= L[(size_t)k * n + k];
    __syncthreads();
    for (int i = k + 1 + t; i < n; i += nt) W.r[i] -= L[(size_t)i * n + k] * W.r[k];
    __syncthreads();
  }
  for (int k = n - 1; k >= 0; k--) {
    if (t == 0) W.r[k] /= L[(size_t)k * n + k];
    __syncthreads();
    for (int i = t; i < k; i += nt) W.r[i] -= L[(size_t)k * n + i] * W.r[k];
    __syncthreads();
  }
  for (int a = t; a < n; a += nt) W.x[W.fr[a]] = W.r[a];
  __syncthreads();
  return true;
}

// The box-constrained LCP over the rows with st != LS_OUT (their start
// partition in st).  Returns the pivoting rounds, or -1 at the cap or on a
// matrix that is not positive definite.  All lanes call it.
__device__ inline int lcp_blcp(const LcpWs &W, const float *b, int m, float bscale, Shared &sh) {
  const int nt = blockDim.x, t = threadIdx.x;
  int best = m + 1, backup = 3;
  const float tw = 1e-5f * (1.f + bscale);
  for (int round = 0; round < 8 * m + 16; round++) {
    // fixed rows at their bound, rows outside at 0; the free list
    for (int i = t; i < m; i += nt) {
      const int s = W.st[i];
      if (s == LS_LO) W.x[i] = W.lo[i];
      else if (s == LS_HI) W.x[i] = W.hi[i];
      else if (s == LS_OUT) W.x[i] = 0.f;
    }
    if (t == 0) {
      int n = 0;
      for (int i = 0; i < m; i++)
        if (W.st[i] == LS_FREE) W.fr[n++] = i;
      sh.ival[13] = n;
    }
    __syncthreads();
    const int n = sh.ival[13];
    if (!lcp_free_solve(W, b, m, n, sh)) return -1;
    // infeasible rows: free ones outside their bounds, bound ones whose w
    // pushes away from the bound
    int nbad = 0, last = -1;
    for (int i = t; i < m; i += nt) {
      const int s = W.st[i];
      bool bad = false;
      if (s == LS_FREE) {
        const float tx = 1e-6f * (1.f + fabsf(W.x[i]));
        bad = W.x[i] < W.lo[i] - tx || W.x[i] > W.hi[i] + tx;
      } else if (s != LS_OUT) {
        float w = -b[i];
        for (int j = 0; j < m; j++)
          if (W.st[j] != LS_OUT) w += W.A[(size_t)i * m + j] * W.x[j];
        bad = (s == LS_LO && w < -tw) || (s == LS_HI && w > tw);
      }
      if (bad) { nbad++; last = i > last ? i : last; }
      W.r[i] = bad ? 1.f : 0.f;   // the free solve's scratch is free again
    }
    const int cnt = block_sum(nbad, sh);
    if (cnt == 0) return round;
    const int lastbad = (int)block_max((float)last, sh);
    // all infeasible rows change sides while their number falls (and for a
    // few rounds after), else only the last one
    bool all;
    if (cnt < best) { best = cnt; backup = 3; all = true; }
    else if (backup > 0) { backup--; all = true; }
    else all = false;
    for (int i = t; i < m; i += nt) {
      if (W.r[i] == 0.f || (!all && i != lastbad)) continue;
      const int s = W.st[i];
      if (s == LS_FREE) W.st[i] = W.x[i] < W.lo[i] ? LS_LO : LS_HI;
      else W.st[i] = LS_FREE;
    }
    __syncthreads();
  }
  return -1;
}

// Every island of the world (isl: [cons0, ncons, row0, nrows, seed] per
// island) solved exactly; lambda into the rows, fc = M^-1 J^T lambda per body
// into bt.  An island larger than the workspace, or a solve that does not
// finish, sets OVF_LCP; its impulses are then the clamped last iterate.
__device__ void lcp_exact_world(const DevBatch &B, int w, Row *rows, float *bt, const int *isl, int ngroups, int nb,
                                Shared &sh) {
  const int nt = blockDim.x, t = threadIdx.x, M = B.lcp_ws_rows;
  const LcpWs W(B.lcp_ws + (size_t)w * B.lcp_ws_stride, M);
  bool fail = false;
  for (int k = 0; k < ngroups; k++) {
    const int r0 = isl[5 * k + 2], m = isl[5 * k + 3];
    if (m <= 0) continue;
    if (m > M) {
      for (int i = t; i < m; i += nt) rows[r0 + i].lambda = 0.f;
      fail = true;
      continue;
    }
    // A_ij = J_i . M^-1 J_j^T over the bodies the rows share, + cfm/h
    for (int e = t; e < m * m; e += nt) {
      const int i = e / m, j = e - i * m;
      const Row &Ri = rows[r0 + i], &Rj = rows[r0 + j];
      float s = 0.f;
      for (int si = 0; si < 2; si++) {
        const int bi = si ? Ri.b2 : Ri.b1;
        if (bi < 0) continue;
        for (int sj = 0; sj < 2; sj++)
          if ((sj ? Rj.b2 : Rj.b1) == bi)
            for (int q = 0; q < 6; q++) s += Ri.J[6 * si + q] * Rj.iMJ[6 * sj + q];
      }
      W.A[(size_t)i * m + j] = s + (i == j ? Ri.cfm : 0.f);
    }
    float bmax = 0.f;
    bool fric = false;
    for (int i = t; i < m; i += nt) {
      const Row &R = rows[r0 + i];
      W.lo[i] = R.lo;
      W.hi[i] = R.hi;
      // rows held at lo = hi (= 0) and, first, the friction rows stay out
      W.st[i] = (R.findex >= 0 || R.lo == R.hi) ? LS_OUT : LS_FREE;
      W.x[i] = 0.f;
      bmax = fmaxf(bmax, fabsf(R.rhs));
      fric |= R.findex >= 0;
    }
    bmax = block_max(bmax, sh);
    const int any_fric = block_sum(fric ? 1 : 0, sh);
    for (int i = t; i < m; i += nt) W.b[i] = rows[r0 + i].rhs;
    __syncthreads();
    int rounds = lcp_blcp(W, W.b, m, bmax, sh);
    if (rounds >= 0 && any_fric) {
      for (int i = t; i < m; i += nt) {
        const Row &R = rows[r0 + i];
        if (R.findex < 0) continue;
        const float xn = W.x[R.findex - r0];
        if (xn == 0.f) {
          W.lo[i] = W.hi[i] = 0.f;
          W.st[i] = LS_OUT;
        } else {
          W.hi[i] = fabsf(R.hi * xn);
          W.lo[i] = -W.hi[i];
          W.st[i] = LS_FREE;
        }
      }
      __syncthreads();
      rounds = lcp_blcp(W, W.b, m, bmax, sh);
    }
    if (rounds < 0) fail = true;
    for (int i = t; i < m; i += nt) rows[r0 + i].lambda = fminf(fmaxf(W.x[i], W.lo[i]), W.hi[i]);
    __syncthreads();
  }
  // fc of every body from the rows it is in
  int nrows_all = 0;
  for (int k = 0; k < ngroups; k++) nrows_all = max(nrows_all, isl[5 * k + 2] + isl[5 * k + 3]);
  for (int b = t; b < nb; b += nt) {
    float f[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int r = 0; r < nrows_all; r++) {
      const Row &R = rows[r];
      if (R.b1 == b)
        for (int q = 0; q < 6; q++) f[q] += R.iMJ[q] * R.lambda;
      else if (R.b2 == b)
        for (int q = 0; q < 6; q++) f[q] += R.iMJ[6 + q] * R.lambda;
    }
    for (int q = 0; q < 6; q++) bt[32 * b + 9 + q] = f[q];
  }
  if (fail && t == 0) atomicOr(&B.counters[w * CNT_N + CNT_OVERFLOW], OVF_LCP);
  __syncthreads();
}
