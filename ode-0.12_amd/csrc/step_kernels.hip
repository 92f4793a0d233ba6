// step_kernels.hip -- the ODE 0.12 step path on MI355X, one workgroup per
// world.
//
//   collide:  geom AABBs -> uniform-grid broadphase (cell keys, bitonic sort,
//             run scan with a home-cell rule) -> narrowphase, one pair per lane
//   step:     body preprocessing -> constraint rows (contact, ball, hinge,
//             hinge2) -> SOR-LCP -> velocity update -> integration
//
// The SOR-LCP runs in one of two orders:
//   ORDER_REFERENCE  islands found by the reference's depth-first walk, rows in
//                    island/joint order, non-friction rows first and friction
//                    rows reversed at the tail, shuffled every 8 iterations by
//                    the reference's LCG (each island's RNG state is obtained by
//                    skip-ahead so islands relax in parallel); candidate pairs
//                    are sorted into the reference hash space's callback order.
//   ORDER_COLORED    constraints coloured so that one colour shares no body;
//                    a colour's constraints relax in parallel, one per lane.
//
// Reference: ode/src/collision_space.cpp:420-583 (hash space),
// ode/src/quickstep.cpp:342-1025 (SOR-LCP, dxQuickStepper),
// ode/src/util.cpp:255-568 (integrator, islands).
#include <hip/hip_runtime.h>

#include "colliders.hpp"
#include "devmath.hpp"
#include "joints_dev.hpp"
#include "layout.hpp"

namespace odemi {

constexpr int WG = 256;
constexpr int MAX_COLORS_LDS = 256;   // colour buckets counted in LDS

struct Shared {
  int wsum[WG / 64];
  float wmax[WG / 64];
  int ival[16];
  float fval[8];
  int ccount[MAX_COLORS_LDS + 1];
  int ccur[MAX_COLORS_LDS];
};

// dynamic LDS pool of the workgroup (size DevBatch::lds_bytes): sort keys,
// colouring state and the coloured solver's working set, one after another
extern __shared__ __align__(16) unsigned char od_pool[];

// optional per-phase wall-clock ticks (thread 0)
struct PhaseClock {
  unsigned long long *dst;
  unsigned long long t;
  __device__ void start(const DevBatch &B, int w) {
    dst = B.prof ? B.prof + (size_t)w * PH_N : nullptr;
    if (dst && threadIdx.x == 0) t = wall_clock64();
  }
  __device__ void lap(int ph) {
    if (dst && threadIdx.x == 0) {
      unsigned long long n = wall_clock64();
      dst[ph] += n - t;
      t = n;
    }
  }
};

// ------------------------------------------------------------------ block helpers

__device__ inline int block_excl_scan(int v, int &total, Shared &sh) {
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
  for (int o = 1; o < 64; o <<= 1) {
    int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) sh.wsum[wid] = x;
  __syncthreads();
  int nw = (blockDim.x + 63) >> 6, off = 0, tot = 0;
  for (int k = 0; k < nw; k++) {
    int s = sh.wsum[k];
    if (k < wid) off += s;
    tot += s;
  }
  __syncthreads();
  total = tot;
  return off + x - v;
}

// colouring: the lowest colour below 64 that neither body's mask (body temp
// words 24-25) holds and that is not full; 64 if none
__device__ inline int first_free_colour(const float *bt, int b1, int b2, unsigned long long full) {
  const unsigned long long used = *(const unsigned long long *)(bt + 32 * b1 + 24) |
                                  (b2 >= 0 ? *(const unsigned long long *)(bt + 32 * b2 + 24) : 0ull) | full;
  return used == ~0ull ? 64 : __ffsll((long long)~used) - 1;
}

__device__ inline void take_colour(float *bt, int b1, int b2, int *col, int c) {
  *col = c;
  *(unsigned long long *)(bt + 32 * b1 + 24) |= 1ull << c;
  if (b2 >= 0) *(unsigned long long *)(bt + 32 * b2 + 24) |= 1ull << c;
}

__device__ inline float block_max(float v, Shared &sh) {
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  if (lane == 0) sh.wmax[wid] = v;
  __syncthreads();
  int nw = (blockDim.x + 63) >> 6;
  float m = sh.wmax[0];
  for (int k = 1; k < nw; k++) m = fmaxf(m, sh.wmax[k]);
  __syncthreads();
  return m;
}

__device__ inline float block_fsum(float v, Shared &sh) {
  int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) sh.wmax[wid] = v;
  __syncthreads();
  int nw = (blockDim.x + 63) >> 6;
  float m = 0.f;
  for (int k = 0; k < nw; k++) m += sh.wmax[k];
  __syncthreads();
  return m;
}

__device__ inline int block_sum(int v, Shared &sh) {
  int t;
  block_excl_scan(v, t, sh);
  return t;
}

// ascending bitonic sort of n (power of two) u64 keys; in LDS when it fits
__device__ inline void bitonic_sort(unsigned long long *a, int n, Shared &sh, int lds_bytes) {
  bool lds = (long long)n * 8 <= lds_bytes;
  unsigned long long *pool = (unsigned long long *)od_pool;
  unsigned long long *k = lds ? pool : a;
  if (lds) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) pool[i] = a[i];
    __syncthreads();
  }
  for (int size = 2; size <= n; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int i = threadIdx.x; i < n; i += blockDim.x) {
        int j = i ^ stride;
        if (j > i) {
          bool asc = (i & size) == 0;
          unsigned long long x = k[i], y = k[j];
          if ((x > y) == asc) { k[i] = y; k[j] = x; }
        }
      }
      __syncthreads();
    }
  }
  if (lds) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = pool[i];
    __syncthreads();
  }
}

// LDS bytes the block radix sort needs for a workgroup of nt lanes: digit
// bases (256) and per-wavefront digit counts of one tile
__host__ __device__ inline int radix_lds_bytes(int nt) { return 4 * (256 + 256 * ((nt + 63) / 64)); }

// Workgroup barrier that orders LDS accesses only: global loads and stores
// in flight stay in flight (a full __syncthreads waits for them).  For
// phases whose cross-lane exchange is in LDS and whose global writes are
// read only after a later full barrier.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Ascending LSD radix sort of n u64 keys in global memory by one workgroup,
// by bits [lo_bit, 64) in 8-bit digits (lo_bit a multiple of 8); tmp holds n
// keys.  A digit that is zero in every key is skipped (a stable pass over
// one digit value is the identity).  Each pass counts digits (LDS atomics),
// scans the 256 counts (block scan) and scatters tile by tile, stably: the
// rank of a key among the equal digits of its wavefront comes from eight
// ballots, the wavefronts of a tile are ordered through per-wavefront counts
// in LDS.  Keys equal in the sorted bits keep their input order.  The result
// is in keys.
__device__ inline void block_radix_sort(unsigned long long *keys, unsigned long long *tmp, int n, int lo_bit,
                                        Shared &sh) {
  const int nt = blockDim.x, nw = (nt + 63) >> 6;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int *base = (int *)od_pool;      // 256 digit bases
  int *wc = base + 256;            // per wavefront: 256 digit counts of the current tile
  unsigned long long any = 0;
  for (int i = threadIdx.x; i < n; i += nt) any |= keys[i];
  // bits present in some key (OR over the workgroup, 32 bits at a time)
  unsigned lo = (unsigned)any, hi = (unsigned)(any >> 32);
  for (int o = 32; o > 0; o >>= 1) { lo |= __shfl_xor(lo, o, 64); hi |= __shfl_xor(hi, o, 64); }
  if (lane == 0) { sh.ival[0] = 0; sh.ival[1] = 0; }
  __syncthreads();
  if (lane == 0) { atomicOr(&sh.ival[0], (int)lo); atomicOr(&sh.ival[1], (int)hi); }
  __syncthreads();
  const unsigned long long all = ((unsigned long long)(unsigned)sh.ival[1] << 32) | (unsigned)sh.ival[0];
  __syncthreads();
  unsigned long long *src = keys, *dst = tmp;
  for (int shift = lo_bit; shift < 64; shift += 8) {
    if (((all >> shift) & 255) == 0) continue;
    for (int d = threadIdx.x; d < 256; d += nt) base[d] = 0;
    for (int d = threadIdx.x; d < 256 * nw; d += nt) wc[d] = 0;
    __syncthreads();
    // digit counts, four keys per lane in flight at a time
    for (int i0 = threadIdx.x; i0 < n; i0 += 4 * nt) {
      unsigned long long k4[4];
      for (int u = 0; u < 4; u++) k4[u] = (i0 + u * nt < n) ? src[i0 + u * nt] : 0ULL;
      for (int u = 0; u < 4; u++)
        if (i0 + u * nt < n) atomicAdd(&base[(int)((k4[u] >> shift) & 255)], 1);
    }
    __syncthreads();
    {
      // exclusive scan of the digit counts, a run of digits per lane
      const int per = (256 + nt - 1) / nt, d0 = threadIdx.x * per;
      const int d1 = d0 + per < 256 ? d0 + per : 256;
      int loc = 0;
      for (int d = d0; d < d1; d++) loc += base[d];
      int tot;
      int off = block_excl_scan(loc, tot, sh);
      for (int d = d0; d < d1; d++) { const int cnt = base[d]; base[d] = off; off += cnt; }
    }
    __syncthreads();
    // tiles of nt keys; the next tile's key is loaded before this tile's
    // LDS-only barriers, the scattered stores are waited for once per pass
    unsigned long long knext = threadIdx.x < n ? src[threadIdx.x] : 0ULL;
    for (int t0 = 0; t0 < n; t0 += nt) {
      const int i = t0 + threadIdx.x;
      const bool valid = i < n;
      const unsigned long long k = knext;
      knext = (i + nt < n) ? src[i + nt] : 0ULL;
      const int d = (int)((k >> shift) & 255);
      unsigned long long peers = __ballot(valid);
      for (int b = 0; b < 8; b++) {
        const unsigned long long m = __ballot(valid && ((d >> b) & 1));
        peers &= ((d >> b) & 1) ? m : ~m;
      }
      const unsigned long long lt = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
      const int rank = __popcll(peers & lt);
      const int cnt = __popcll(peers);
      if (valid && rank == cnt - 1) wc[wid * 256 + d] = cnt;
      lds_barrier();
      if (valid) {
        int off = base[d] + rank;
        for (int w = 0; w < wid; w++) off += wc[w * 256 + d];
        dst[off] = k;
      }
      lds_barrier();
      for (int dd = threadIdx.x; dd < 256; dd += nt) {
        int tot = 0;
        for (int w = 0; w < nw; w++) { tot += wc[w * 256 + dd]; wc[w * 256 + dd] = 0; }
        base[dd] += tot;
      }
      lds_barrier();
    }
    __syncthreads();   // this pass's stores before the next pass reads them
    unsigned long long *t = src; src = dst; dst = t;
  }
  if (src != keys) {
    for (int i = threadIdx.x; i < n; i += nt) keys[i] = src[i];
    __syncthreads();
  }
}

__device__ inline int pow2_at_least(int n) {
  int p = 1;
  while (p < n) p <<= 1;
  return p;
}

// ------------------------------------------------------------------ world context

struct Ctx {
  int w;
  WorldInfo W;
  ScratchLayout L;
  char *S;
  mutable PhaseClock clk;
  template <class T> __device__ T *at(long long off) const { return (T *)(S + off); }
};

__device__ inline GeomView load_geom(const DevBatch &B, const Ctx &c, int g) {
  int gg = c.W.g0 + g;
  GeomView v;
  v.type = B.gtype[gg];
  v.prm = B.gparam[gg];
  int b = B.gbody[gg];
  if (b >= 0) {
    int gb = c.W.b0 + b;
    v.p = xyz(B.pos[gb]);
    v.R = load_mat(B.R + 3 * (size_t)gb);
  } else {
    v.p = xyz(B.gpos[gg]);
    v.R = load_mat(B.gR + 3 * (size_t)gg);
  }
  return v;
}

__device__ inline bool aabb_overlap(const float *a, const float *b) {
  return !(a[0] > b[1] || a[1] < b[0] || a[2] > b[3] || a[3] < b[2] || a[4] > b[5] || a[5] < b[4]);
}

// collideAABBs filters (collision_space_internal.h:48-82)
__device__ inline bool pair_passes(const DevBatch &B, const Ctx &c, int g1, int g2, const float *aabb) {
  int b1 = B.gbody[c.W.g0 + g1], b2 = B.gbody[c.W.g0 + g2];
  if (b1 == b2 && b1 >= 0) return false;
  unsigned c1 = B.gcat[c.W.g0 + g1], c2 = B.gcat[c.W.g0 + g2];
  unsigned k1 = B.gcol[c.W.g0 + g1], k2 = B.gcol[c.W.g0 + g2];
  if (((c1 & k2) || (c2 & k1)) == 0) return false;
  return aabb_overlap(aabb + 6 * g1, aabb + 6 * g2);
}

// the reference's near-callback filter: skip bodies joined by a permanent joint
__device__ inline bool connected(const DevBatch &B, const Ctx &c, int b1, int b2) {
  if (b1 < 0 || b2 < 0 || !B.adj_start) return false;
  int s = B.adj_start[c.W.b0 + b1], e = B.adj_start[c.W.b0 + b1 + 1];
  for (int k = s; k < e; k++)
    if (B.adj_other[k] == b2) return true;
  return false;
}

__device__ inline unsigned cell_hash(int x, int y, int z) {
  return ((unsigned)x * 73856093u) ^ ((unsigned)y * 19349663u) ^ ((unsigned)z * 83492791u);
}

// ------------------------------------------------------------------ collide

// AABBs, big-geom flags, reference hash levels and discrete bounds
// (collision_space.cpp:331-351, 444-476).
__device__ void phase_aabb(const DevBatch &B, const Ctx &c, const WorldParams &P, Shared &sh) {
  float *aabb = c.at<float>(c.L.aabb);
  int *gbig = c.at<int>(c.L.gbig);
  int *glevel = c.at<int>(c.L.glevel);
  int *gdb = c.at<int>(c.L.gdb);
  for (int g = threadIdx.x; g < c.W.ng; g += blockDim.x) {
    GeomView v = load_geom(B, c, g);
    float a[6];
    geom_aabb(v, a);
    for (int k = 0; k < 6; k++) aabb[6 * g + k] = a[k];
    bool inf = false;
    for (int k = 0; k < 6; k++) inf |= isinf(a[k]);
    int lvl = 0x7fffffff;
    if (!inf) {
      float q = fmaxf(a[1] - a[0], fmaxf(a[3] - a[2], a[5] - a[4]));
      int e;
      frexpf(q, &e);
      lvl = e < P.hash_min ? P.hash_min : e;
    }
    bool big = inf || lvl > P.hash_max;
    gbig[g] = big ? 1 : 0;
    glevel[g] = lvl;
    if (!big) {
      float cell = ldexpf(1.f, lvl);
      for (int k = 0; k < 6; k++) gdb[6 * g + k] = (int)floorf(a[k] / cell);
    }
  }
  __syncthreads();
}

// reference hash index of every non-big geom (position in the space list
// among non-big geoms) and the big-geom list in reverse space order.
__device__ void phase_ref_index(const DevBatch &B, const Ctx &c, Shared &sh) {
  int *gbig = c.at<int>(c.L.gbig);
  int *gidx = c.at<int>(c.L.gidx);
  int *inv = c.at<int>(c.L.order);   // rank -> geom, then big list
  const int ng = c.W.ng;
  for (int g = threadIdx.x; g < ng; g += blockDim.x) inv[B.grank[c.W.g0 + g]] = g;
  __syncthreads();
  // in space-list order (block scans): non-big geoms get consecutive hash
  // indices; big geoms note their position among the big ones (-2 - pos)
  int n = 0, nbig = 0;
  for (int r0 = 0; r0 < ng; r0 += blockDim.x) {
    const int r = r0 + threadIdx.x;
    const int g = r < ng ? inv[r] : -1;
    const int isb = (g >= 0 && gbig[g]) ? 1 : 0;
    const int nonb = (g >= 0 && !isb) ? 1 : 0;
    int t1, t2;
    const int o1 = block_excl_scan(nonb, t1, sh);
    const int o2 = block_excl_scan(isb, t2, sh);
    if (nonb) gidx[g] = n + o1;
    if (isb) gidx[g] = -2 - (nbig + o2);
    n += t1;
    nbig += t2;
  }
  __syncthreads();
  // big geoms in reverse space order: big rank k = nbig - 1 - pos
  int *big = inv + ng;
  for (int g = threadIdx.x; g < ng; g += blockDim.x) {
    const int v = gidx[g];
    if (v <= -2) {
      const int k = nbig - 1 - (-2 - v);
      big[k] = g;
      gbig[g] = 1 + k;   // 1 + big rank
      gidx[g] = -1;
    }
  }
  if (threadIdx.x == 0) {
    sh.ival[0] = n;
    sh.ival[1] = nbig;
  }
  __syncthreads();
}

// key placing pair (a,b) of non-big geoms at the reference's first encounter
// in the hash-space traversal; also returns the emitted orientation.
__device__ inline unsigned long long ref_pair_key(const Ctx &c, int a, int b, int n, int &g1, int &g2) {
  const int *gidx = c.at<int>(c.L.gidx);
  const int *glevel = c.at<int>(c.L.glevel);
  const int *gdb = c.at<int>(c.L.gdb);
  int X = gidx[a] > gidx[b] ? a : b;    // larger index: earlier in the traversal
  int Y = X == a ? b : a;
  int outer, part;
  if (glevel[Y] >= glevel[X]) { outer = X; part = Y; }
  else { outer = Y; part = X; }
  int L = glevel[part];
  int sh = L - glevel[outer];
  int d[3];
  for (int k = 0; k < 3; k++) {
    int o0 = gdb[6 * outer + 2 * k] >> sh;
    int p0 = gdb[6 * part + 2 * k];
    int c0 = o0 > p0 ? o0 : p0;
    d[k] = c0 - o0;
  }
  g1 = outer;
  g2 = part;
  unsigned long long rank = (unsigned long long)(n - 1 - gidx[outer]);
  return (rank << 40) | ((unsigned long long)(sh & 31) << 35) | ((unsigned long long)(d[0] & 1) << 34) |
         ((unsigned long long)(d[1] & 1) << 33) | ((unsigned long long)(d[2] & 1) << 32) |
         (unsigned long long)gidx[part];
}

// Uniform-grid broadphase.  Each non-big geom enters the (at most 8) cells
// its AABB covers; entries are sorted by cell key; a pair is reported from the
// run of its home cell (the low corner of the two cell ranges' overlap), so
// every overlapping pair appears exactly once.  Big geoms (unbounded AABBs)
// are tested against everything.  Returns the pair count.
__device__ int phase_broadphase(const DevBatch &B, const Ctx &c, const WorldParams &P, Shared &sh, bool ref) {
  const float *aabb = c.at<float>(c.L.aabb);
  const int *gbig = c.at<int>(c.L.gbig);
  unsigned long long *cells = c.at<unsigned long long>(c.L.cells);
  int *ecount = c.at<int>(c.L.ecount);
  PairRec *pairs = c.at<PairRec>(c.L.pairs);
  unsigned long long *pkeys = c.at<unsigned long long>(c.L.pkeys);
  const int ng = c.W.ng;
  int nref = 0;
  if (ref) {
    phase_ref_index(B, c, sh);
    nref = sh.ival[0];
  }
  // cell size from the largest extent
  float ext = 0.f;
  for (int g = threadIdx.x; g < ng; g += blockDim.x) {
    if (gbig[g]) continue;
    const float *a = aabb + 6 * g;
    ext = fmaxf(ext, fmaxf(a[1] - a[0], fmaxf(a[3] - a[2], a[5] - a[4])));
  }
  ext = block_max(ext, sh);
  float cell = fmaxf(ext * P.cell_scale * 1.0001f, 1e-6f);
  float inv = 1.f / cell;
  // cell entries: key = hash << 32 | geom << 3 | subcell
  int base = 0;
  for (int start = 0; start < ng; start += blockDim.x) {
    int g = start + threadIdx.x;
    int cnt = 0, lo[3] = {0, 0, 0}, span[3] = {1, 1, 1};
    if (g < ng && !gbig[g]) {
      const float *a = aabb + 6 * g;
      for (int k = 0; k < 3; k++) {
        lo[k] = (int)floorf(a[2 * k] * inv);
        int hi = (int)floorf(a[2 * k + 1] * inv);
        span[k] = hi > lo[k] ? 2 : 1;
      }
      cnt = span[0] * span[1] * span[2];
    }
    int tot;
    int off = block_excl_scan(cnt, tot, sh);
    if (cnt) {
      int e = base + off;
      for (int dz = 0; dz < span[2]; dz++)
        for (int dy = 0; dy < span[1]; dy++)
          for (int dx = 0; dx < span[0]; dx++) {
            unsigned h = cell_hash(lo[0] + dx, lo[1] + dy, lo[2] + dz);
            cells[e++] = ((unsigned long long)h << 32) | ((unsigned long long)g << 3) | (unsigned)(dx | dy << 1 | dz << 2);
          }
    }
    base += tot;
  }
  const int ne = base;
  const int np2 = pow2_at_least(ne < 2 ? 2 : ne);
  if ((long long)np2 * 8 > B.lds_bytes && B.lds_bytes >= radix_lds_bytes(blockDim.x) && ne <= c.W.pcap) {
    // too many keys for the LDS pool: radix sort in global memory (pkeys,
    // pcap entries written only by the pair scan below, holds the second
    // buffer), by the cell hash only: the pair scan needs the entries of a
    // cell together, and the pairs it finds are sorted by their own keys
    // below, so the order inside a cell does not reach the result
    __syncthreads();
    block_radix_sort(cells, pkeys, ne, 32, sh);
  } else {
    for (int i = ne + threadIdx.x; i < np2; i += blockDim.x) cells[i] = ~0ULL;
    __syncthreads();
    bitonic_sort(cells, np2, sh, B.lds_bytes);
  }
  // pair scan: one lane per cell entry walks the rest of its cell's run.
  // A pair is found once, from the entry of the pair's home cell (the low
  // corner of the two cell ranges' overlap) that comes first in the run.
  auto pair_key = [&](int gi, int gj) -> unsigned long long {
    if (ref) {
      int a, b;
      return ref_pair_key(c, gi, gj, nref, a, b);
    }
    return gi < gj ? ((unsigned long long)gi << 32) | (unsigned)gj : ((unsigned long long)gj << 32) | (unsigned)gi;
  };
  auto visit = [&](int i, auto &&emit) {
    const unsigned long long ei = cells[i];
    const unsigned hi = (unsigned)(ei >> 32);
    const int gi = (int)((ei & 0xffffffffu) >> 3), si = (int)(ei & 7);
    const int G = c.W.g0;
    float ai[6];
    for (int k = 0; k < 6; k++) ai[k] = aabb[6 * gi + k];
    const int bi = B.gbody[G + gi];
    const unsigned ci = B.gcat[G + gi], ki = B.gcol[G + gi];
    int li[3];
    for (int k = 0; k < 3; k++) li[k] = (int)floorf(ai[2 * k] * inv) + ((si >> k) & 1);
    for (int j = i + 1; j < ne; j++) {
      const unsigned long long ej = cells[j];
      if ((unsigned)(ej >> 32) != hi) break;
      const int gj = (int)((ej & 0xffffffffu) >> 3), sj = (int)(ej & 7);
      if (gj == gi) continue;
      float aj[6];
      for (int k = 0; k < 6; k++) aj[k] = aabb[6 * gj + k];
      const int bj = B.gbody[G + gj];
      const unsigned cj = B.gcat[G + gj], kj = B.gcol[G + gj];
      bool same = true, home = true;
      for (int k = 0; k < 3; k++) {
        int ljk0 = (int)floorf(aj[2 * k] * inv);
        int lj = ljk0 + ((sj >> k) & 1);
        int lik0 = li[k] - ((si >> k) & 1);
        int hc = lik0 > ljk0 ? lik0 : ljk0;
        same &= lj == li[k];
        home &= hc == li[k];
      }
      if (!same || !home) continue;
      // collideAABBs filters (collision_space_internal.h:54-71)
      if (bi == bj && bi >= 0) continue;
      if (((ci & kj) || (cj & ki)) == 0) continue;
      if (!aabb_overlap(ai, aj)) continue;
      emit(gi, gj);
    }
  };
  // every pair takes the next slot of a workgroup counter: the keys are
  // sorted below, so the slot order does not reach the result
  if (threadIdx.x == 0) sh.ival[2] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < ne; i += blockDim.x)
    visit(i, [&](int gi, int gj) {
      const int o = atomicAdd(&sh.ival[2], 1);
      if (o < c.W.pcap) pkeys[o] = pair_key(gi, gj);
    });
  __syncthreads();
  int npairs = sh.ival[2];
  __syncthreads();
  if (npairs > c.W.pcap) {
    // more pairs than slots: scan again in entry order (counting pass, then
    // writing pass) so that the pairs kept are the same on every run
    npairs = 0;
    for (int start = 0; start < ne; start += blockDim.x) {
      const int i = start + threadIdx.x;
      int cnt = 0;
      if (i < ne) visit(i, [&](int, int) { cnt++; });
      int tot;
      const int off = block_excl_scan(cnt, tot, sh);
      if (i < ne && cnt) {
        int o = npairs + off;
        visit(i, [&](int gi, int gj) {
          if (o < c.W.pcap) pkeys[o] = pair_key(gi, gj);
          o++;
        });
      }
      npairs += tot;
    }
  }
  // big geoms against the rest, then big against big; the big geoms first
  // listed in ascending index order (ecount is free scratch here)
  int *blist = ecount;
  int nbig = 0;
  for (int start = 0; start < ng; start += blockDim.x) {
    int g = start + threadIdx.x;
    int isb = (g < ng && gbig[g]) ? 1 : 0;
    int tot;
    int off = block_excl_scan(isb, tot, sh);
    if (isb) blist[nbig + off] = g;
    nbig += tot;
  }
  __syncthreads();
  for (int start = 0; start < ng; start += blockDim.x) {
    int g = start + threadIdx.x;
    int cnt = 0;
    if (g < ng)
      for (int t = 0; t < nbig; t++) {
        const int b = blist[t];
        if (b == g) continue;
        if (gbig[g] && !(b > g)) continue;  // big-big once
        if (pair_passes(B, c, g, b, aabb)) cnt++;
      }
    int tot;
    int off = block_excl_scan(cnt, tot, sh);
    if (cnt) {
      int o = npairs + off;
      for (int t = 0; t < nbig; t++) {
        const int b = blist[t];
        if (b == g) continue;
        if (gbig[g] && !(b > g)) continue;
        if (!pair_passes(B, c, g, b, aabb)) continue;
        if (o < c.W.pcap) {
          if (ref) {
            const int *gidx = c.at<int>(c.L.gidx);
            int bb = gbig[b] - 1;
            if (!gbig[g]) {
              pkeys[o] = (1ULL << 63) | ((unsigned long long)(nref - 1 - gidx[g]) << 32) | (unsigned)bb;
            } else {
              int ga = gbig[g] - 1;
              int r1 = ga < bb ? ga : bb, r2 = ga < bb ? bb : ga;
              pkeys[o] = (3ULL << 62) | ((unsigned long long)r1 << 24) | (unsigned)r2;
            }
          } else {
            pkeys[o] = g < b ? ((unsigned long long)g << 32) | (unsigned)b
                             : ((unsigned long long)b << 32) | (unsigned)g;
          }
        }
        o++;
      }
    }
    npairs += tot;
  }
  if (npairs > c.W.pcap) {
    if (threadIdx.x == 0) atomicOr(&B.counters[c.w * CNT_N + CNT_OVERFLOW], OVF_PAIRS);
    npairs = c.W.pcap;
  }
  {
    // sort the keys: in reference mode into the callback order, in coloured
    // mode into (g1, g2) order with g1 < g2, so that the contact order (and
    // with it the colouring) does not depend on the grid
    int np = pow2_at_least(npairs < 2 ? 2 : npairs);
    if (np > c.W.pcap) np = c.W.pcap;  // pcap is a power of two
    if ((long long)np * 8 > B.lds_bytes && B.lds_bytes >= radix_lds_bytes(blockDim.x)) {
      // radix sort in global memory; the pair records (written after the
      // sort) hold the second buffer
      __syncthreads();
      block_radix_sort(pkeys, (unsigned long long *)pairs, npairs, 0, sh);
    } else {
      for (int i = npairs + threadIdx.x; i < np; i += blockDim.x) pkeys[i] = ~0ULL;
      __syncthreads();
      bitonic_sort(pkeys, np, sh, B.lds_bytes);
    }
  }
  if (!ref) {
    for (int i = threadIdx.x; i < npairs; i += blockDim.x) {
      unsigned long long k = pkeys[i];
      pairs[i].g1 = (int)(k >> 32);
      pairs[i].g2 = (int)(k & 0xffffffffu);
      pairs[i].key = k;
    }
  } else {
    const int *gidx = c.at<int>(c.L.gidx);
    // inverse maps: hash index -> geom, big rank -> geom
    int *inv = c.at<int>(c.L.order);
    int *byidx = c.at<int>(c.L.island);
    for (int g = threadIdx.x; g < ng; g += blockDim.x)
      if (gidx[g] >= 0) byidx[gidx[g]] = g;
    __syncthreads();
    const int *big = inv + ng;
    for (int i = threadIdx.x; i < npairs; i += blockDim.x) {
      unsigned long long k = pkeys[i];
      int a, b;
      if ((k >> 62) == 3ULL) { a = big[(k >> 24) & 0xffffff]; b = big[k & 0xffffff]; }
      else if (k >> 63) { a = byidx[nref - 1 - (int)((k >> 32) & 0x3fffffff)]; b = big[k & 0xffffffff]; }
      else { a = byidx[nref - 1 - (int)(k >> 40)]; b = byidx[k & 0xffffffff]; }
      pairs[i].g1 = a;
      pairs[i].g2 = b;
      pairs[i].key = k;
    }
  }
  __syncthreads();
  return npairs;
}

__device__ inline void store_contacts(const DevBatch &B, const Ctx &c, int g1, int g2, const CPoint *pts, int n,
                                      ContactRec *dst, int room) {
  int b1 = B.gbody[c.W.g0 + g1], b2 = B.gbody[c.W.g0 + g2];
  int flags = 0;
  if (b1 < 0) { b1 = b2; b2 = -1; flags = JF_REVERSE; }
  for (int i = 0; i < n && i < room; i++) {
    ContactRec &r = dst[i];
    r.pos[0] = pts[i].pos.x; r.pos[1] = pts[i].pos.y; r.pos[2] = pts[i].pos.z;
    r.depth = pts[i].depth;
    r.normal[0] = pts[i].n.x; r.normal[1] = pts[i].n.y; r.normal[2] = pts[i].n.z;
    r.g1 = g1; r.g2 = g2; r.b1 = b1; r.b2 = b2; r.flags = flags | (i << 8);
  }
}

// narrowphase over the pair list, one pair per lane; contacts keep pair order.
__device__ int phase_narrowphase(const DevBatch &B, const Ctx &c, const WorldParams &P, int npairs, Shared &sh,
                                 int maxc, bool apply_policy) {
  const PairRec *pairs = c.at<PairRec>(c.L.pairs);
  int *pcount = c.at<int>(c.L.pcount);
  ContactRec *con = c.at<ContactRec>(c.L.contacts);
  int nk = 0;
  if (maxc > MAX_SLOTS) maxc = MAX_SLOTS;
  for (int start = 0; start < npairs; start += blockDim.x) {
    int i = start + threadIdx.x;
    CPoint pts[MAX_SLOTS];
    int n = 0;
    if (i < npairs) {
      int g1 = pairs[i].g1, g2 = pairs[i].g2;
      bool skip = false;
      if (apply_policy && P.skip_connected) {
        int b1 = B.gbody[c.W.g0 + g1], b2 = B.gbody[c.W.g0 + g2];
        skip = connected(B, c, b1, b2);
      }
      if (!skip) {
        GeomView a = load_geom(B, c, g1), b = load_geom(B, c, g2);
        n = collide_pair(a, b, maxc, pts);
      }
      pcount[i] = n;
    }
    int tot;
    int off = block_excl_scan(n, tot, sh);
    if (n) {
      int o = nk + off;
      int room = c.W.kcap - o;
      if (room > 0) store_contacts(B, c, pairs[i].g1, pairs[i].g2, pts, n, con + o, room);
    }
    nk += tot;
  }
  if (nk > c.W.kcap) {
    if (threadIdx.x == 0) atomicOr(&B.counters[c.w * CNT_N + CNT_OVERFLOW], OVF_CONTACTS);
    nk = c.W.kcap;
  }
  __syncthreads();
  return nk;
}

// ------------------------------------------------------------------ step

struct StepIn {
  int nk;               // contacts
  bool from_host;       // contacts are ContactIn records of the host
  int seq_base;         // attach sequence of the first fused contact
};

__device__ inline Surf world_surface(const WorldParams &P) {
  Surf s;
  s.mode = P.surf_mode; s.mu = P.mu; s.mu2 = P.mu2; s.bounce = P.bounce; s.bounce_vel = P.bounce_vel;
  s.soft_erp = P.soft_erp; s.soft_cfm = P.soft_cfm; s.motion1 = P.motion1; s.motion2 = P.motion2;
  s.motionN = P.motionN; s.slip1 = P.slip1; s.slip2 = P.slip2; s.fdir1 = mk3(0.f, 0.f, 0.f);
  return s;
}

__device__ inline Surf host_surface(const ContactIn &ci) {
  Surf s;
  s.mode = ci.mode; s.mu = ci.mu; s.mu2 = ci.mu2; s.bounce = ci.bounce; s.bounce_vel = ci.bounce_vel;
  s.soft_erp = ci.soft_erp; s.soft_cfm = ci.soft_cfm; s.motion1 = ci.motion1; s.motion2 = ci.motion2;
  s.motionN = ci.motionN; s.slip1 = ci.slip1; s.slip2 = ci.slip2;
  s.fdir1 = mk3(ci.fdir1[0], ci.fdir1[1], ci.fdir1[2]);
  return s;
}

// relax one row: ode/src/quickstep.cpp:484-582 (J and b scaled by ad there).
// A body's 6 constraint-force floats are at bt[32 b + 9] (per-body scratch).
__device__ inline void relax_row(Row *rows, int r, float *bt) {
  Row &R = rows[r];
  float old = R.lambda;
  float *f1 = bt + 32 * R.b1 + 9;
  float s = R.J[0] * f1[0] + R.J[1] * f1[1] + R.J[2] * f1[2] + R.J[3] * f1[3] + R.J[4] * f1[4] + R.J[5] * f1[5];
  float *f2 = nullptr;
  if (R.b2 >= 0) {
    f2 = bt + 32 * R.b2 + 9;
    s += R.J[6] * f2[0] + R.J[7] * f2[1] + R.J[8] * f2[2] + R.J[9] * f2[3] + R.J[10] * f2[4] + R.J[11] * f2[5];
  }
  float delta = R.ad * (R.rhs - old * R.cfm - s);
  float lo, hi;
  if (R.findex >= 0) {
    hi = fabsf(R.hi * rows[R.findex].lambda);
    lo = -hi;
  } else {
    lo = R.lo;
    hi = R.hi;
  }
  float nl = old + delta;
  if (nl < lo) { delta = lo - old; R.lambda = lo; }
  else if (nl > hi) { delta = hi - old; R.lambda = hi; }
  else R.lambda = nl;
  for (int k = 0; k < 6; k++) f1[k] += delta * R.iMJ[k];
  if (f2)
    for (int k = 0; k < 6; k++) f2[k] += delta * R.iMJ[6 + k];
}

// LDS bytes of the reference-order solve of a world: its rows, the row
// order and 6 fc floats per body
__device__ inline long long ref_sor_lds_bytes(int nrows, int nb) {
  return (long long)sizeof(Row) * nrows + 8LL * nrows + 24LL * nb;
}
// the same with the rows left in global memory: impulses, order and fc only
__device__ inline long long ref_sor_lds_bytes_streamed(int nrows, int nb) {
  return 8LL * nrows + 24LL * nb;
}

// the read-only part of a row (everything but lambda), held in registers
struct RowC {
  float J[12], iMJ[12];
  float rhs, cfm, lo, hi, ad;
  int findex, b1, b2;
};

__device__ inline RowC load_rowc(const Row &R) {
  RowC q;
  for (int k = 0; k < 12; k++) { q.J[k] = R.J[k]; q.iMJ[k] = R.iMJ[k]; }
  q.rhs = R.rhs; q.cfm = R.cfm; q.lo = R.lo; q.hi = R.hi; q.ad = R.ad;
  q.findex = R.findex; q.b1 = R.b1; q.b2 = R.b2;
  return q;
}

// relax_row on a row held in registers, its impulse in lam[r] and fc in
// F[6 b]: the same arithmetic in the same order.  All reads (the row's
// impulse, its normal row's impulse, both bodies' fc) are issued before any
// write, so they share one LDS round trip.
__device__ inline void relax_rowc(const RowC &R, int r, float *lam, float *F) {
  const bool two = R.b2 >= 0;
  const float old = lam[r];
  const float lf = lam[R.findex >= 0 ? R.findex : r];
  float *f1 = F + 6 * R.b1;
  float *f2 = F + 6 * (two ? R.b2 : R.b1);
  float fa[6], fb[6];
  for (int k = 0; k < 6; k++) { fa[k] = f1[k]; fb[k] = f2[k]; }
  float s = R.J[0] * fa[0] + R.J[1] * fa[1] + R.J[2] * fa[2] + R.J[3] * fa[3] + R.J[4] * fa[4] + R.J[5] * fa[5];
  if (two) s += R.J[6] * fb[0] + R.J[7] * fb[1] + R.J[8] * fb[2] + R.J[9] * fb[3] + R.J[10] * fb[4] + R.J[11] * fb[5];
  float delta = R.ad * (R.rhs - old * R.cfm - s);
  float lo, hi;
  if (R.findex >= 0) {
    hi = fabsf(R.hi * lf);
    lo = -hi;
  } else {
    lo = R.lo;
    hi = R.hi;
  }
  float nl = old + delta;
  if (nl < lo) { delta = lo - old; lam[r] = lo; }
  else if (nl > hi) { delta = hi - old; lam[r] = hi; }
  else lam[r] = nl;
  for (int k = 0; k < 6; k++) f1[k] = fa[k] + delta * R.iMJ[k];
  if (two)
    for (int k = 0; k < 6; k++) f2[k] = fb[k] + delta * R.iMJ[6 + k];
}

// SOR_LCP in the reference order (quickstep.cpp:406-582): one island per
// lane; rows ordered with the findex = -1 rows first and the others from the
// tail, reshuffled every 8 iterations with the island's own RNG state.
// MODE 1: rows, impulses, order and fc in the LDS pool (copied in first and
// back after the iterations); MODE 2 (worlds whose rows exceed the pool):
// impulses, order and fc in LDS, the read-only rows streamed from global
// memory; in both the next row's read-only part is loaded while the current
// one is relaxed.  MODE 0: everything in the world's scratch (rows, ord, and
// fc at bt[32 b + 9]).  The arithmetic and its order are the same in all
// three.
template <int MODE>
__device__ void sor_reference(Row *rows, float *bt, const int *isl, int *ord, int nrows, int ngroups, int nb,
                              int iters) {
  constexpr bool LDS = MODE == 1;
  Row *R = LDS ? (Row *)od_pool : rows;
  int *O = LDS ? (int *)(R + nrows) : (MODE == 2 ? (int *)od_pool : ord);
  float *F = MODE == 0 ? bt : (float *)(O + nrows);
  float *LAM = MODE == 0 ? nullptr : F + 6 * nb;
  const int nt = blockDim.x;
  if (MODE != 0) {
    if (LDS)
      for (int i = threadIdx.x; i < nrows; i += nt) R[i] = rows[i];
    for (int i = threadIdx.x; i < nrows; i += nt) LAM[i] = rows[i].lambda;
    for (int b = threadIdx.x; b < nb; b += nt)
      for (int k = 0; k < 6; k++) F[6 * b + k] = bt[32 * b + 9 + k];
    __syncthreads();
  }
  for (int k = threadIdx.x; k < ngroups; k += nt) {
    int r0 = isl[5 * k + 2], m = isl[5 * k + 3];
    if (m <= 0) continue;
    unsigned seed = (unsigned)isl[5 * k + 4];
    int *o = O + r0;
    int head = 0, tail = m - 1;
    for (int i = 0; i < m; i++) {
      if (R[r0 + i].findex == -1) o[head++] = r0 + i;
      else o[tail--] = r0 + i;
    }
    for (int it = 0; it < iters; it++) {
      if ((it & 7) == 0) {
        for (int i = 1; i < m; i++) {
          int sw = lcg_below(seed, i + 1);
          int t = o[i]; o[i] = o[sw]; o[sw] = t;
        }
      }
      if (MODE != 0) {
        // two row buffers in turn: the next row loads while this one relaxes
        int ra = o[0], rb = ra;
        RowC qa = load_rowc(R[ra]), qb;
        for (int i = 0; i < m; i += 2) {
          if (i + 1 < m) { rb = o[i + 1]; qb = load_rowc(R[rb]); }
          relax_rowc(qa, ra, LAM, F);
          if (i + 1 >= m) break;
          if (i + 2 < m) { ra = o[i + 2]; qa = load_rowc(R[ra]); }
          relax_rowc(qb, rb, LAM, F);
        }
      } else {
        for (int i = 0; i < m; i++) relax_row(R, o[i], F);
      }
    }
  }
  // islands are solved by lanes of every wavefront; the integration reads
  // the fc of all bodies
  __syncthreads();
  if (MODE != 0) {
    for (int i = threadIdx.x; i < nrows; i += nt) rows[i].lambda = LAM[i];
    for (int b = threadIdx.x; b < nb; b += nt)
      for (int k = 0; k < 6; k++) bt[32 * b + 9 + k] = F[6 * b + k];
    __syncthreads();
  }
}

// |lambda - clamp(lambda - d, lo, hi)| evaluated without the cancellation of
// lambda - (lambda - d): |d| when the step stays inside the bounds, else the
// distance from lambda to the bound it crosses
__device__ inline float proj_residual(float lambda, float d, float lo, float hi) {
  const float p = lambda - d;
  if (p < lo) return fabsf(lambda - lo);
  if (p > hi) return fabsf(hi - lambda);
  return fabsf(d);
}

// complementarity residual of one row: |lambda - clamp(lambda - (A lambda - b)/A_ii)|
__device__ inline float row_residual(const Row *rows, int r, const float *fc) {
  const Row &R = rows[r];
  const float *f1 = fc + 32 * R.b1 + 9;
  float s = 0.f, d = 0.f;
  for (int k = 0; k < 6; k++) { s += R.J[k] * f1[k]; d += R.J[k] * R.iMJ[k]; }
  if (R.b2 >= 0) {
    const float *f2 = fc + 32 * R.b2 + 9;
    for (int k = 0; k < 6; k++) { s += R.J[6 + k] * f2[k]; d += R.J[6 + k] * R.iMJ[6 + k]; }
  }
  float Aii = d + R.cfm;
  float y = s + R.cfm * R.lambda - R.rhs;
  float lo = R.lo, hi = R.hi;
  if (R.findex >= 0) { hi = fabsf(R.hi * rows[R.findex].lambda); lo = -hi; }
  return proj_residual(R.lambda, y / Aii, lo, hi);
}

// Coloured SOR-LCP.  The working set -- per-body fc, world inverse inertia
// (symmetric) and inverse mass; the constraints sorted by colour with their
// rows stored contiguously in that order -- is copied into the LDS pool when
// it fits, else into the world's scratch.  One lane relaxes all rows of one
// constraint with both bodies' fc in registers; the constraints of a colour
// share no body, so a colour is one parallel pass.  Each row follows the
// reference's update (quickstep.cpp:484-582) with inv(M) J^T rebuilt from
// the body data (every supported row has J_lin2 = -J_lin1).
constexpr int SOR_MAX_ROWS = 6;   // rows of one constraint (contact 3, ball 3, hinge 6, hinge2 6)

// J_1 . f1 (+ J_2 . f2) of one row (J_l2 = -J_l1), summed in the order written
// Rows of the coloured solver's working set are stored with their four
// 16-byte chunks rotated by ((r >> 1) & 3) for row r: neighbouring lanes read
// rows 3 (contacts) or 1 apart, whose 64-byte starts fall on two LDS banks
// only; the rotation spreads one chunk read of eight such lanes over all 32
// banks.  The layout is the same in global scratch and in LDS.
__device__ inline int srow_chunk(int r, int j) { return j ^ ((r >> 1) & 3); }

__device__ inline void store_srow(SRow *base, int r, const SRow &v) {
  char *dst = (char *)(base + r);
  const char *src = (const char *)&v;
  for (int j = 0; j < 4; j++) __builtin_memcpy(dst + 16 * srow_chunk(r, j), src + 16 * j, 16);
}

__device__ inline SRow load_srow(const SRow *base, int r) {
  SRow v;
  const char *src = (const char *)(base + r);
  char *dst = (char *)&v;
  for (int j = 0; j < 4; j++) __builtin_memcpy(dst + 16 * j, src + 16 * srow_chunk(r, j), 16);
  return v;
}

__device__ inline float sor_dot(const SRow &r, const float *f1, bool two, const float *f2) {
  float s = r.l[0] * f1[0] + r.l[1] * f1[1] + r.l[2] * f1[2] + r.a1[0] * f1[3] + r.a1[1] * f1[4] + r.a1[2] * f1[5];
  if (two)
    s += -r.l[0] * f2[0] - r.l[1] * f2[1] - r.l[2] * f2[2] + r.a2[0] * f2[3] + r.a2[1] * f2[4] + r.a2[2] * f2[5];
  return s;
}

// largest complementarity residual |lambda - clamp(lambda - (A lambda - b)_i / A_ii)|
// over the rows of one constraint of the coloured solver (A_ii = w / ad)
__device__ inline float sor_cons_residual(const SBody *sb, const SRow *sr, const float *slam, const SCons q,
                                          float sor_w) {
  float f1[6], f2[6];
  for (int k = 0; k < 6; k++) f1[k] = sb[q.b1].fc[k];
  const bool two = q.b2 >= 0;
  if (two)
    for (int k = 0; k < 6; k++) f2[k] = sb[q.b2].fc[k];
  float r = 0.f;
  for (int k = 0; k < q.m; k++) {
    const SRow w = load_srow(sr, q.row0 + k);
    const float lam = slam[q.row0 + k];
    const float y = sor_dot(w, f1, two, f2) + w.cfm * lam - w.rhs;
    float lo = w.lo, hi = w.hi;
    if (w.findex >= 0) {
      hi = fabsf(w.hi * slam[q.row0 + w.findex]);
      lo = -hi;
    }
    r = fmaxf(r, proj_residual(lam, y * (w.ad / sor_w), lo, hi));
  }
  return r;
}

// fc += delta * inv(M) J^T of one row for both bodies of its constraint
__device__ inline void sor_push(const SRow &r, float delta, float im1, const float *iI1, float *f1, bool two,
                                float im2, const float *iI2, float *f2) {
  float d1 = delta * im1;
  f1[0] += d1 * r.l[0]; f1[1] += d1 * r.l[1]; f1[2] += d1 * r.l[2];
  f1[3] += delta * (iI1[0] * r.a1[0] + iI1[1] * r.a1[1] + iI1[2] * r.a1[2]);
  f1[4] += delta * (iI1[1] * r.a1[0] + iI1[3] * r.a1[1] + iI1[4] * r.a1[2]);
  f1[5] += delta * (iI1[2] * r.a1[0] + iI1[4] * r.a1[1] + iI1[5] * r.a1[2]);
  if (two) {
    float d2 = delta * im2;
    f2[0] -= d2 * r.l[0]; f2[1] -= d2 * r.l[1]; f2[2] -= d2 * r.l[2];
    f2[3] += delta * (iI2[0] * r.a2[0] + iI2[1] * r.a2[1] + iI2[2] * r.a2[2]);
    f2[4] += delta * (iI2[1] * r.a2[0] + iI2[3] * r.a2[1] + iI2[4] * r.a2[2]);
    f2[5] += delta * (iI2[2] * r.a2[0] + iI2[4] * r.a2[1] + iI2[5] * r.a2[2]);
  }
}

// Relaxes the rows of one constraint with one lane (quickstep.cpp:484-582):
// both bodies' fc in registers, the rows read three at a time so that the
// LDS latency is paid once per group of rows.  tol_mode: fc follows the
// stored impulse.
__device__ inline void sor_relax(SBody *__restrict__ sb, const SRow *__restrict__ sr, float *__restrict__ slam,
                                 float2 *__restrict__ ld, const SCons q, bool accel, bool tol_mode) {
  const bool two = q.b2 >= 0;
  const SBody B1 = sb[q.b1];
  SBody B2;
  if (two) B2 = sb[q.b2];
  float f1[6], iI1[6], f2[6], iI2[6], im2 = 0.f;
  for (int k = 0; k < 6; k++) { f1[k] = B1.fc[k]; iI1[k] = B1.iI[k]; f2[k] = 0.f; iI2[k] = 0.f; }
  const float im1 = B1.im;
  if (two) {
    for (int k = 0; k < 6; k++) { f2[k] = B2.fc[k]; iI2[k] = B2.iI[k]; }
    im2 = B2.im;
  }
  float lam0 = 0.f;
  for (int k0 = 0; k0 < q.m; k0 += 3) {
    SRow R3[3];
    float L3[3];
#pragma unroll
    for (int j = 0; j < 3; j++)
      if (k0 + j < q.m) { R3[j] = load_srow(sr, q.row0 + k0 + j); L3[j] = slam[q.row0 + k0 + j]; }
#pragma unroll
    for (int j = 0; j < 3; j++) {
      const int k = k0 + j;
      if (k >= q.m) break;
      const SRow &r = R3[j];
      const float s = sor_dot(r, f1, two, f2);
      const float old = L3[j];
      if (accel) ld[q.row0 + k].x = old;
      float delta = r.ad * (r.rhs - old * r.cfm - s);
      float lo = r.lo, hi = r.hi;
      if (r.findex >= 0) {
        float ln = (r.findex == 0) ? lam0 : slam[q.row0 + r.findex];
        hi = fabsf(r.hi * ln);
        lo = -hi;
      }
      float nl = old + delta;
      if (nl < lo) { delta = lo - old; nl = lo; }
      else if (nl > hi) { delta = hi - old; nl = hi; }
      else if (tol_mode) delta = nl - old;
      slam[q.row0 + k] = nl;
      if (k == 0) lam0 = nl;
      sor_push(r, delta, im1, iI1, f1, two, im2, iI2, f2);
    }
  }
  for (int k = 0; k < 6; k++) sb[q.b1].fc[k] = f1[k];
  if (two)
    for (int k = 0; k < 6; k++) sb[q.b2].fc[k] = f2[k];
}

// Starting impulses of one constraint's rows from the warm-start records of
// the previous step: a contact by its key (records sorted by key, binary
// search), a permanent joint by index when its row count is unchanged.
__device__ inline void warm_lookup(const WorldParams &P, const Cons &q, const ContactRec *con, const WarmRec *wrec,
                                   int nprev, const float *wjnt, float *wl) {
  if (q.kind == JT_CONTACT) {
    const ContactRec &cr = con[q.src];
    const unsigned long long key = contact_key(cr.g1, cr.g2, cr.flags >> 8);
    int lo = 0, hi = nprev;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (wrec[mid].key < key) lo = mid + 1;
      else hi = mid;
    }
    if (lo < nprev && wrec[lo].key == key)
      for (int k = 0; k < q.m && k < 3; k++) wl[k] = P.warm * wrec[lo].lam[k];
  } else {
    const float *jw = wjnt + 8 * q.src;
    if (__float_as_int(jw[6]) == q.m)
      for (int k = 0; k < q.m && k < 6; k++) wl[k] = P.warm * jw[k];
  }
}

// The coloured solver's working set: the rows (read-only during the sweeps),
// bodies, constraints in colour order, colour starts, the row impulses and
// (accelerated mode) the conjugate-gradient state: per row lambda before the
// sweep and the direction, per body fc before the sweep and the direction's
// fc (last, so that a pool can hold all but it).  Built in the world's
// scratch; copied to the LDS pool for the sweeps when it fits there (same
// layout).
struct SorSet {
  SRow *sr;
  SBody *sb;
  SCons *sc;
  int *sstart;
  float *slam;
  SAccBody *sa;
  float2 *ld;
};

__device__ inline SorSet sor_set(char *base, SRow *sr, int nb, int nrows, int nact) {
  SorSet S;
  S.sr = sr;
  S.sb = (SBody *)base;
  S.sc = (SCons *)(S.sb + nb);
  S.sstart = (int *)(S.sc + nact);
  S.slam = (float *)((char *)S.sstart + ((4LL * (2LL * nact + 8) + 15) & ~15LL));
  S.ld = (float2 *)((char *)S.slam + ((4LL * nrows + 15) & ~15LL));
  S.sa = (SAccBody *)((char *)S.ld + ((8LL * nrows + 15) & ~15LL));
  return S;
}

__device__ inline SorSet sor_global(const Ctx &c, int nrows, int nact) {
  SRow *sr = c.at<SRow>(c.L.sor);
  return sor_set((char *)(sr + nrows), sr, c.W.nb, nrows, nact);
}

// per-world values handed from the solver's set-up to the sweeps and on to
// the integration (split launches), in the world's misc block
enum { MISC_SOR_ROWS = 8, MISC_SOR_GROUPS = 9, MISC_SOR_RESIDUAL = 10, MISC_SOR_READY = 11 };

// Set-up: the working set in the world's scratch, in colour order, with the
// starting impulses (zero, or the warm start from the previous step).
__device__ void sor_prepare(const DevBatch &B, const Ctx &c, const WorldParams &P, const Cons *cons, const Row *rows,
                            const float *bt, const int *corder, const int *cstart, int nrows, int ngroups,
                            bool host_contacts, Shared &sh) {
  const int nb = c.W.nb;
  const int nact = cstart[ngroups];   // constraints with rows, in colour order
  const bool warm_on = P.warm != 0.f && !host_contacts && B.warm_n != nullptr;
  const int nprev = warm_on ? B.warm_n[c.w] : -1;
  const bool warm = warm_on && nprev >= 0;
  const WarmRec *wrec = warm_on ? (const WarmRec *)(B.warm + c.W.warm) : nullptr;
  const float *wjnt = warm_on ? (const float *)(B.warm + c.W.warm + od_align(24LL * c.W.kcap)) : nullptr;
  const ContactRec *con = c.at<ContactRec>(c.L.contacts);
  const int nt = blockDim.x;
  const SorSet S = sor_global(c, nrows, nact);
  for (int b = threadIdx.x; b < nb; b += nt) {
    const float *t = bt + 32 * b;
    SBody q;
    for (int k = 0; k < 6; k++) q.fc[k] = 0.f;
    q.iI[0] = t[0]; q.iI[1] = t[1]; q.iI[2] = t[2]; q.iI[3] = t[4]; q.iI[4] = t[5]; q.iI[5] = t[8];
    q.im = B.invMass[c.W.b0 + b];
    q.pad[0] = q.pad[1] = q.pad[2] = 0.f;
    S.sb[b] = q;
  }
  for (int k = threadIdx.x; k <= ngroups; k += nt) S.sstart[k] = cstart[k];
  int rbase = 0;
  for (int x0 = 0; x0 < nact; x0 += nt) {
    int x = x0 + threadIdx.x;
    const Cons *q = x < nact ? &cons[corder[x]] : nullptr;
    int m = q ? q->m : 0;
    int tot;
    int off = block_excl_scan(m, tot, sh);
    if (q) {
      SCons o;
      o.row0 = rbase + off; o.m = m; o.b1 = q->b1; o.b2 = q->b2;
      S.sc[x] = o;
      float wl[SOR_MAX_ROWS] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (warm) warm_lookup(P, *q, con, wrec, nprev, wjnt, wl);
      for (int k = 0; k < m; k++) {
        const Row &R = rows[q->row0 + k];
        SRow w;
        w.l[0] = R.J[0]; w.l[1] = R.J[1]; w.l[2] = R.J[2];
        w.a1[0] = R.J[3]; w.a1[1] = R.J[4]; w.a1[2] = R.J[5];
        w.a2[0] = R.J[9]; w.a2[1] = R.J[10]; w.a2[2] = R.J[11];
        w.rhs = R.rhs; w.cfm = R.cfm; w.ad = R.ad; w.lo = R.lo; w.hi = R.hi;
        w.pad = 0.f;
        w.findex = R.findex >= 0 ? R.findex - q->row0 : -1;
        store_srow(S.sr, o.row0 + k, w);
        S.slam[o.row0 + k] = k < SOR_MAX_ROWS ? wl[k] : 0.f;
      }
    }
    rbase += tot;
  }
  __syncthreads();
}

// Placement of the coloured solver's working set during the sweeps: all in
// global scratch, all in the LDS pool, everything but the (read-only) rows
// in the pool, everything but the rows and the bodies' conjugate-gradient
// state (read and written once per sweep, by one coalesced pass) in the
// pool, or also the rows' conjugate-gradient state in global scratch (inside
// the colour chain a relaxation only stores to it).
enum { SOR_IN_GLOBAL = 0, SOR_IN_LDS = 1, SOR_ROWS_GLOBAL = 2, SOR_ROWS_ACC_GLOBAL = 3, SOR_ROWS_CG_GLOBAL = 4 };

// The sweeps.  The working set (rows included, or all but the rows) is
// copied to the LDS pool first and back at the end; the placement is a
// template argument so that every access compiles to LDS (or global)
// instructions rather than generic flat ones.  Writes the iteration count
// and leaves the last residual in the world's misc block.
template <bool TOL, int PLACE>
__device__ __attribute__((always_inline)) void sor_solve(const DevBatch &B, const Ctx &c, const WorldParams &P,
                                                         int nrows, int ngroups, int nact, bool host_contacts,
                                                         Shared &sh) {
  constexpr bool LDS = PLACE != SOR_IN_GLOBAL;
  constexpr bool ROWS_LDS = PLACE == SOR_IN_LDS;
  const int nb = c.W.nb;
  const bool use_tol = TOL && P.tol > 0.f;
  const bool accel = use_tol && P.accel != 0;
  const bool warm = P.warm != 0.f && !host_contacts && B.warm_n != nullptr && B.warm_n[c.w] >= 0;
  const int nt = blockDim.x;
  const SorSet G = sor_global(c, nrows, nact);
  const SorSet S = ROWS_LDS ? sor_set((char *)((SRow *)od_pool + nrows), (SRow *)od_pool, nb, nrows, nact)
                 : LDS      ? sor_set((char *)od_pool, G.sr, nb, nrows, nact)
                            : G;
  SRow *__restrict__ sr = S.sr;
  SBody *__restrict__ sb = S.sb;
  SCons *__restrict__ sc = S.sc;
  int *__restrict__ sstart = S.sstart;
  float *__restrict__ slam = S.slam;
  SAccBody *__restrict__ sa = PLACE >= SOR_ROWS_ACC_GLOBAL ? G.sa : S.sa;
  float2 *__restrict__ ld = PLACE == SOR_ROWS_CG_GLOBAL ? G.ld : S.ld;
  if (LDS) {
    if (ROWS_LDS)
      for (int r = threadIdx.x; r < nrows; r += nt) sr[r] = G.sr[r];
    for (int b = threadIdx.x; b < nb; b += nt) sb[b] = G.sb[b];
    for (int x = threadIdx.x; x < nact; x += nt) sc[x] = G.sc[x];
    for (int k = threadIdx.x; k <= ngroups; k += nt) sstart[k] = G.sstart[k];
    for (int r = threadIdx.x; r < nrows; r += nt) slam[r] = G.slam[r];
  }
  if (accel) {
    for (int b = threadIdx.x; b < nb; b += nt)
      for (int k = 0; k < 6; k++) { sa[b].fdir[k] = 0.f; sa[b].fp[k] = 0.f; }
    for (int r = threadIdx.x; r < nrows; r += nt) ld[r] = make_float2(0.f, 0.f);
  }
  __syncthreads();
  if (warm) {
    // fc = inv(M) J^T lambda of the starting impulses, colour by colour
    // (a colour's constraints share no body)
    for (int g = 0; g < ngroups; g++) {
      const int x1 = sstart[g + 1];
      for (int x = sstart[g] + threadIdx.x; x < x1; x += nt) {
        const SCons q = sc[x];
        const bool two = q.b2 >= 0;
        float f1[6], f2[6] = {0, 0, 0, 0, 0, 0}, iI2[6] = {0, 0, 0, 0, 0, 0}, im2 = 0.f;
        for (int k = 0; k < 6; k++) f1[k] = sb[q.b1].fc[k];
        if (two)
          for (int k = 0; k < 6; k++) { f2[k] = sb[q.b2].fc[k]; iI2[k] = sb[q.b2].iI[k]; }
        if (two) im2 = sb[q.b2].im;
        float iI1[6];
        for (int k = 0; k < 6; k++) iI1[k] = sb[q.b1].iI[k];
        const float im1 = sb[q.b1].im;
        for (int k = 0; k < q.m && k < SOR_MAX_ROWS; k++) {
          const float lam = slam[q.row0 + k];
          if (lam != 0.f) sor_push(load_srow(sr, q.row0 + k), lam, im1, iI1, f1, two, im2, iI2, f2);
        }
        for (int k = 0; k < 6; k++) sb[q.b1].fc[k] = f1[k];
        if (two)
          for (int k = 0; k < 6; k++) sb[q.b2].fc[k] = f2[k];
      }
      __syncthreads();
    }
  }
  // fixed iteration count, or (tol > 0) iterate until the LCP residual of
  // the world is at most tol, checked every check_every iterations, up to
  // max_iterations; accelerated: between sweeps a nonsmooth nonlinear
  // conjugate-gradient step (Silcowitz, Niebe, Erleben 2010): with
  // d = lambda_after - lambda_before, beta = |d|^2/|d_prev|^2 (0 if > 1),
  // lambda += beta p_prev (applied before the next sweep, so the solver
  // always ends on a projected sweep), p = d + beta p_prev; fc follows
  // through M^-1 J^T p
  const int maxit = use_tol ? P.max_iterations : P.iterations;
  const int every = use_tol ? (P.check_every > 0 ? P.check_every : 1) : (maxit > 0 ? maxit : 1);
  float rr_prev = 0.f, res_prev = INFINITY;
  bool acc = accel;        // CG steps still on for this solve
  float res_last = 0.f;    // residual at the last check
  int it = 0;
  while (it < maxit) {
    const int n = every < maxit - it ? every : maxit - it;
    for (int rep = 0; rep < n; rep++) {
      if (acc && rep == 0) {
        // the pending momentum beta p_prev of the previous sweep (kept in
        // ld.x for the rows and fp for the bodies), then fc before this
        // sweep; inside a block of sweeps between two checks this is folded
        // into the previous sweep's CG update below
        for (int r = threadIdx.x; r < nrows; r += nt)
          if (ld[r].x != 0.f) slam[r] += ld[r].x;
        for (int e = threadIdx.x; e < 6 * nb; e += nt) {   // one fc component per lane
          const int b = e / 6, k = e - 6 * b;
          sb[b].fc[k] += sa[b].fp[k];
          sa[b].fp[k] = sb[b].fc[k];
        }
        __syncthreads();
      }
      for (int g = 0; g < ngroups; g++) {
        const int x1 = sstart[g + 1];
        for (int x = sstart[g] + threadIdx.x; x < x1; x += nt) sor_relax(sb, sr, slam, ld, sc[x], accel, use_tol);
        __syncthreads();
      }
      if (acc) {
        // d = this sweep's change; beta = |d|^2 / |d_prev|^2 (restart when
        // > 1); the momentum beta p_prev is applied before the next sweep
        float loc = 0.f;
        for (int r = threadIdx.x; r < nrows; r += nt) {
          float d = slam[r] - ld[r].x;
          loc += d * d;
        }
        const float rr = block_fsum(loc, sh);
        float beta = (rr_prev > 0.f) ? rr / rr_prev : 0.f;
        if (!(beta <= 1.f)) beta = 0.f;
        // the next sweep of this block takes the momentum now (the same
        // operations as the pending-momentum step above); the last sweep
        // before a check leaves it pending, so the check sees the sweep's
        // own result
        const bool now = rep + 1 < n;
        for (int r = threadIdx.x; r < nrows; r += nt) {
          float d = slam[r] - ld[r].x;
          float pend = beta * ld[r].y;   // momentum along the previous direction
          ld[r].y = d + pend;            // new direction
          if (!now) ld[r].x = pend;
          else if (pend != 0.f) slam[r] += pend;
        }
        for (int e = threadIdx.x; e < 6 * nb; e += nt) {   // one fc component per lane
          const int b = e / 6, k = e - 6 * b;
          float pend = beta * sa[b].fdir[k];
          sa[b].fdir[k] = (sb[b].fc[k] - sa[b].fp[k]) + pend;
          if (!now) {
            sa[b].fp[k] = pend;
          } else {
            sb[b].fc[k] += pend;
            sa[b].fp[k] = sb[b].fc[k];
          }
        }
        rr_prev = rr;
        __syncthreads();
      }
    }
    it += n;
    if (use_tol) {
      float r = 0.f;
      for (int x = threadIdx.x; x < nact; x += nt) r = fmaxf(r, sor_cons_residual(sb, sr, slam, sc[x], P.sor_w));
      r = block_max(r, sh);
      res_last = r;
      if (r <= P.tol) break;
      if (acc && !(r <= res_prev)) {
        // the residual rose since the last check: restart the CG steps, and
        // stop them for the rest of the solve when the residual is still far
        // (> 1000 tol) from the tolerance
        for (int q = threadIdx.x; q < nrows; q += nt) ld[q] = make_float2(0.f, 0.f);
        for (int b = threadIdx.x; b < nb; b += nt)
          for (int k = 0; k < 6; k++) { sa[b].fdir[k] = 0.f; sa[b].fp[k] = 0.f; }
        rr_prev = 0.f;
        if (r > 1000.f * P.tol) acc = false;
        __syncthreads();
      }
      res_prev = r;
    }
  }
  if (threadIdx.x == 0) {
    B.counters[c.w * CNT_N + CNT_ITERS] = it;
    c.at<int>(c.L.misc)[MISC_SOR_RESIDUAL] = __float_as_int(use_tol ? res_last : 0.f);
    if (use_tol) {
      int *cw = B.counters + c.w * CNT_N;
      if (!(res_last <= P.tol)) cw[CNT_UNCONV] += 1;
      // non-negative floats order like their bit patterns
      const int rb = __float_as_int(res_last == res_last ? res_last : INFINITY);
      if (rb > cw[CNT_RESMAX]) cw[CNT_RESMAX] = rb;
    }
  }
  if (LDS) {
    for (int b = threadIdx.x; b < nb; b += nt)
      for (int k = 0; k < 6; k++) G.sb[b].fc[k] = sb[b].fc[k];
    for (int r = threadIdx.x; r < nrows; r += nt) G.slam[r] = slam[r];
  }
  __syncthreads();
}

// the sweeps on the working set in global scratch (worlds larger than the LDS
// pool); kept out of line so that the LDS variant is not merged with it
template <bool TOL>
__device__ __attribute__((noinline)) void sor_solve_global(const DevBatch &B, const Ctx &c, const WorldParams &P,
                                                           int nrows, int ngroups, int nact, bool host_contacts,
                                                           Shared &sh) {
  sor_solve<TOL, SOR_IN_GLOBAL>(B, c, P, nrows, ngroups, nact, host_contacts, sh);
}

// the sweeps with the rows read from global scratch and the rest of the
// working set in the pool (worlds whose rows do not fit); out of line too
template <bool TOL>
__device__ __attribute__((noinline)) void sor_solve_rows_global(const DevBatch &B, const Ctx &c,
                                                                const WorldParams &P, int nrows, int ngroups,
                                                                int nact, bool host_contacts, Shared &sh) {
  sor_solve<TOL, SOR_ROWS_GLOBAL>(B, c, P, nrows, ngroups, nact, host_contacts, sh);
}

template <bool TOL>
__device__ __attribute__((noinline)) void sor_solve_rows_acc_global(const DevBatch &B, const Ctx &c,
                                                                    const WorldParams &P, int nrows, int ngroups,
                                                                    int nact, bool host_contacts, Shared &sh) {
  sor_solve<TOL, SOR_ROWS_ACC_GLOBAL>(B, c, P, nrows, ngroups, nact, host_contacts, sh);
}

template <bool TOL>
__device__ __attribute__((noinline)) void sor_solve_rows_cg_global(const DevBatch &B, const Ctx &c,
                                                                   const WorldParams &P, int nrows, int ngroups,
                                                                   int nact, bool host_contacts, Shared &sh) {
  sor_solve<TOL, SOR_ROWS_CG_GLOBAL>(B, c, P, nrows, ngroups, nact, host_contacts, sh);
}

template <bool TOL>
__device__ inline void sor_solve_any(const DevBatch &B, const Ctx &c, const WorldParams &P, int nrows, int ngroups,
                                     int nact, bool host_contacts, Shared &sh) {
  const bool accel = TOL && P.tol > 0.f && P.accel != 0;
  if (sor_lds_bytes(c.W.nb, nrows, nact, accel) <= B.lds_bytes)
    sor_solve<TOL, SOR_IN_LDS>(B, c, P, nrows, ngroups, nact, host_contacts, sh);
  else if (sor_lds_bytes_rows_global(c.W.nb, nrows, nact, accel) <= B.lds_bytes)
    sor_solve_rows_global<TOL>(B, c, P, nrows, ngroups, nact, host_contacts, sh);
  else if (accel && sor_lds_bytes_rows_acc_global(c.W.nb, nrows, nact, accel) <= B.lds_bytes)
    sor_solve_rows_acc_global<TOL>(B, c, P, nrows, ngroups, nact, host_contacts, sh);
  else if (accel && sor_lds_bytes_rows_cg_global(c.W.nb, nrows, nact) <= B.lds_bytes)
    sor_solve_rows_cg_global<TOL>(B, c, P, nrows, ngroups, nact, host_contacts, sh);
  else
    sor_solve_global<TOL>(B, c, P, nrows, ngroups, nact, host_contacts, sh);
}

// After the sweeps: impulses back to the rows (constraint order), fc to the
// per-body scratch, and the warm-start records of the next step.
__device__ void sor_finish(const DevBatch &B, const Ctx &c, const WorldParams &P, const Cons *cons, Row *rows,
                           float *bt, const int *corder, int nrows, int nact, int nk, int ncons, bool host_contacts) {
  const int nb = c.W.nb;
  const int nt = blockDim.x;
  const SorSet G = sor_global(c, nrows, nact);
  const bool warm_on = P.warm != 0.f && !host_contacts && B.warm_n != nullptr;
  WarmRec *wrec = warm_on ? (WarmRec *)(B.warm + c.W.warm) : nullptr;
  float *wjnt = warm_on ? (float *)(B.warm + c.W.warm + od_align(24LL * c.W.kcap)) : nullptr;
  const ContactRec *con = c.at<ContactRec>(c.L.contacts);
  const float res_last = __int_as_float(c.at<int>(c.L.misc)[MISC_SOR_RESIDUAL]);
  for (int x = threadIdx.x; x < nact; x += nt) {
    const Cons &q = cons[corder[x]];
    const SCons o = G.sc[x];
    for (int k = 0; k < o.m; k++) rows[q.row0 + k].lambda = G.slam[o.row0 + k];
  }
  for (int b = threadIdx.x; b < nb; b += nt)
    for (int k = 0; k < 6; k++) bt[32 * b + 9 + k] = G.sb[b].fc[k];
  __syncthreads();
  const bool use_tol = P.tol > 0.f;
  if (warm_on && use_tol && !(res_last <= 100.f * P.tol)) {
    // a solve that ended at the cap far from the tolerance (residual > 100
    // tol) is no starting point: the next solve starts from zero
    if (threadIdx.x == 0) B.warm_n[c.w] = -1;
    __syncthreads();
  } else if (warm_on) {
    // this step's impulses, the next step's starting point; contacts are in
    // pair order, so the records come out sorted by key
    for (int i = threadIdx.x; i < nk; i += nt) {
      const Cons &q = cons[i];
      const ContactRec &cr = con[q.src];
      WarmRec r;
      r.key = contact_key(cr.g1, cr.g2, cr.flags >> 8);
      for (int k = 0; k < 3; k++) r.lam[k] = k < q.m ? rows[q.row0 + k].lambda : 0.f;
      r.pad = 0;
      wrec[i] = r;
    }
    for (int i = nk + threadIdx.x; i < ncons; i += nt) {
      const Cons &q = cons[i];
      float *jw = wjnt + 8 * q.src;
      for (int k = 0; k < 6; k++) jw[k] = k < q.m ? rows[q.row0 + k].lambda : 0.f;
      jw[6] = __int_as_float(q.m);
      jw[7] = 0.f;
    }
    if (threadIdx.x == 0) B.warm_n[c.w] = nk;
    __syncthreads();
  }
}

// the coloured solver in one kernel: set-up, sweeps, write-back
template <bool TOL>
__device__ void sor_colored(const DevBatch &B, const Ctx &c, const WorldParams &P, const Cons *cons, Row *rows,
                            float *bt, const int *corder, const int *cstart, int nrows, int ngroups, int nk,
                            int ncons, bool host_contacts, Shared &sh) {
  const int nact = cstart[ngroups];
  sor_prepare(B, c, P, cons, rows, bt, corder, cstart, nrows, ngroups, host_contacts, sh);
  sor_solve_any<TOL>(B, c, P, nrows, ngroups, nact, host_contacts, sh);
  sor_finish(B, c, P, cons, rows, bt, corder, nrows, nact, nk, ncons, host_contacts);
}

// LDS bytes of the reference-order constraint ordering: 12 ints per
// constraint, 5 per body, one per joint
__device__ inline long long ref_order_lds_bytes(int nb, int ncons, int nj) {
  return 4LL * (12LL * ncons + 5LL * nb + nj + 16);
}

// The reference order of constraints and rows in the LDS pool: the same
// steps as the global-memory version in phase_step (adjacency newest attach
// first, island walk, rows in island order, per-island RNG state and Info2
// erp), with the constraint fields gathered into LDS by all lanes first and
// one lane walking the islands there.  The lane writes the body visit order
// and the island table straight to global memory (nothing reads them back
// during the walk); all lanes then write the constraint order, row offsets,
// erp, and m = 0 for the constraints no island reaches.
__device__ void ref_order_lds(const DevBatch &B, const Ctx &c, WorldParams &P, Cons *cons, int nk, int ncons,
                              int *corder, int *isl, float *erp_out, int *misc, Shared &sh) {
  const int nb = c.W.nb, nj = ncons - nk, nt = blockDim.x;
  int *cb1 = (int *)od_pool, *cb2 = cb1 + ncons, *cm = cb2 + ncons, *ckind = cm + ncons;
  int *cseq = ckind + ncons;               // attach sequence, then first row
  int *cen = cseq + ncons, *ctag = cen + ncons, *cord = ctag + ncons;
  float *cerp = (float *)(cord + ncons);   // Info2 erp seen by the constraint
  float *cbe = cerp + ncons;               // a ball joint's own erp
  int *adj = (int *)(cbe + ncons);         // 2 per constraint
  int *adjs = adj + 2 * ncons;             // nb + 1
  int *fill = adjs + nb + 1;               // fill cursors, then body tags
  int *tag = fill;
  int *stack = fill + nb;
  int *ispan = stack + nb;                 // first constraint and count per island
  int *jord = ispan + 2 * nb;              // nj
  int *border = corder + ncons;            // body visit order (global)
  for (int i = threadIdx.x; i < ncons; i += nt) {
    const Cons q = cons[i];
    cb1[i] = q.b1; cb2[i] = q.b2; cm[i] = q.m; ckind[i] = q.kind; cseq[i] = q.seq;
    cen[i] = q.b1 >= 0 && (B.invMass[c.W.b0 + q.b1] > 0.f || (q.b2 >= 0 && B.invMass[c.W.b0 + q.b2] > 0.f));
    cbe[i] = q.kind == JT_BALL ? B.joints[c.W.j0 + q.src].erp : 0.f;
    ctag[i] = 0;
  }
  for (int b = threadIdx.x; b <= nb; b += nt) adjs[b] = 0;
  __syncthreads();
  // adjacency list sizes (LDS atomics), then their prefix (block scan) and
  // the fill cursors, all lanes
  for (int i = threadIdx.x; i < ncons; i += nt) {
    if (cm[i] <= 0 && ckind[i] != JT_CONTACT) continue;
    if (cb1[i] < 0) continue;
    atomicAdd(&adjs[cb1[i] + 1], 1);
    if (cb2[i] >= 0) atomicAdd(&adjs[cb2[i] + 1], 1);
  }
  __syncthreads();
  {
    const int per = (nb + 1 + nt - 1) / nt, d0 = threadIdx.x * per;
    const int d1 = d0 + per < nb + 1 ? d0 + per : nb + 1;
    int loc = 0;
    for (int d = d0; d < d1; d++) loc += adjs[d];
    int tot;
    int run = block_excl_scan(loc, tot, sh);
    for (int d = d0; d < d1; d++) { run += adjs[d]; adjs[d] = run; }
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += nt) fill[b] = adjs[b];
  __syncthreads();
  if (threadIdx.x == 0) {
    int nj2 = 0;
    for (int i = nk; i < ncons; i++) {
      if (cb1[i] < 0) continue;
      int s = cseq[i], p = nj2++;
      while (p > 0 && cseq[jord[p - 1]] > s) { jord[p] = jord[p - 1]; p--; }
      jord[p] = i;
    }
    int ci = nk - 1, jj = nj2 - 1;
    while (ci >= 0 || jj >= 0) {
      int pick;
      if (jj < 0) pick = ci--;
      else if (ci < 0) pick = jord[jj--];
      else if (cseq[ci] > cseq[jord[jj]]) pick = ci--;
      else pick = jord[jj--];
      if (cb1[pick] < 0) continue;
      adj[fill[cb1[pick]]++] = pick;
      if (cb2[pick] >= 0) adj[fill[cb2[pick]]++] = pick;
    }
    for (int b = 0; b < nb; b++) tag[b] = 0;
    int nic = 0, nib = 0, nisl = 0;
    for (int bb = nb - 1; bb >= 0; bb--) {
      if (tag[bb]) continue;
      tag[bb] = 1;
      int c0 = nic;
      border[nib++] = bb;
      int sp = 0, b = bb;
      while (true) {
        for (int k = adjs[b]; k < adjs[b + 1]; k++) {
          int qi = adj[k];
          if (ctag[qi]) continue;
          if (!cen[qi]) { ctag[qi] = 2; continue; }
          ctag[qi] = 1;
          cord[nic++] = qi;
          int other = (cb1[qi] == b) ? cb2[qi] : cb1[qi];
          if (other >= 0 && tag[other] <= 0) {
            tag[other] = 1;
            stack[sp++] = other;
          }
        }
        if (sp == 0) break;
        b = stack[--sp];
        border[nib++] = b;
      }
      ispan[2 * nisl + 0] = c0;
      ispan[2 * nisl + 1] = nic - c0;
      nisl++;
    }
    unsigned seed = P.seed;
    unsigned long long shuffles = (unsigned long long)((P.iterations + 7) / 8);
    int r = 0;
    for (int k = 0; k < nisl; k++) {
      int c0 = ispan[2 * k], cn = ispan[2 * k + 1];
      int r0 = r;
      float erp = P.erp;
      for (int x = 0; x < cn; x++) {
        const int i = cord[c0 + x];
        cseq[i] = r;
        r += cm[i];
        if (ckind[i] == JT_BALL) erp = cbe[i];
        cerp[i] = erp;
      }
      int m = r - r0;
      isl[5 * k + 0] = c0;
      isl[5 * k + 1] = cn;
      isl[5 * k + 2] = r0;
      isl[5 * k + 3] = m;
      isl[5 * k + 4] = (int)seed;
      if (m > 0 && m - 1 > 0) seed = lcg_skip(seed, shuffles * (unsigned long long)(m - 1));
    }
    P.seed = seed;
    misc[0] = r;
    misc[1] = nisl;
    misc[2] = nic;
    misc[3] = nib;
    sh.ival[5] = nic;
  }
  __syncthreads();
  const int nic = sh.ival[5];
  for (int i = threadIdx.x; i < ncons; i += nt) {
    if (ctag[i] != 1) {
      cons[i].m = 0;
    } else {
      cons[i].row0 = cseq[i];
      erp_out[i] = cerp[i];
    }
  }
  for (int x = threadIdx.x; x < nic; x += nt) corder[x] = cord[x];
  __syncthreads();
}

#include "lcp_exact.hpp"

// STAGE 0: the whole step; 1: up to the coloured solver's set-up (the sweeps
// and the rest run in later kernels); 2: after the sweeps (write-back,
// velocities, integration)
template <bool TOL, int STAGE>
__device__ __attribute__((always_inline)) void phase_step(const DevBatch &B, const Ctx &c, WorldParams &P,
                                                          const StepIn &in, Shared &sh, bool ref) {
  const int nb = c.W.nb, nj = c.W.nj, nk = in.nk;
  const float h = P.dt, hinv = 1.f / P.dt;
  float *bt = c.at<float>(c.L.btmp);   // per body: invIw 0..8, fc 9..14, tmp 15..20, extra torque 21..23
  Cons *cons = c.at<Cons>(c.L.cons);
  Row *rows = c.at<Row>(c.L.rows);
  const ContactRec *con = c.at<ContactRec>(c.L.contacts);
  int *misc = c.at<int>(c.L.misc);
  const int ncons = nk + nj;
  int *corder = c.at<int>(c.L.order);          // constraint order (island order when ref)
  int *isl = c.at<int>(c.L.island);            // ref: [cons0, ncons, row0, nrows, seed] per island
  int *color = c.at<int>(c.L.color);
  int nrows = 0, ngroups = 0;
  if (STAGE == 2) {
    if (misc[MISC_SOR_READY] == 0) return;     // the step left the world unchanged
    nrows = misc[MISC_SOR_ROWS];
    ngroups = misc[MISC_SOR_GROUPS];
    const int nact = isl[ngroups];
    const bool accel = TOL && P.tol > 0.f && P.accel != 0;
    if (sor_lds_bytes(nb, nrows, nact, accel) > B.sor_lds_bytes) {
      // not solved by the sweep kernel (its LDS pool is too small)
      sor_solve_global<TOL>(B, c, P, nrows, ngroups, nact, in.from_host, sh);
      c.clk.lap(PH_SOR);
    }
    sor_finish(B, c, P, cons, rows, bt, corder, nrows, nact, nk, ncons, in.from_host);
  } else {
  // --- bodies: world inverse inertia, gyroscopic torque, gravity
  //     (quickstep.cpp:609-665)
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int gb = c.W.b0 + b;
    Mat3 R = load_mat(B.R + 3 * (size_t)gb);
    Mat3 iI = rotate_tensor(R, load_mat(B.invI + 3 * (size_t)gb));
    float *t = bt + 32 * b;
    t[0] = iI.r0.x; t[1] = iI.r0.y; t[2] = iI.r0.z;
    t[3] = iI.r1.x; t[4] = iI.r1.y; t[5] = iI.r1.z;
    t[6] = iI.r2.x; t[7] = iI.r2.y; t[8] = iI.r2.z;
    for (int k = 9; k < 24; k++) t[k] = 0.f;
    float3 w = xyz(B.avel[gb]);
    float4 ta = B.tacc[gb];
    Mat3 Iw = rotate_tensor(R, load_mat(B.I + 3 * (size_t)gb));
    float3 gyro = cross(w, Iw.mul(w));
    ta.x -= gyro.x; ta.y -= gyro.y; ta.z -= gyro.z;
    B.tacc[gb] = ta;
    float4 fa = B.facc[gb];
    float m = B.mass[gb];
    if (P.gravity[0] != 0.f) fa.x += m * P.gravity[0];
    if (P.gravity[1] != 0.f) fa.y += m * P.gravity[1];
    if (P.gravity[2] != 0.f) fa.z += m * P.gravity[2];
    B.facc[gb] = fa;
  }
  // --- constraints: contacts (k < nk) then permanent joints
  for (int i = threadIdx.x; i < ncons; i += blockDim.x) {
    Cons &q = cons[i];
    if (i < nk) {
      q.kind = JT_CONTACT;
      q.src = i;
      if (in.from_host) {
        const ContactIn &ci = B.cin[B.cin_start[c.w] + i];
        q.b1 = ci.b1; q.b2 = ci.b2; q.seq = ci.seq; q.fb = ci.fb;
      } else {
        q.b1 = con[i].b1; q.b2 = con[i].b2; q.seq = in.seq_base + i; q.fb = -1;
      }
    } else {
      const JointDev &j = B.joints[c.W.j0 + i - nk];
      q.kind = j.type; q.src = i - nk; q.b1 = j.b1; q.b2 = j.b2; q.seq = j.seq; q.fb = j.fb;
    }
    q.m = 0;
    q.row0 = 0;
  }
  __syncthreads();
  // --- getInfo1 (quickstep.cpp:670-688); joints update their limit state
  for (int i = threadIdx.x; i < ncons; i += blockDim.x) {
    Cons &q = cons[i];
    int m = 0;
    if (q.b1 >= 0) {
      if (q.kind == JT_CONTACT) {
        Surf s = in.from_host ? host_surface(B.cin[B.cin_start[c.w] + i]) : world_surface(P);
        m = contact_info1(s);
      } else {
        JointDev &j = B.joints[c.W.j0 + q.src];
        BodyPose s1 = load_pose(B, c.W.b0 + q.b1);
        BodyPose s2;
        if (q.b2 >= 0) s2 = load_pose(B, c.W.b0 + q.b2);
        if (!(j.type == JT_HINGE2 && q.b2 < 0)) m = joint_info1(j, s1, q.b2 >= 0 ? &s2 : nullptr);
      }
    }
    q.m = m;
  }
  __syncthreads();
  c.clk.lap(PH_BODIES);
  // --- constraint order, row offsets, groups
  if (ref && ref_order_lds_bytes(nb, ncons, nj) <= B.lds_bytes) {
    ref_order_lds(B, c, P, cons, nk, ncons, corder, isl, (float *)(color + nb), misc, sh);
    nrows = misc[0];
    ngroups = misc[1];
  } else if (ref) {
    if (threadIdx.x == 0) {
      // adjacency, newest attach first: body joint lists prepend (ode.cpp:1382-1389)
      int *adjs = c.at<int>(c.L.adjs);
      int *adj = c.at<int>(c.L.adj);
      for (int b = 0; b <= nb; b++) adjs[b] = 0;
      for (int i = 0; i < ncons; i++) {
        if (cons[i].m <= 0 && cons[i].kind != JT_CONTACT) continue;
        if (cons[i].b1 < 0) continue;
        adjs[cons[i].b1 + 1]++;
        if (cons[i].b2 >= 0) adjs[cons[i].b2 + 1]++;
      }
      for (int b = 0; b < nb; b++) adjs[b + 1] += adjs[b];
      // visit constraints in descending attach sequence: contacts carry
      // increasing sequence numbers in order; joints are merged in by seq
      int *jord = color;   // joints sorted by seq (insertion sort, few joints)
      int nj2 = 0;
      for (int i = nk; i < ncons; i++) {
        if (cons[i].b1 < 0) continue;
        int s = cons[i].seq, p = nj2++;
        while (p > 0 && cons[jord[p - 1]].seq > s) { jord[p] = jord[p - 1]; p--; }
        jord[p] = i;
      }
      int *fill = isl;     // per-body fill cursor (scratch reuse)
      for (int b = 0; b < nb; b++) fill[b] = adjs[b];
      int ci = nk - 1, jj = nj2 - 1;
      while (ci >= 0 || jj >= 0) {
        int pick;
        if (jj < 0) pick = ci--;
        else if (ci < 0) pick = jord[jj--];
        else if (cons[ci].seq > cons[jord[jj]].seq) pick = ci--;
        else pick = jord[jj--];
        const Cons &q = cons[pick];
        if (q.b1 < 0) continue;
        adj[fill[q.b1]++] = pick;
        if (q.b2 >= 0) adj[fill[q.b2]++] = pick;
      }
      // islands: depth-first walk from bodies in world-list order, newest
      // body first (util.cpp:420-487)
      int *tag = color;           // body tags, then constraint tags
      int *ctag = color + nb;
      for (int b = 0; b < nb; b++) tag[b] = 0;
      for (int i = 0; i < ncons; i++) ctag[i] = 0;
      int *border = corder + ncons;  // body visit order
      int *stack = isl + 5 * (nb + 1);  // island table first, then the stack
      int nic = 0, nib = 0, nisl = 0;
      for (int bb = nb - 1; bb >= 0; bb--) {
        if (tag[bb]) continue;
        tag[bb] = 1;
        int c0 = nic;
        border[nib++] = bb;
        int sp = 0, b = bb;
        while (true) {
          for (int k = adjs[b]; k < adjs[b + 1]; k++) {
            int qi = adj[k];
            if (ctag[qi]) continue;
            const Cons &q = cons[qi];
            bool enabled = B.invMass[c.W.b0 + q.b1] > 0.f || (q.b2 >= 0 && B.invMass[c.W.b0 + q.b2] > 0.f);
            if (!enabled) { ctag[qi] = 2; continue; }
            ctag[qi] = 1;
            corder[nic++] = qi;
            int other = (q.b1 == b) ? q.b2 : q.b1;
            if (other >= 0 && tag[other] <= 0) {
              tag[other] = 1;
              stack[sp++] = other;
            }
          }
          if (sp == 0) break;
          b = stack[--sp];
          border[nib++] = b;
        }
        isl[5 * nisl + 0] = c0;
        isl[5 * nisl + 1] = nic - c0;
        nisl++;
      }
      for (int i = 0; i < ncons; i++)
        if (ctag[i] != 1) cons[i].m = 0;
      // rows in island order; the RNG state of every island by skip-ahead:
      // an island with m rows makes ceil(iterations/8) * (m-1) draws
      // (quickstep.cpp:473-482)
      unsigned seed = P.seed;
      unsigned long long shuffles = (unsigned long long)((P.iterations + 7) / 8);
      int r = 0;
      float erp = P.erp;
      for (int k = 0; k < nisl; k++) {
        int c0 = isl[5 * k], cn = isl[5 * k + 1];
        int r0 = r;
        erp = P.erp;   // Info2 erp is reset per island (quickstep.cpp:777-779)
        for (int x = 0; x < cn; x++) {
          Cons &q = cons[corder[c0 + x]];
          q.row0 = r;
          r += q.m;
          // a ball joint sets the Info2 erp seen by the joints after it
          if (q.kind == JT_BALL) erp = B.joints[c.W.j0 + q.src].erp;
          ((float *)ctag)[corder[c0 + x]] = erp;
        }
        int m = r - r0;
        isl[5 * k + 2] = r0;
        isl[5 * k + 3] = m;
        isl[5 * k + 4] = (int)seed;
        if (m > 0 && m - 1 > 0) seed = lcg_skip(seed, shuffles * (unsigned long long)(m - 1));
      }
      P.seed = seed;
      misc[0] = r;
      misc[1] = nisl;
      misc[2] = nic;
      misc[3] = nib;
    }
    __syncthreads();
    nrows = misc[0];
    ngroups = misc[1];
  } else {
    // natural order; rows by a block scan
    int r0 = 0;
    for (int start = 0; start < ncons; start += blockDim.x) {
      int i = start + threadIdx.x;
      int m = (i < ncons) ? cons[i].m : 0;
      int tot;
      int off = block_excl_scan(m, tot, sh);
      if (i < ncons) { cons[i].row0 = r0 + off; corder[i] = i; }
      r0 += tot;
    }
    nrows = r0;
    // colouring: each round takes the constraints whose priority is the
    // largest among the still uncoloured constraints at both their bodies;
    // the round's winners share no body.  Working set in the LDS pool when
    // it fits.
    const bool in_lds = 4LL * (nb + 3LL * ncons) <= B.lds_bytes;
    int *bmax = in_lds ? (int *)od_pool : c.at<int>(c.L.adjs);
    int *col = in_lds ? bmax + nb : color;
    int *cb1 = in_lds ? col + ncons : c.at<int>(c.L.adj);
    int *cb2 = cb1 + ncons;
    for (int i = threadIdx.x; i < ncons; i += blockDim.x) {
      col[i] = (cons[i].m > 0) ? -1 : -2;
      cb1[i] = cons[i].b1;
      cb2[i] = cons[i].b2;
    }
    // per body: the colours taken by its coloured constraints (64-bit mask
    // in the body's unused temp words 24-25)
    for (int b = threadIdx.x; b < nb; b += blockDim.x) *(unsigned long long *)(bt + 32 * b + 24) = 0ull;
    // members of each of the first 64 colours, and the mask of full ones
    for (int k = threadIdx.x; k < 64; k += blockDim.x) sh.ccount[k] = 0;
    if (threadIdx.x == 0) sh.ival[9] = sh.ival[10] = 0;
    __syncthreads();
    int round = 0;
    while (true) {
      for (int b = threadIdx.x; b < nb; b += blockDim.x) bmax[b] = -1;
      __syncthreads();
      int left = 0;
      for (int i = threadIdx.x; i < ncons; i += blockDim.x) {
        if (col[i] != -1) continue;
        left++;
        int pr = (int)(((unsigned)i * 2654435761u) & 0x7fffffffu);   // unique hashed priority
        atomicMax(&bmax[cb1[i]], pr);
        if (cb2[i] >= 0) atomicMax(&bmax[cb2[i]], pr);
      }
      int nleft = block_sum(left, sh);
      if (nleft == 0) break;
      for (int i = threadIdx.x; i < ncons; i += blockDim.x) {
        if (col[i] != -1) continue;
        int pr = (int)(((unsigned)i * 2654435761u) & 0x7fffffffu);
        if (bmax[cb1[i]] == pr && (cb2[i] < 0 || bmax[cb2[i]] == pr)) col[i] = -3;   // a winner, colour pending
      }
      __syncthreads();
      // first fit, capped: the winners, in constraint order, each take the
      // lowest colour that neither body has taken and that has fewer than
      // blockDim.x members (a colour is then one pass of the block); past 64
      // colours, one colour of their own per round.  Colours are settled in
      // increasing order: below the lowest colour the winners would overfill
      // every proposal stands; that colour goes to its first proposers in
      // constraint order and is full, and the rest propose again.  A
      // winner's proposal only rises, and a round's winners share no body
      // (each mask has one writer per round), so this equals the sequential
      // assignment.
      while (true) {
        const unsigned long long full = ((unsigned long long)(unsigned)sh.ival[10] << 32) | (unsigned)sh.ival[9];
        for (int k = threadIdx.x; k < 64; k += blockDim.x) sh.ccur[k] = 0;
        __syncthreads();
        for (int i = threadIdx.x; i < ncons; i += blockDim.x)
          if (col[i] == -3) {
            const int p = first_free_colour(bt, cb1[i], cb2[i], full);
            if (p < 64) atomicAdd(&sh.ccur[p], 1);
          }
        __syncthreads();
        if (threadIdx.x == 0) {
          int cs = 64;
          for (int k = 0; k < 64; k++)
            if (sh.ccount[k] + sh.ccur[k] > (int)blockDim.x) { cs = k; break; }
          sh.ival[8] = cs;
        }
        __syncthreads();
        const int cs = sh.ival[8];
        for (int i = threadIdx.x; i < ncons; i += blockDim.x)
          if (col[i] == -3) {
            const int p = first_free_colour(bt, cb1[i], cb2[i], full);
            if (p < cs) take_colour(bt, cb1[i], cb2[i], col + i, p);
            else if (p >= 64) col[i] = 64 + round;
          }
        for (int k = threadIdx.x; k < cs; k += blockDim.x) sh.ccount[k] += sh.ccur[k];
        if (cs == 64) break;
        const int room = (int)blockDim.x - sh.ccount[cs];
        int base = 0;
        for (int start = 0; start < ncons && base < room; start += blockDim.x) {
          const int i = start + threadIdx.x;
          const bool want = i < ncons && col[i] == -3 && first_free_colour(bt, cb1[i], cb2[i], full) == cs;
          int tot;
          const int r = base + block_excl_scan(want ? 1 : 0, tot, sh);
          if (want && r < room) take_colour(bt, cb1[i], cb2[i], col + i, cs);
          base += tot;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
          sh.ccount[cs] = (int)blockDim.x;
          if (cs < 32) sh.ival[9] |= (int)(1u << cs);
          else sh.ival[10] |= (int)(1u << (cs - 32));
        }
        __syncthreads();
      }
      __syncthreads();
      round++;
    }
    {
      int mc = -1;
      for (int i = threadIdx.x; i < ncons; i += blockDim.x) mc = max(mc, col[i]);
      ngroups = (int)block_max((float)mc, sh) + 1;
    }
    // bucket constraints by colour, stably (index order inside a colour):
    // the sweep result does not depend on the order inside a colour, but
    // the block sums of the stopping test and the CG step do, so the
    // layout is kept reproducible from run to run
    int *cstart = c.at<int>(c.L.island);   // ngroups+1 starts
    if (ngroups <= MAX_COLORS_LDS) {
      for (int k = threadIdx.x; k <= ngroups; k += blockDim.x) sh.ccount[k] = 0;
      __syncthreads();
      for (int i = threadIdx.x; i < ncons; i += blockDim.x)
        if (col[i] >= 0) atomicAdd(&sh.ccount[col[i] + 1], 1);
      __syncthreads();
      if (threadIdx.x == 0) {
        for (int k = 0; k < ngroups; k++) sh.ccount[k + 1] += sh.ccount[k];
        for (int k = 0; k <= ngroups; k++) cstart[k] = sh.ccount[k];
      }
      __syncthreads();
      for (int k = threadIdx.x; k < ngroups; k += blockDim.x) sh.ccur[k] = sh.ccount[k];
      __syncthreads();
      for (int start = 0; start < ncons; start += blockDim.x) {
        const int i = start + threadIdx.x;
        const int ci = (i < ncons) ? col[i] : -1;
        if (ci >= 0) {
          int rank = 0;
          for (int j = start; j < i; j++) rank += (col[j] == ci);
          corder[sh.ccur[ci] + rank] = i;
        }
        __syncthreads();
        if (ci >= 0) atomicAdd(&sh.ccur[ci], 1);
        __syncthreads();
      }
    } else if (threadIdx.x == 0) {
      for (int k = 0; k <= ngroups; k++) cstart[k] = 0;
      for (int i = 0; i < ncons; i++)
        if (col[i] >= 0) cstart[col[i] + 1]++;
      for (int k = 0; k < ngroups; k++) cstart[k + 1] += cstart[k];
      int *cur = cstart + ngroups + 1;
      for (int k = 0; k < ngroups; k++) cur[k] = cstart[k];
      for (int i = 0; i < ncons; i++)
        if (col[i] >= 0) corder[cur[col[i]]++] = i;
    }
    __syncthreads();
  }
  c.clk.lap(PH_ORDER);
  if (nrows > c.W.rcap) {
    // not enough row storage: leave the world unchanged for this step (and
    // start the next solve from zero)
    if (threadIdx.x == 0) {
      atomicOr(&B.counters[c.w * CNT_N + CNT_OVERFLOW], OVF_ROWS);
      if (B.warm_n) B.warm_n[c.w] = -1;
      if (STAGE == 1) misc[MISC_SOR_READY] = 0;
    }
    __syncthreads();
    return;
  }
  // --- getInfo2 (quickstep.cpp:764-830): one constraint per lane
  const float *erp_of = ref ? (const float *)(color + nb) : nullptr;
  for (int i = threadIdx.x; i < ncons; i += blockDim.x) {
    Cons &q = cons[i];
    if (q.m <= 0) continue;
    Row *r = rows + q.row0;
    float cfm0 = P.cfm;
    for (int k = 0; k < q.m; k++) {
      row_clear(r[k], cfm0);
      r[k].b1 = q.b1;
      r[k].b2 = q.b2;
    }
    BodyPose s1 = load_pose(B, c.W.b0 + q.b1);
    BodyPose s2;
    if (q.b2 >= 0) s2 = load_pose(B, c.W.b0 + q.b2);
    const BodyPose *p2 = q.b2 >= 0 ? &s2 : nullptr;
    float erp = erp_of ? erp_of[i] : P.erp;
    if (q.kind == JT_CONTACT) {
      float3 cpos, cn;
      float cd;
      bool rev;
      Surf s;
      if (in.from_host) {
        const ContactIn &ci = B.cin[B.cin_start[c.w] + q.src];
        cpos = mk3(ci.pos[0], ci.pos[1], ci.pos[2]);
        cn = mk3(ci.normal[0], ci.normal[1], ci.normal[2]);
        cd = ci.depth;
        rev = (ci.flags & JF_REVERSE) != 0;
        s = host_surface(ci);
      } else {
        const ContactRec &cr = con[q.src];
        cpos = mk3(cr.pos[0], cr.pos[1], cr.pos[2]);
        cn = mk3(cr.normal[0], cr.normal[1], cr.normal[2]);
        cd = cr.depth;
        rev = (cr.flags & JF_REVERSE) != 0;
        s = world_surface(P);
      }
      int m = contact_info1(s);
      contact_info2(r, m, s, cpos, cn, cd, rev, s1, p2, hinv, erp, P.min_depth, P.max_vel);
    } else {
      JointDev &j = B.joints[c.W.j0 + q.src];
      float3 e1 = mk3(0.f, 0.f, 0.f), e2 = mk3(0.f, 0.f, 0.f);
      float e = erp;
      joint_info2(j, r, s1, p2, hinv, e, e1, e2);
      float *t1 = bt + 32 * q.b1 + 21;
      if (e1.x != 0.f || e1.y != 0.f || e1.z != 0.f) {
        // a motor pushing against its limit: torque applied below in order
        atomicAdd(&t1[0], e1.x); atomicAdd(&t1[1], e1.y); atomicAdd(&t1[2], e1.z);
        if (q.b2 >= 0) {
          float *t2 = bt + 32 * q.b2 + 21;
          atomicAdd(&t2[0], e2.x); atomicAdd(&t2[1], e2.y); atomicAdd(&t2[2], e2.z);
        }
      }
    }
    for (int k = 0; k < q.m; k++)
      if (r[k].findex >= 0) r[k].findex += q.row0;
  }
  __syncthreads();
  // --- rhs = c/h - J (v/h + inv(M) f_ext); cfm /= h (quickstep.cpp:832-857)
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int gb = c.W.b0 + b;
    float *t = bt + 32 * b;
    float4 ta = B.tacc[gb];
    ta.x += t[21]; ta.y += t[22]; ta.z += t[23];
    B.tacc[gb] = ta;
    float im = B.invMass[gb];
    float4 fa = B.facc[gb];
    float3 v = xyz(B.lvel[gb]), w = xyz(B.avel[gb]);
    t[15] = fa.x * im + v.x * hinv;
    t[16] = fa.y * im + v.y * hinv;
    t[17] = fa.z * im + v.z * hinv;
    float3 it = mk3(t[0] * ta.x + t[1] * ta.y + t[2] * ta.z, t[3] * ta.x + t[4] * ta.y + t[5] * ta.z,
                    t[6] * ta.x + t[7] * ta.y + t[8] * ta.z);
    t[18] = it.x + w.x * hinv;
    t[19] = it.y + w.y * hinv;
    t[20] = it.z + w.z * hinv;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nrows; i += blockDim.x) {
    Row &R = rows[i];
    const float *t1 = bt + 32 * R.b1;
    float s = 0.f;
    for (int k = 0; k < 6; k++) s += R.J[k] * t1[15 + k];
    float d = 0.f;
    float im1 = B.invMass[c.W.b0 + R.b1];
    for (int k = 0; k < 3; k++) R.iMJ[k] = im1 * R.J[k];
    for (int k = 0; k < 3; k++) R.iMJ[3 + k] = t1[3 * k] * R.J[3] + t1[3 * k + 1] * R.J[4] + t1[3 * k + 2] * R.J[5];
    for (int k = 0; k < 6; k++) d += R.iMJ[k] * R.J[k];
    if (R.b2 >= 0) {
      const float *t2 = bt + 32 * R.b2;
      for (int k = 0; k < 6; k++) s += R.J[6 + k] * t2[15 + k];
      float im2 = B.invMass[c.W.b0 + R.b2];
      for (int k = 0; k < 3; k++) R.iMJ[6 + k] = im2 * R.J[6 + k];
      for (int k = 0; k < 3; k++) R.iMJ[9 + k] = t2[3 * k] * R.J[9] + t2[3 * k + 1] * R.J[10] + t2[3 * k + 2] * R.J[11];
      for (int k = 6; k < 12; k++) d += R.iMJ[k] * R.J[k];
    } else {
      for (int k = 6; k < 12; k++) R.iMJ[k] = 0.f;
    }
    R.rhs = R.rhs * hinv - s;
    R.cfm = R.cfm * hinv;
    R.ad = P.sor_w / (d + R.cfm);
    R.lambda = 0.f;
  }
  __syncthreads();
  c.clk.lap(PH_ROWS);
  // --- SOR-LCP
  const int iters = P.iterations;
  if (ref) {
    int *ord = corder + ncons + nb;
    if (P.exact && B.lcp_ws)
      lcp_exact_world(B, c.w, rows, bt, isl, ngroups, nb, sh);
    else if (ref_sor_lds_bytes(nrows, nb) <= B.lds_bytes)
      sor_reference<1>(rows, bt, isl, ord, nrows, ngroups, nb, iters);
    else if (ref_sor_lds_bytes_streamed(nrows, nb) <= B.lds_bytes)
      sor_reference<2>(rows, bt, isl, ord, nrows, ngroups, nb, iters);
    else
      sor_reference<0>(rows, bt, isl, ord, nrows, ngroups, nb, iters);
    if (threadIdx.x == 0) B.counters[c.w * CNT_N + CNT_ITERS] = iters;
  } else if (STAGE == 1) {
    // the sweeps run in their own kernel
    sor_prepare(B, c, P, cons, rows, bt, corder, isl, nrows, ngroups, in.from_host, sh);
    if (threadIdx.x == 0) {
      misc[MISC_SOR_ROWS] = nrows;
      misc[MISC_SOR_GROUPS] = ngroups;
      misc[MISC_SOR_READY] = 1;
    }
    __syncthreads();
    return;
  } else {
    sor_colored<TOL>(B, c, P, cons, rows, bt, corder, isl, nrows, ngroups, nk, ncons, in.from_host, sh);
  }
  }
  c.clk.lap(PH_SOR);
  // --- residual (optional), feedback
  if (B.track_residual) {
    float mr = 0.f;
    for (int i = threadIdx.x; i < nrows; i += blockDim.x) mr = fmaxf(mr, row_residual(rows, i, bt));
    mr = block_max(mr, sh);
    if (threadIdx.x == 0) B.residual[c.w] = mr;
  }
  if (B.feedback) {
    for (int i = threadIdx.x; i < ncons; i += blockDim.x) {
      const Cons &q = cons[i];
      if (q.fb < 0 || q.m <= 0) continue;
      float d[12];
      for (int k = 0; k < 12; k++) d[k] = 0.f;
      for (int r = 0; r < q.m; r++) {
        const Row &R = rows[q.row0 + r];
        for (int k = 0; k < 12; k++) d[k] += R.J[k] * R.lambda;
      }
      float *o = B.feedback + 12 * (size_t)q.fb;
      for (int k = 0; k < 6; k++) o[k] = d[k];
      if (q.b2 >= 0)
        for (int k = 0; k < 6; k++) o[6 + k] = d[6 + k];
    }
  }
  // --- velocity update, integration (quickstep.cpp:905-1021, util.cpp:255-360)
  if (threadIdx.x == 0) sh.ival[13] = 0;
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    int gb = c.W.b0 + b;
    float *t = bt + 32 * b;
    float3 v = xyz(B.lvel[gb]), w = xyz(B.avel[gb]);
    v = v + h * mk3(t[9], t[10], t[11]);
    w = w + h * mk3(t[12], t[13], t[14]);
    float im = B.invMass[gb] * h;
    float4 fa = B.facc[gb], ta = B.tacc[gb];
    v = v + im * xyz(fa);
    float3 th = h * xyz(ta);
    w = w + mk3(t[0] * th.x + t[1] * th.y + t[2] * th.z, t[3] * th.x + t[4] * th.y + t[5] * th.z,
                t[6] * th.x + t[7] * th.y + t[8] * th.z);
    float3 p = xyz(B.pos[gb]) + h * v;
    float4 q = B.q[gb];
    float4 dq = make_float4(0.5f * (-w.x * q.y - w.y * q.z - w.z * q.w), 0.5f * (w.x * q.x + w.y * q.w - w.z * q.z),
                            0.5f * (-w.x * q.w + w.y * q.x + w.z * q.y), 0.5f * (w.x * q.z - w.y * q.y + w.z * q.x));
    q = make_float4(q.x + h * dq.x, q.y + h * dq.y, q.z + h * dq.z, q.w + h * dq.w);
    q = quat_normalize(q);
    B.pos[gb] = f4(p);
    B.q[gb] = q;
    store_mat(B.R + 3 * (size_t)gb, quat_to_mat(q));
    B.lvel[gb] = f4(v);
    B.avel[gb] = f4(w);
    B.facc[gb] = make_float4(0.f, 0.f, 0.f, 0.f);
    B.tacc[gb] = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!(isfinite(p.x) && isfinite(p.y) && isfinite(p.z) && isfinite(q.x) && isfinite(v.x) && isfinite(v.y) &&
          isfinite(v.z) && isfinite(w.x) && isfinite(w.y) && isfinite(w.z)))
      sh.ival[13] = 1;
  }
  // --- space list after dGeomMoved of every stepped body: moved geoms move
  //     to the front in reverse move order (collision_space.cpp:47-75, 209-213)
  if (ref) {
    // in parallel: old rank -> geom; the moved geoms in move order (bodies
    // in island visit order, each body's geoms newest first) take ranks
    // M-1..0; the others follow in their old relative order
    const int nib = misc[3];
    const int *border = corder + ncons;
    int *old = isl + 5 * (ngroups + 1);
    const int ng = c.W.ng;
    int *inv = old;                     // old rank -> geom
    int *moved = old + ng;
    int *flag = moved + ng;
    for (int g = threadIdx.x; g < ng; g += blockDim.x) {
      inv[B.grank[c.W.g0 + g]] = g;
      flag[g] = 0;
    }
    __syncthreads();
    int M = 0;
    for (int x0 = 0; x0 < nib; x0 += blockDim.x) {
      const int x = x0 + threadIdx.x;
      int k0 = 0, cnt = 0;
      if (x < nib) {
        const int gb = c.W.b0 + border[x];
        k0 = B.bgeom_start[gb];
        cnt = B.bgeom_start[gb + 1] - k0;
      }
      int tot;
      const int off = block_excl_scan(cnt, tot, sh);
      for (int j = 0; j < cnt; j++) moved[M + off + j] = B.bgeom[k0 + j];
      M += tot;
    }
    __syncthreads();
    for (int x = threadIdx.x; x < M; x += blockDim.x) {
      B.grank[c.W.g0 + moved[x]] = M - 1 - x;
      flag[moved[x]] = 1;
    }
    __syncthreads();
    int r = M;
    for (int k0 = 0; k0 < ng; k0 += blockDim.x) {
      const int k = k0 + threadIdx.x;
      const int g = k < ng ? inv[k] : -1;
      const int keep = (g >= 0 && !flag[g]) ? 1 : 0;
      int tot;
      const int off = block_excl_scan(keep, tot, sh);
      if (keep) B.grank[c.W.g0 + g] = r + off;
      r += tot;
    }
  }
  if (threadIdx.x == 0) {
    B.counters[c.w * CNT_N + CNT_ROWS] = nrows;
    B.counters[c.w * CNT_N + CNT_GROUPS] = ngroups;
  }
  __syncthreads();
  // a world whose state became non-finite is frozen (and reported) from now on
  if (threadIdx.x == 0 && sh.ival[13]) {
    P.dead = 1;
    atomicOr(&B.counters[c.w * CNT_N + CNT_OVERFLOW], OVF_NONFINITE);
  }
  c.clk.lap(PH_INTEGRATE);
}

// joint controls (steer stops, wheel torques) at the start of a step
__device__ void phase_controls(const DevBatch &B, const Ctx &c) {
  if (!B.ctrl) return;
  if (threadIdx.x == 0) {
    for (int i = 0; i < c.W.nj; i++) {
      JointDev &j = B.joints[c.W.j0 + i];
      float4 u = B.ctrl[c.W.j0 + i];
      if (!isnan(u.x)) j.lm1.lostop = u.x;
      if (!isnan(u.y)) j.lm1.histop = u.y;
      if (j.type == JT_HINGE2 && (u.z != 0.f || u.w != 0.f) && j.b1 >= 0 && j.b2 >= 0) {
        // dJointAddHinge2Torques (hinge2.cpp:393-410)
        int g1 = c.W.b0 + j.b1, g2 = c.W.b0 + j.b2;
        float3 a1 = load_mat(B.R + 3 * (size_t)g1).mul(mk3(j.axis1[0], j.axis1[1], j.axis1[2]));
        float3 a2 = load_mat(B.R + 3 * (size_t)g2).mul(mk3(j.axis2[0], j.axis2[1], j.axis2[2]));
        float3 t = u.z * a1 + u.w * a2;
        float4 t1 = B.tacc[g1], t2 = B.tacc[g2];
        t1.x += t.x; t1.y += t.y; t1.z += t.z;
        t2.x -= t.x; t2.y -= t.y; t2.z -= t.z;
        B.tacc[g1] = t1;
        B.tacc[g2] = t2;
      }
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------------ kernels

// Kernel modes (bit set): COLLIDE runs AABBs, broadphase and narrowphase;
// STEP runs the quickstep.  API_OUT hands pairs and contacts of one world to
// the host (dSpaceCollide); HOST_CONTACTS steps with contact joints the host
// created (dWorldQuickStep).  Batch steps run COLLIDE|STEP in one kernel, or
// COLLIDE then STEP as two kernels with the contact count passed through
// the world's counters.
// TOL: the coloured solver honours the world's residual tolerance (its own
// instantiation, so the iterate-to-tolerance step is a separate kernel).
enum { MF_COLLIDE = 1, MF_STEP = 2, MF_API_OUT = 4, MF_HOST_CONTACTS = 8, MF_TOL = 16, MF_PRE = 32, MF_POST = 64 };
enum {
  MODE_BATCH_COLLIDE = MF_COLLIDE,
  MODE_BATCH_STEP = MF_STEP,
  MODE_FUSED = MF_COLLIDE | MF_STEP,
  MODE_API_COLLIDE = MF_COLLIDE | MF_API_OUT,
  MODE_API_STEP = MF_STEP | MF_HOST_CONTACTS,
  MODE_BATCH_STEP_TOL = MF_STEP | MF_TOL,
  MODE_FUSED_TOL = MF_COLLIDE | MF_STEP | MF_TOL,
  MODE_FUSED_PRE = MF_COLLIDE | MF_STEP | MF_TOL | MF_PRE,   // collide and step up to the solver set-up
  MODE_POST = MF_STEP | MF_TOL | MF_POST,                    // the step after the sweeps
  KM_SOR = 1000                                               // the sweeps (sor_kernel)
};

template <int MODE>
__device__ __attribute__((always_inline)) void run_world(const DevBatch &B, int w, Shared &sh, int api_maxc) {
  Ctx c;
  c.w = w;
  c.W = B.wi[w];
  c.L = scratch_layout(c.W.nb, c.W.ng, c.W.nj, c.W.pcap, c.W.kcap, c.W.rcap);
  c.S = B.scratch + c.W.scratch;
  c.clk.start(B, w);
  WorldParams &P = B.wp[w];
  const bool ref = B.order_mode == ORDER_REFERENCE;
  constexpr bool api = (MODE & (MF_API_OUT | MF_HOST_CONTACTS)) != 0;
  if (!api && P.dead) {
    // frozen after its state became non-finite
    if (threadIdx.x == 0) B.counters[w * CNT_N + CNT_OVERFLOW] = OVF_NONFINITE;
    return;
  }
  // the first kernel of a step clears the overflow flags
  if (((MODE & MF_COLLIDE) || api) && threadIdx.x == 0) B.counters[w * CNT_N + CNT_OVERFLOW] = 0;
  if ((MODE & MF_COLLIDE) && !api) phase_controls(B, c);
  int nk = 0;
  if (MODE & MF_COLLIDE) {
    phase_aabb(B, c, P, sh);
    c.clk.lap(PH_AABB);
    int np = phase_broadphase(B, c, P, sh, ref);
    c.clk.lap(PH_BROAD);
    int maxc = api ? api_maxc : P.max_contacts;
    nk = phase_narrowphase(B, c, P, np, sh, maxc, !api);
    c.clk.lap(PH_NARROW);
    if (threadIdx.x == 0) {
      B.counters[w * CNT_N + CNT_PAIRS] = np;
      B.counters[w * CNT_N + CNT_CONTACTS] = nk;
    }
    if (MODE & MF_API_OUT) {
      // hand pairs and contacts to the host (one world per launch)
      const PairRec *pr = c.at<PairRec>(c.L.pairs);
      const int *pc = c.at<int>(c.L.pcount);
      const ContactRec *cr = c.at<ContactRec>(c.L.contacts);
      for (int i = threadIdx.x; i < np; i += blockDim.x) {
        B.pairs_out[i] = pr[i];
        B.pcount_out[i] = pc[i];
      }
      for (int i = threadIdx.x; i < nk; i += blockDim.x) B.contacts_out[i] = cr[i];
    }
  }
  if (MODE & MF_STEP) {
    StepIn in;
    in.from_host = (MODE & MF_HOST_CONTACTS) != 0;
    if (in.from_host) in.nk = B.cin_count[w];
    else if (MODE & MF_COLLIDE) in.nk = nk;
    else in.nk = B.counters[w * CNT_N + CNT_CONTACTS];
    in.seq_base = 1 << 20;
    if (in.from_host && in.nk > c.W.kcap) {
      if (threadIdx.x == 0) atomicOr(&B.counters[w * CNT_N + CNT_OVERFLOW], OVF_CONTACTS);
      in.nk = c.W.kcap;
    }
    phase_step<(MODE & MF_TOL) != 0, (MODE & MF_PRE) ? 1 : ((MODE & MF_POST) ? 2 : 0)>(B, c, P, in, sh, ref);
  }
  __syncthreads();
}

template <int MODE>
__global__ __launch_bounds__(WG) void world_kernel(DevBatch B, int first_world, int persistent, int api_maxc) {
  __shared__ Shared sh;
  if (!persistent) {
    int w = first_world + blockIdx.x;
    if (w < B.nworlds) run_world<MODE>(B, B.world_order ? B.world_order[w] : w, sh, api_maxc);
    return;
  }
  // persistent: workgroups pull worlds (largest first) from a queue
  while (true) {
    if (threadIdx.x == 0) sh.ival[15] = atomicAdd(B.work_counter, 1);
    __syncthreads();
    int k = sh.ival[15];
    __syncthreads();
    if (k >= B.nworlds) break;
    run_world<MODE>(B, B.world_order ? B.world_order[k] : k, sh, api_maxc);
  }
}

// The coloured solver's sweeps of every world (between MODE_FUSED_PRE and
// MODE_POST): a lean kernel, so that many workgroups fit on a CU.
template <bool TOL>
__device__ void sor_world(const DevBatch &B, int w, Shared &sh) {
  const WorldParams &P = B.wp[w];
  if (P.dead) return;
  Ctx c;
  c.w = w;
  c.W = B.wi[w];
  c.L = scratch_layout(c.W.nb, c.W.ng, c.W.nj, c.W.pcap, c.W.kcap, c.W.rcap);
  c.S = B.scratch + c.W.scratch;
  const int *misc = c.at<int>(c.L.misc);
  if (misc[MISC_SOR_READY] == 0) return;
  const int nrows = misc[MISC_SOR_ROWS], ngroups = misc[MISC_SOR_GROUPS];
  const int nact = c.at<int>(c.L.island)[ngroups];
  const bool accel = TOL && P.tol > 0.f && P.accel != 0;
  // a world whose working set does not fit the pool is solved by the
  // integration kernel, in global scratch
  if (sor_lds_bytes(c.W.nb, nrows, nact, accel) > B.lds_bytes) return;
  c.clk.start(B, w);
  sor_solve<TOL, SOR_IN_LDS>(B, c, P, nrows, ngroups, nact, false, sh);
  c.clk.lap(PH_SOR);
}

template <bool TOL>
__global__ __launch_bounds__(WG) void sor_kernel(DevBatch B, int first_world, int persistent) {
  __shared__ Shared sh;
  if (!persistent) {
    int w = first_world + blockIdx.x;
    if (w < B.nworlds) sor_world<TOL>(B, B.world_order ? B.world_order[w] : w, sh);
    return;
  }
  while (true) {
    if (threadIdx.x == 0) sh.ival[15] = atomicAdd(B.work_counter, 1);
    __syncthreads();
    int k = sh.ival[15];
    __syncthreads();
    if (k >= B.nworlds) break;
    sor_world<TOL>(B, B.world_order ? B.world_order[k] : k, sh);
  }
}

__global__ void reset_counter_kernel(int *ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *ctr = 0;
}

// Processing order of the persistent queue by expected cost, most expensive
// first, ties in the static order (largest first, `rank`).  The estimate of
// a world (est, kept per world) is the larger of its last step's rows times
// solver sweeps and 3/4 of its previous estimate: a tolerance solve that
// ended at the sweep cap far from the tolerance starts the next one cold,
// which may then converge quickly, so a world's cost alternates and the
// last step alone predicts the next one badly; plus bodies and geoms,
// which is all the first step has.  One workgroup; a bitonic sort of
// n <= 8192 64-bit keys (cost << 32 | position from the end of the static
// order) in the dynamic LDS (8 bytes per key, n rounded up to a power of two).
__global__ __launch_bounds__(1024) void order_worlds_kernel(DevBatch B, const int *rank, float *est, int n) {
  extern __shared__ unsigned long long okey[];
  int N = 1;
  while (N < n) N <<= 1;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    unsigned long long k = 0;   // padding sorts last
    if (i < n) {
      const int w = rank[i];
      const int *cnt = B.counters + (size_t)w * CNT_N;
      const float work = (float)max(cnt[CNT_ROWS], 0) * (float)max(cnt[CNT_ITERS], 1);
      const float e = fmaxf(work, 0.75f * est[w]);
      est[w] = e;
      const float c = e + (float)(B.wi[w].nb + B.wi[w].ng);
      const unsigned long long cost = c < 4.0e9f ? (unsigned long long)c : 0xffffffffull;
      k = (cost << 32) | (unsigned long long)(unsigned)(N - 1 - i);
    }
    okey[i] = k;
  }
  // descending bitonic sort
  for (int size = 2; size <= N; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (int i = threadIdx.x; i < (N >> 1); i += blockDim.x) {
        const int lo = 2 * i - (i & (stride - 1)), hi = lo + stride;
        const bool desc = (lo & size) == 0;
        const unsigned long long a = okey[lo], b = okey[hi];
        if ((a < b) == desc) {
          okey[lo] = b;
          okey[hi] = a;
        }
      }
    }
  __syncthreads();
  int *order = const_cast<int *>(B.world_order);
  for (int k = threadIdx.x; k < n; k += blockDim.x) order[k] = rank[N - 1 - (int)(okey[k] & 0xffffffffull)];
}

hipError_t launch_order_worlds(const DevBatch &B, const int *rank, float *est, hipStream_t st) {
  const int n = B.nworlds;
  if (n < 2) return hipSuccess;
  if (n > 8192) return hipErrorInvalidValue;
  int N = 1;
  while (N < n) N <<= 1;
  order_worlds_kernel<<<1, 1024, sizeof(unsigned long long) * (size_t)N, st>>>(B, rank, est, n);
  return hipGetLastError();
}

// one dCollide call on the device (ODE API path outside a cached batch)
__global__ void collide_one_kernel(GeomView a, GeomView b, int maxc, CPoint *out, int *n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    CPoint pts[MAX_SLOTS];
    int k = collide_pair(a, b, maxc > MAX_SLOTS ? MAX_SLOTS : maxc, pts);
    for (int i = 0; i < k; i++) out[i] = pts[i];
    *n = k;
  }
}

// the narrowphase of a list of geom pairs, one pair per lane (the
// geom-geom pairs dSpaceCollide2 reports): out holds MAX_SLOTS contacts per
// pair, n the count of each
__global__ void collide_list_kernel(const GeomView *a, const GeomView *b, int npairs, int maxc, CPoint *out,
                                    int *n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npairs) return;
  CPoint pts[MAX_SLOTS];
  const int k = collide_pair(a[i], b[i], maxc > MAX_SLOTS ? MAX_SLOTS : maxc, pts);
  for (int j = 0; j < k; j++) out[(size_t)MAX_SLOTS * i + j] = pts[j];
  n[i] = k;
}

// body-frame setup: inverse inertia is computed on the host; here the
// rotation matrices are derived from the quaternions
__global__ void init_rot_kernel(DevBatch B, int nbodies) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nbodies) store_mat(B.R + 3 * (size_t)b, quat_to_mat(B.q[b]));
}

}  // namespace odemi

// ------------------------------------------------------------------ launch API
namespace odemi {

int world_kernel_static_lds() { return (int)sizeof(Shared); }

template <int MODE>
static const void *world_kernel_fn() { return (const void *)world_kernel<MODE>; }

static const void *kernel_of(int mode) {
  switch (mode) {
    case MODE_BATCH_COLLIDE: return world_kernel_fn<MODE_BATCH_COLLIDE>();
    case MODE_BATCH_STEP: return world_kernel_fn<MODE_BATCH_STEP>();
    case MODE_FUSED: return world_kernel_fn<MODE_FUSED>();
    case MODE_API_COLLIDE: return world_kernel_fn<MODE_API_COLLIDE>();
    case MODE_API_STEP: return world_kernel_fn<MODE_API_STEP>();
    case MODE_BATCH_STEP_TOL: return world_kernel_fn<MODE_BATCH_STEP_TOL>();
    case MODE_FUSED_TOL: return world_kernel_fn<MODE_FUSED_TOL>();
    case MODE_FUSED_PRE: return world_kernel_fn<MODE_FUSED_PRE>();
    case MODE_POST: return world_kernel_fn<MODE_POST>();
    case KM_SOR: return (const void *)sor_kernel<true>;
    default: return nullptr;
  }
}

hipError_t allow_world_kernel_lds(int bytes) {
  const int modes[] = {MODE_BATCH_COLLIDE, MODE_BATCH_STEP, MODE_FUSED, MODE_API_COLLIDE, MODE_API_STEP,
                       MODE_BATCH_STEP_TOL, MODE_FUSED_TOL, MODE_FUSED_PRE, MODE_POST, KM_SOR};
  for (int m : modes) {
    hipError_t e = hipFuncSetAttribute(kernel_of(m), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

int world_kernel_occupancy(int mode, int wg, int lds_bytes) {
  int n = 0;
  const void *f = kernel_of(mode);
  if (!f) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, f, wg, (size_t)lds_bytes) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

// queue: which of the batch's work counters a persistent launch pulls from
// (kernels of one step use different counters)
hipError_t launch_world_kernel(int mode, const DevBatch &B0, int first_world, int nworlds_launch, int grid,
                               int wg, int api_maxc, hipStream_t st, int queue) {
  int persistent = grid > 0 ? 1 : 0;
  int blocks = persistent ? grid : nworlds_launch;
  if (blocks <= 0) return hipSuccess;
  if (wg > WG || wg < 64 || (wg & 63)) return hipErrorInvalidValue;
  if (B0.lds_bytes < 0 || (B0.lds_bytes & 15)) return hipErrorInvalidValue;
  if (queue < 0 || queue >= WORK_QUEUES) return hipErrorInvalidValue;
  DevBatch B = B0;
  if (persistent) {
    B.work_counter += 16 * queue;
    reset_counter_kernel<<<1, 64, 0, st>>>(B.work_counter);
  }
  size_t lds = (size_t)B.lds_bytes;
  switch (mode) {
    case MODE_BATCH_COLLIDE:
      world_kernel<MODE_BATCH_COLLIDE><<<blocks, wg, lds, st>>>(B, first_world, persistent, api_maxc);
      break;
    case MODE_BATCH_STEP:
      world_kernel<MODE_BATCH_STEP><<<blocks, wg, lds, st>>>(B, first_world, persistent, api_maxc);
      break;
    case MODE_FUSED: world_kernel<MODE_FUSED><<<blocks, wg, lds, st>>>(B, first_world, persistent, api_maxc); break;
    case MODE_API_COLLIDE:
      world_kernel<MODE_API_COLLIDE><<<blocks, wg, lds, st>>>(B, first_world, persistent, api_maxc);
      break;
    case MODE_API_STEP:
      world_kernel<MODE_API_STEP><<<blocks, wg, lds, st>>>(B, first_world, persistent, api_maxc);
      break;
    case MODE_BATCH_STEP_TOL:
      world_kernel<MODE_BATCH_STEP_TOL><<<blocks, wg, lds, st>>>(B, first_world, persistent, api_maxc);
      break;
    case MODE_FUSED_TOL:
      world_kernel<MODE_FUSED_TOL><<<blocks, wg, lds, st>>>(B, first_world, persistent, api_maxc);
      break;
    case MODE_FUSED_PRE:
      world_kernel<MODE_FUSED_PRE><<<blocks, wg, lds, st>>>(B, first_world, persistent, api_maxc);
      break;
    case MODE_POST: world_kernel<MODE_POST><<<blocks, wg, lds, st>>>(B, first_world, persistent, api_maxc); break;
    case KM_SOR: sor_kernel<true><<<blocks, wg, lds, st>>>(B, first_world, persistent); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_collide_list(const GeomView *a, const GeomView *b, int npairs, int maxc, CPoint *out, int *n,
                               hipStream_t st) {
  if (npairs <= 0) return hipSuccess;
  collide_list_kernel<<<(npairs + 63) / 64, 64, 0, st>>>(a, b, npairs, maxc, out, n);
  return hipGetLastError();
}

hipError_t launch_collide_one(const GeomView &a, const GeomView &b, int maxc, CPoint *out, int *n, hipStream_t st) {
  collide_one_kernel<<<1, 64, 0, st>>>(a, b, maxc, out, n);
  return hipGetLastError();
}

hipError_t launch_init_rot(const DevBatch &B, int nbodies, hipStream_t st) {
  if (nbodies <= 0) return hipSuccess;
  init_rot_kernel<<<(nbodies + 255) / 256, 256, 0, st>>>(B, nbodies);
  return hipGetLastError();
}

}  // namespace odemi
