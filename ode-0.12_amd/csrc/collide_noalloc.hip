// collide_noalloc.hip -- the device memory of dCollideNoAlloc: the contacts
// of one pair and their count live in a file-scope device array (one copy per
// GPU, held until the process ends), so that the entry point itself never
// allocates device memory.  Every call on one GPU shares that GPU's copy:
// calls that overlap on one GPU (from several host threads) race.
#include <hip/hip_runtime.h>

#include "colliders.hpp"
#include "layout.hpp"

namespace {

// one pair's contacts (at most MAX_SLOTS, the most a collider produces) and
// their count, written by the one-pair collide kernel on every launch
struct OnePairScratch {
  odemi::CPoint pts[odemi::MAX_SLOTS];
  int n;
};

}  // namespace

static __device__ __align__(256) OnePairScratch scratch;

namespace odemi {

hipError_t launch_collide_one(const GeomView &a, const GeomView &b, int maxc, CPoint *out, int *n, hipStream_t st);

namespace {

constexpr int kMaxDevices = 64;
// device address of `scratch` on each GPU (the array's host-side name is not
// a device address); null until looked up
void *g_scratch_addr[kMaxDevices];

}  // namespace

// Looks up this GPU's copy of the scratch array once (called from
// dInitODE2 for the current device, and on a device's first call).
// Returns its device address, or null with the HIP error in *err.
void *noalloc_scratch(hipError_t *err) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    if (err) *err = e != hipSuccess ? e : hipErrorInvalidDevice;
    return nullptr;
  }
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      if (err) *err = e != hipSuccess ? e : hipErrorInvalidSymbol;
      return nullptr;
    }
    g_scratch_addr[dev] = p;
  }
  if (err) *err = hipSuccess;
  return g_scratch_addr[dev];
}

// The one-pair collide into the scratch at `addr` (from noalloc_scratch),
// then the count and the contacts copied to the host (stream-ordered,
// synchronised).  The kernel writes the count on every launch, so nothing
// needs clearing between calls.
hipError_t collide_one_noalloc(void *addr, const GeomView &a, const GeomView &b, int maxc, CPoint *host_pts,
                               int *host_n, hipStream_t st) {
  OnePairScratch *s = static_cast<OnePairScratch *>(addr);
  hipError_t e = launch_collide_one(a, b, maxc, s->pts, &s->n, st);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(host_n, &s->n, sizeof(int), hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(host_pts, s->pts, sizeof(CPoint) * MAX_SLOTS, hipMemcpyDeviceToHost, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

}  // namespace odemi
