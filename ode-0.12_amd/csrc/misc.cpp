// misc.cpp -- error routing, the process-global RNG of the ODE API, library
// information.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>

#include "host_common.hpp"
#include "odemi.h"

namespace {
dMessageFunction *g_error = nullptr, *g_debug = nullptr, *g_message = nullptr;
unsigned long g_seed = 0;   // ode/src/misc.cpp:31
}

namespace odemi {
// Library errors go to the installed error handler; without one they are
// printed and the call that failed returns its failure value (the reference
// would exit, ode/src/error.cpp:86-103; a GPU library must not take the
// caller's process down for a recoverable device error).
void report_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  if (g_error) {
    g_error(1, fmt, ap);
  } else {
    std::fprintf(stderr, "odemi: ");
    std::vfprintf(stderr, fmt, ap);
    std::fprintf(stderr, "\n");
  }
  va_end(ap);
}
void report_debug(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  if (g_debug) g_debug(2, fmt, ap);
  else {
    std::fprintf(stderr, "odemi debug: ");
    std::vfprintf(stderr, fmt, ap);
    std::fprintf(stderr, "\n");
  }
  va_end(ap);
}
unsigned long &global_seed() { return g_seed; }
}  // namespace odemi

extern "C" {
void dSetErrorHandler(dMessageFunction *fn) { g_error = fn; }
void dSetDebugHandler(dMessageFunction *fn) { g_debug = fn; }
void dSetMessageHandler(dMessageFunction *fn) { g_message = fn; }

// the reference's RNG (ode/src/misc.cpp:33-117)
unsigned long dRand(void) {
  g_seed = (1664525UL * g_seed + 1013904223UL) & 0xffffffffUL;
  return g_seed;
}
unsigned long dRandGetSeed(void) { return g_seed; }
void dRandSetSeed(unsigned long s) { g_seed = s & 0xffffffffUL; }
int dRandInt(int n) {
  unsigned long un = (unsigned long)n;
  unsigned long r = dRand();
  if (un <= 0x10UL) {
    r ^= r >> 16; r ^= r >> 8; r ^= r >> 4;
    if (un <= 0x2UL) { r ^= r >> 2; r ^= r >> 1; }
    else if (un <= 0x4UL) { r ^= r >> 2; }
  } else if (un <= 0x100UL) {
    r ^= r >> 16; r ^= r >> 8;
  } else if (un <= 0x10000UL) {
    r ^= r >> 16;
  }
  return (int)(r % un);
}
int dTestRand(void) {
  unsigned long old = g_seed;
  g_seed = 0;
  int ok = dRand() == 0x3c6ef35fUL && dRand() == 0x47502932UL && dRand() == 0xd1ccf6e9UL &&
           dRand() == 0xaaf95334UL && dRand() == 0x6252e503UL;
  g_seed = old;
  return ok;
}

const char *odemiVersion(void) { return "odemi 0.1 (ODE 0.12 step path, gfx950)"; }
int odemiDeviceCount(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}
}
