// TEST INFRASTRUCTURE ONLY — the libm attribution build of the oracle. The product never links this file.
//
// The CPU tick compiled with RS_LIBM_MIXED: each of the seven transcendental calls chooses per call between the correctly
// rounded function and the platform C library, from a bit mask (1 sin, 2 cos, 4 tan, 8 atan2, 16 pow, 32 atan2_64,
// 64 hypot64; a set bit is the platform library). It restates the reference's mixed and platform builds
// (crates/rocketsim-cuda/csrc/phase_mixed.cpp:17-29, phase_platform.cpp:17-23): mask 0 is the correctly rounded build,
// mask 127 the platform build, a single bit the build that attributes a word to one function (tests/common/mod.rs:458-503).
// Built into its own library with hidden visibility, so its copies of the core's inline functions never mix with
// liboracle.so's.
#include <math.h>

#define RS_LIBM_MIXED 1
#include "core/rs_tick.h"

using namespace rs;

#define RS_EXPORT extern "C" __attribute__((visibility("default")))

static constexpr unsigned PLATFORM_ALL = 127u;

// phase_mixed.cpp:17-20 rscuda_phase_mixed
RS_EXPORT int rsim_cpu_phase_mixed(int k, ArenaState* s, const StaticData* w, unsigned mask) {
  libm_platform_mask = mask;
  return run_phase(k, *s, *w);
}
// phase_mixed.cpp:21-24 rscuda_step_tick_mixed
RS_EXPORT int rsim_cpu_step_tick_mixed(ArenaState* s, const StaticData* w, unsigned mask) {
  libm_platform_mask = mask;
  return step_tick(*s, *w);
}
// phase_mixed.cpp:25-29 rscuda_step_ticks_mixed (the arguments of rsim_cpu_step_ticks, plus the mask)
RS_EXPORT int rsim_cpu_step_ticks_mixed(ArenaState* s, const StaticData* w, const Controls* controls, uint32_t n,
                                        uint32_t* scored_tick, ArenaState* trace, unsigned mask) {
  libm_platform_mask = mask;
  uint32_t r = run_decision(*s, *w, controls, n, [trace](uint32_t t, const ArenaState& st) {
    if (trace) trace[t - 1] = st;
  });
  *scored_tick = r & SIM8_SCORED_MASK;
  return (r & SIM8_UNIMPLEMENTED) ? -1 : 0;
}
// phase_platform.cpp:17-23: the platform build is every bit set
RS_EXPORT int rsim_cpu_phase_platform(int k, ArenaState* s, const StaticData* w) {
  return rsim_cpu_phase_mixed(k, s, w, PLATFORM_ALL);
}
RS_EXPORT int rsim_cpu_step_tick_platform(ArenaState* s, const StaticData* w) {
  return rsim_cpu_step_tick_mixed(s, w, PLATFORM_ALL);
}
// The function behind bit `fn` (0 sin ... 6 hypot64) under `mask`, on one input (the f32 functions take and return values
// exactly representable in f32): checks that a bit routes exactly one call.
RS_EXPORT double rsim_cpu_libm_mixed(int fn, unsigned mask, double x, double y) {
  libm_platform_mask = mask;
  switch (fn) {
    case 0: return fsin((float)x);
    case 1: return fcos((float)x);
    case 2: return ftan((float)x);
    case 3: return fatan2((float)x, (float)y);
    case 4: return fpow((float)x, (float)y);
    case 5: return datan2(x, y);
    case 6: return dhypot(x, y);
    default: return 0.0;
  }
}
