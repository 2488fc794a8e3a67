// Launchers of the gfx950 kernels (sim_kernels.hip), called by the C ABI (host/rsim.cpp). All are asynchronous on `st`.
#pragma once
#include <hip/hip_runtime.h>

#include "core/rs_tick.h"

namespace rs {
namespace dev {

constexpr uint32_t AOSOA_LANES = 64;  // arenas per AoSoA block: one wave64
constexpr uint32_t SIM8_BLOCK = 64;   // production work-group: one wave, one AoSoA block
constexpr uint32_t MAX_BLOCK = 256;   // the largest work-group the kernels are compiled for

RS_HD inline size_t aosoa_words(uint32_t n) { return (size_t)((n + AOSOA_LANES - 1) / AOSOA_LANES) * ARENA_WORDS * AOSOA_LANES; }
RS_HD inline size_t aosoa_index(uint32_t a, uint32_t w) {
  return (size_t)(a / AOSOA_LANES) * ARENA_WORDS * AOSOA_LANES + (size_t)w * AOSOA_LANES + (a % AOSOA_LANES);
}

hipError_t launch_identity(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n, bool tick_only);
hipError_t launch_copy_words(hipStream_t st, const uint32_t* in, uint32_t* out, size_t words);  // words % 4 == 0
hipError_t launch_phase(hipStream_t st, int k, uint32_t* buf, const StaticData* w, uint32_t n, int32_t* status);
hipError_t launch_step_tick(hipStream_t st, uint32_t* buf, const StaticData* w, uint32_t n, int32_t* status);
hipError_t launch_sim8(hipStream_t st, uint32_t* buf, const StaticData* w, const Controls* ctl, uint32_t* result, uint32_t n,
                       uint32_t block, uint32_t* trace);

}  // namespace dev
}  // namespace rs
