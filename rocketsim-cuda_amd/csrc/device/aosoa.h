// Copies between an AoSoA-64 state buffer and a lane's private ArenaState (word w of arena a at
// (a / 64) * ARENA_WORDS * 64 + w * 64 + a % 64): the kernels' copy-in and copy-out, and the same for user kernels
// (include/rsim_device.h). Reference: cuda/sim/aosoa.cuh:10-41.
#pragma once
#include "device/launch.h"

namespace rs {
namespace dev {

RS_HD inline size_t block_base(uint32_t a) { return (size_t)(a / AOSOA_LANES) * ARENA_WORDS * AOSOA_LANES + (a % AOSOA_LANES); }

__device__ inline void load_all(ArenaState& s, const uint32_t* __restrict__ buf, uint32_t a) {
  uint32_t* d = reinterpret_cast<uint32_t*>(&s);
  const uint32_t* src = buf + block_base(a);
  for (int w = 0; w < ARENA_WORDS; ++w) d[w] = src[(size_t)w * AOSOA_LANES];
}
__device__ inline void store_all(uint32_t* __restrict__ buf, uint32_t a, const ArenaState& s) {
  const uint32_t* d = reinterpret_cast<const uint32_t*>(&s);
  uint32_t* dst = buf + block_base(a);
  for (int w = 0; w < ARENA_WORDS; ++w) dst[(size_t)w * AOSOA_LANES] = d[w];
}
// Tick-boundary copies: the TICK_WORDS persistent words move; the mid-tick tail is zero between ticks (arena_state.cuh:221).
// The buffer side is read once and written once per launch: non-temporal, so the stream does not evict what the ticks
// reuse from the caches (the static data, the lanes' private memory).
__device__ inline void load_tick(ArenaState& s, const uint32_t* __restrict__ buf, uint32_t a) {
  uint32_t* d = reinterpret_cast<uint32_t*>(&s);
  const uint32_t* src = buf + block_base(a);
  for (int w = 0; w < TICK_WORDS; ++w) d[w] = __builtin_nontemporal_load(src + (size_t)w * AOSOA_LANES);
  for (int w = TICK_WORDS; w < ARENA_WORDS; ++w) d[w] = 0;
}
__device__ inline void store_tick(uint32_t* __restrict__ buf, uint32_t a, const ArenaState& s) {
  const uint32_t* d = reinterpret_cast<const uint32_t*>(&s);
  uint32_t* dst = buf + block_base(a);
  for (int w = 0; w < TICK_WORDS; ++w) __builtin_nontemporal_store(d[w], dst + (size_t)w * AOSOA_LANES);
}


}  // namespace dev
}  // namespace rs
