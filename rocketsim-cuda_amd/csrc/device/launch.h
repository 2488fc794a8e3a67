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
// The result-word bit of an arena whose variant index is past the world table (the arena is not stepped).
constexpr uint32_t SIM8_BAD_VARIANT = 0x20000000u;

struct Sim8Args {
  uint32_t* buf;             // AoSoA-64 states, n arenas, stepped in place
  const StaticData* worlds;  // nworlds StaticData back to back
  uint32_t nworlds;
  const uint32_t* variant;   // null, or n world indices
  const Controls* ctl;       // Controls[n][NUM_CARS]
  uint32_t* result;          // n result words
  uint32_t n;
  uint32_t* trace;           // null, or TICKS_PER_DECISION slabs of aosoa_words(n)
  uint64_t* cycles;          // null, or n: the lane's wall-clock ticks (100 MHz) for its decision
};
hipError_t launch_sim8(hipStream_t st, const Sim8Args& p, uint32_t block);
hipError_t launch_run(hipStream_t st, const Sim8Args& p, uint32_t decisions, uint32_t block);  // ctl/result: per decision
// One decision as ceil(8 / ticks_per_launch) launches of k_sim8_ticks (no trace, no cycles).
hipError_t launch_sim8_ticks(hipStream_t st, const Sim8Args& p, uint32_t block, uint32_t ticks_per_launch);
// k_run_queue: `workgroups` persistent work-groups (block a multiple of 64), the queue counter `next` reset on the stream
// first; order: null or a permutation of the ceil(n / 64) block indices.
hipError_t launch_run_queue(hipStream_t st, const Sim8Args& p, uint32_t decisions, uint32_t block, const uint32_t* order,
                            uint32_t* next, uint32_t workgroups);
// The largest static LDS (group segment of fixed size) of the kernels that step arenas: their device functions address the
// dynamic LDS from LDS address 0, which holds only while this is 0.
hipError_t stepping_static_lds(size_t* bytes);
// How many work-groups of `block` lanes of k_run_queue are resident at once on the current GPU (0 if unknown).
uint32_t run_queue_capacity(uint32_t block);

}  // namespace dev
}  // namespace rs
