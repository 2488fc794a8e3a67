/* rsim_device.h: the decision stepper for user HIP kernels (observations, rewards and resets on the GPU, next to the
 * arenas), for gfx950.
 *
 * The reference's counterpart: "include cuda/sim/arena_step.cuh and call step_ticks per arena" (README.md:132-135), with
 * ArenaState / StaticData as plain structs (cuda/sim/arena_state.cuh) and the observation's fp16 conversion lf_f2h
 * (cuda/common/hd.cuh:151-176). Here the same per-arena step is rsim_dev::step_ticks, the structs are rs::ArenaState /
 * rs::StaticData (word for word the reference's), and the state buffers are librsim's AoSoA-64 buffers
 * (rsim_states_device_ptr), read and written with load_arena / store_arena.
 *
 * Build: hipcc --offload-arch=gfx950 -I <repo>/include -I <repo>/rocketsim-cuda_amd/csrc with librsim's flags
 * (-ffp-contract=off -fno-fast-math -fno-gpu-flush-denormals-to-zero -fhip-fp32-correctly-rounded-divide-sqrt,
 * -Xclang -target-feature -Xclang -packed-fp32-ops, and -fno-strict-aliasing: load_arena fills the struct through u32
 * words, so a typed read of it straight after is only defined without strict aliasing); other flags can change results.
 * Launch: one lane per arena; a kernel that steps arenas takes rsim_dev::lds_bytes(blockDim.x) bytes of dynamic shared
 * memory (the stepper's per-lane LDS slices: traversal stacks, pooled narrowphase stages, solver bank). Lanes may leave
 * early (a >= n) or step different arenas; the wave-pooled stages use whichever lanes are stepping. */
#ifndef RSIM_DEVICE_H
#define RSIM_DEVICE_H

#include "core/rs_tick.h"
#include "device/aosoa.h"

namespace rsim_dev {

using rs::ArenaState;
using rs::Controls;
using rs::StaticData;

/* Dynamic shared memory a kernel that steps arenas must be launched with. */
__host__ __device__ inline size_t lds_bytes(uint32_t block) { return rs::trav_lds_bytes(block); }

/* Arena a of an AoSoA-64 buffer into s at a tick boundary (the persistent words; the mid-tick tail zeroed), and back. */
__device__ inline void load_arena(ArenaState& s, const uint32_t* states, uint32_t a) { rs::dev::load_tick(s, states, a); }
__device__ inline void store_arena(uint32_t* states, uint32_t a, const ArenaState& s) { rs::dev::store_tick(states, a, s); }

/* n ticks of one arena with controls[0..NUM_CARS-1] set before each (cuda/sim/arena_step.cuh:83-105 step_ticks). s ends as
 * the goal state (the state after the tick the ball was first scored on) or the state after tick n. Returns the sim8
 * result word: the scored tick (RSIM_SIM8_SCORED_MASK, 0 for none) | RSIM_SIM8_UNIMPLEMENTED | RSIM_SIM8_OVERFLOW. */
__device__ inline uint32_t step_ticks(ArenaState& s, const StaticData& w, const Controls* controls, uint32_t n) {
  return rs::run_decision(s, w, controls, n, rs::NoTrace{});
}

/* IEEE binary32 -> binary16 bits, round to nearest even, in integer arithmetic (the same bits on the host and the
 * device); NaN becomes 0x7FFF, the reference's convention for observations. */
__host__ __device__ inline uint16_t f2h(float f) {
  const uint32_t x = __builtin_bit_cast(uint32_t, f), sign = (x >> 16) & 0x8000u, e = (x >> 23) & 0xFFu, m = x & 0x7FFFFFu;
  if (e == 0xFFu) return m ? (uint16_t)0x7FFF : (uint16_t)(sign | 0x7C00u);
  const int32_t E = (int32_t)e - 127;
  if (E > 15) return (uint16_t)(sign | 0x7C00u);
  if (E >= -14) {  // a normal half: 13 significand bits dropped, ties to even; a carry moves into the exponent (to inf)
    uint32_t h = ((uint32_t)(E + 15) << 10) | (m >> 13);
    const uint32_t rest = m & 0x1FFFu;
    if (rest > 0x1000u || (rest == 0x1000u && (h & 1u))) h += 1u;
    return (uint16_t)(sign | h);
  }
  if (E < -25) return (uint16_t)sign;  // below 2^-25: rounds to zero
  const uint32_t full = m | 0x800000u, shift = (uint32_t)(-E - 1);  // the value in units of 2^-24 is full >> shift
  uint32_t q = full >> shift;
  const uint32_t rem = full & ((1u << shift) - 1u), half = 1u << (shift - 1u);
  if (rem > half || (rem == half && (q & 1u))) q += 1u;  // 0x400 is the smallest normal, correctly encoded
  return (uint16_t)(sign | q);
}

}  // namespace rsim_dev

#endif
