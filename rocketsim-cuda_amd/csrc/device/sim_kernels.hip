// The gfx950 kernels of the decision path: one lane per arena over an AoSoA-64 state buffer.
//
// Word w of arena a sits at (a / 64) * ARENA_WORDS * 64 + w * 64 + a % 64, so the 64 lanes of a wave that copy their arenas'
// word w touch one contiguous 256-byte run: every copy-in and copy-out instruction is one fully coalesced HBM transaction.
// Each lane steps its arena in a private ArenaState (the physics core of rs_tick.h, the same source the CPU oracle compiles);
// no lane reads another's arena, so neither the work-group size nor the arena order can change a result.
//
// Kernels (reference cuda/sim/sim.cu:8-112):
//   k_identity / k_identity_tick   copy-in and copy-out only (the layout check)
//   k_phase                         one phase k of the tick on whole states (per-phase parity)
//   k_step_tick                     one full tick at a tick boundary
//   k_sim8 / k_sim8_trace           one decision: 8 ticks, stopping after the scoring tick
//   k_run                           several decisions per launch, the arena held in the lane from first to last
//   k_copy_words                    a streaming copy of a state buffer (restores between timed runs)
#include "device/launch.h"
#include "device/aosoa.h"

namespace rs {
namespace dev {

__global__ void __launch_bounds__(MAX_BLOCK) k_identity(const uint32_t* in, uint32_t* out, uint32_t n) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  ArenaState s;
  load_all(s, in, a);
  store_all(out, a, s);
}
__global__ void __launch_bounds__(MAX_BLOCK) k_identity_tick(const uint32_t* in, uint32_t* out, uint32_t n) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  ArenaState s;
  load_tick(s, in, a);
  store_tick(out, a, s);
}
__global__ void __launch_bounds__(MAX_BLOCK) k_phase(int k, uint32_t* buf, const StaticData* __restrict__ w, uint32_t n,
                                                     int32_t* status) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  ArenaState s;
  load_all(s, buf, a);
  status[a] = run_phase(k, s, *w);
  store_all(buf, a, s);
}
__global__ void __launch_bounds__(MAX_BLOCK) k_step_tick(uint32_t* buf, const StaticData* __restrict__ w, uint32_t n,
                                                         int32_t* status) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  ArenaState s;
  load_tick(s, buf, a);
  status[a] = step_tick(s, *w);
  store_tick(buf, a, s);
}

struct TraceOut {  // the tick-boundary words after tick t into slab t - 1 of the trace
  uint32_t* trace;
  size_t slab_words;
  uint32_t a;
  __device__ void operator()(uint32_t t, const ArenaState& s) const { store_tick(trace + (size_t)(t - 1) * slab_words, a, s); }
};

// One decision of every arena. With a variant array, arena a steps in worlds[variant[a]] (an index past the table leaves
// the arena untouched and flags its result word); without one, every arena steps in worlds[0].
template <bool TRACE, bool VARIANTS>
__global__ void __launch_bounds__(MAX_BLOCK) k_sim8(Sim8Args p) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= p.n) return;
  uint32_t v = 0;
  if (VARIANTS) {
    v = p.variant[a];
    if (v >= p.nworlds) {
      p.result[a] = SIM8_BAD_VARIANT;
      return;
    }
  }
  const StaticData& w = p.worlds[v];  // without variants the same world for every lane: a uniform pointer
  const uint64_t t0 = p.cycles ? wall_clock64() : 0;
  ArenaState s;
  load_tick(s, p.buf, a);
  Controls c[NUM_CARS];
  for (int i = 0; i < NUM_CARS; ++i) c[i] = p.ctl[(size_t)a * NUM_CARS + i];
  if (TRACE) p.result[a] = run_decision(s, w, c, TICKS_PER_DECISION, TraceOut{p.trace, aosoa_words(p.n), a});
  else p.result[a] = run_decision(s, w, c, TICKS_PER_DECISION, NoTrace{});
  store_tick(p.buf, a, s);
  if (p.cycles) p.cycles[a] = wall_clock64() - t0;  // a wave's lanes share its clock: with one lane per work-group
                                                     // (block 1) this is the arena's own cost
}

// Several decisions in one launch: each lane keeps its arena in its private copy from the first decision to the last
// (one copy-in, one copy-out), with the controls of decision d at ctl[d * n * NUM_CARS ...] and its result words at
// result[d * n ...]. Decision d + 1 starts where decision d ended (the goal state after a goal), as consecutive sim8
// launches do; the mid-tick tail is cleared between decisions, as the tick-boundary copy clears it.
__global__ void __launch_bounds__(MAX_BLOCK) k_run(Sim8Args p, uint32_t decisions) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= p.n) return;
  uint32_t v = p.variant ? p.variant[a] : 0u;
  if (v >= p.nworlds) {
    for (uint32_t d = 0; d < decisions; ++d) p.result[(size_t)d * p.n + a] = SIM8_BAD_VARIANT;
    return;
  }
  const StaticData& w = p.worlds[v];
  ArenaState s;
  load_tick(s, p.buf, a);
  uint32_t* tail = reinterpret_cast<uint32_t*>(&s);
  for (uint32_t d = 0; d < decisions; ++d) {
    if (d > 0)
      for (int i = TICK_WORDS; i < ARENA_WORDS; ++i) tail[i] = 0;
    Controls c[NUM_CARS];
    const Controls* cd = p.ctl + ((size_t)d * p.n + a) * NUM_CARS;
    for (int i = 0; i < NUM_CARS; ++i) c[i] = cd[i];
    p.result[(size_t)d * p.n + a] = run_decision(s, w, c, TICKS_PER_DECISION, NoTrace{});
  }
  store_tick(p.buf, a, s);
}

// A plain streaming copy of whole state buffers (16-byte vectors, grid-stride): the restore between timed runs, and the
// reference point for measured HBM bytes (it moves exactly 2 x its size).
__global__ void __launch_bounds__(256) k_copy_words(const uint4* __restrict__ in, uint4* __restrict__ out, size_t n16) {
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) out[i] = in[i];
}

// ---- launchers ------------------------------------------------------------------------------------------------------------------
static dim3 grid_of(uint32_t n, uint32_t block) { return dim3((n + block - 1) / block > 0 ? (n + block - 1) / block : 1); }

hipError_t launch_identity(hipStream_t st, const uint32_t* in, uint32_t* out, uint32_t n, bool tick_only) {
  if (tick_only) k_identity_tick<<<grid_of(n, SIM8_BLOCK), SIM8_BLOCK, 0, st>>>(in, out, n);
  else k_identity<<<grid_of(n, SIM8_BLOCK), SIM8_BLOCK, 0, st>>>(in, out, n);
  return hipGetLastError();
}
hipError_t launch_copy_words(hipStream_t st, const uint32_t* in, uint32_t* out, size_t words) {
  size_t n16 = words / 4;  // AoSoA buffers are whole blocks of 2784 x 64 words: a multiple of four
  size_t blocks = (n16 + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks == 0) blocks = 1;
  k_copy_words<<<dim3((unsigned)blocks), 256, 0, st>>>(reinterpret_cast<const uint4*>(in), reinterpret_cast<uint4*>(out), n16);
  return hipGetLastError();
}
hipError_t launch_phase(hipStream_t st, int k, uint32_t* buf, const StaticData* w, uint32_t n, int32_t* status) {
  k_phase<<<grid_of(n, SIM8_BLOCK), SIM8_BLOCK, trav_lds_bytes(SIM8_BLOCK), st>>>(k, buf, w, n, status);
  return hipGetLastError();
}
hipError_t launch_step_tick(hipStream_t st, uint32_t* buf, const StaticData* w, uint32_t n, int32_t* status) {
  k_step_tick<<<grid_of(n, SIM8_BLOCK), SIM8_BLOCK, trav_lds_bytes(SIM8_BLOCK), st>>>(buf, w, n, status);
  return hipGetLastError();
}
hipError_t launch_run(hipStream_t st, const Sim8Args& p, uint32_t decisions, uint32_t block) {
  k_run<<<grid_of(p.n, block), block, trav_lds_bytes(block), st>>>(p, decisions);
  return hipGetLastError();
}
hipError_t launch_sim8(hipStream_t st, const Sim8Args& p, uint32_t block) {
  const size_t lds = trav_lds_bytes(block);  // the BVH walks' traversal stacks
  const dim3 g = grid_of(p.n, block);
  if (p.trace && p.variant) k_sim8<true, true><<<g, block, lds, st>>>(p);
  else if (p.trace) k_sim8<true, false><<<g, block, lds, st>>>(p);
  else if (p.variant) k_sim8<false, true><<<g, block, lds, st>>>(p);
  else k_sim8<false, false><<<g, block, lds, st>>>(p);
  return hipGetLastError();
}

}  // namespace dev
}  // namespace rs
