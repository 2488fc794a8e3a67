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
//   k_sim8_ticks                    one decision as several launches of a few ticks (the state in HBM between them)
//   k_run                           several decisions per launch, the arena held in the lane from first to last
//   k_run_queue                     k_run with persistent work-groups taking AoSoA blocks from a queue
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
  status[a] = run_phase<true>(k, s, *w);
  store_all(buf, a, s);
}
__global__ void __launch_bounds__(MAX_BLOCK) k_step_tick(uint32_t* buf, const StaticData* __restrict__ w, uint32_t n,
                                                         int32_t* status) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  ArenaState s;
  load_tick(s, buf, a);
  status[a] = step_tick<true>(s, *w);
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
  // without a variant array every lane steps in worlds[0]: the uniform-world instantiation
  if (TRACE) p.result[a] = run_decision<!VARIANTS>(s, w, c, TICKS_PER_DECISION, TraceOut{p.trace, aosoa_words(p.n), a});
  else p.result[a] = run_decision<!VARIANTS>(s, w, c, TICKS_PER_DECISION, NoTrace{});
  store_tick(p.buf, a, s);
  if (p.cycles) p.cycles[a] = wall_clock64() - t0;  // a wave's lanes share its clock: with one lane per work-group
                                                     // (block 1) this is the arena's own cost
}

// One decision split over launches: this launch runs ticks first..first+count-1 (1-based) of it, the state going back to
// the AoSoA buffer in between (the per-tick split of state between HBM and the lane). result[a] carries the decision's
// word from launch to launch: the launch of tick 1 starts it, a later launch skips an arena whose word already holds its
// scored tick (the decision ended there), and the overflow bit joins it with the state the decision ends in. Same words
// and states as k_sim8.
template <bool VARIANTS>
__global__ void __launch_bounds__(MAX_BLOCK) k_sim8_ticks(Sim8Args p, uint32_t first, uint32_t count) {
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
  uint32_t word = first == 1 ? 0u : p.result[a];
  if (word & SIM8_SCORED_MASK) return;
  const StaticData& w = p.worlds[v];
  ArenaState s;
  load_tick(s, p.buf, a);
  Controls c[NUM_CARS];
  for (int i = 0; i < NUM_CARS; ++i) c[i] = p.ctl[(size_t)a * NUM_CARS + i];
  const uint32_t last = first + count - 1;
  for (uint32_t t = first; t <= last; ++t) {
    for (int ci = 0; ci < NUM_CARS; ++ci) s.cars[ci].controls = c[ci];
    if (step_tick<!VARIANTS>(s, w) != PHASE_OK) word |= SIM8_UNIMPLEMENTED;
    if (ball_scored(s, w)) {
      word |= t;
      break;
    }
  }
  if ((word & SIM8_SCORED_MASK) || last == TICKS_PER_DECISION) word |= any_overflow(s) ? SIM8_OVERFLOW : 0u;
  store_tick(p.buf, a, s);
  p.result[a] = word;
}

// Several decisions in one launch: each lane keeps its arena in its private copy from the first decision to the last
// (one copy-in, one copy-out), with the controls of decision d at ctl[d * n * NUM_CARS ...] and its result words at
// result[d * n ...]. Decision d + 1 starts where decision d ended (the goal state after a goal), as consecutive sim8
// launches do; the mid-tick tail is cleared between decisions, as the tick-boundary copy clears it.
template <bool UW>  // UW: no variant array, every lane in worlds[0]
__device__ inline void run_arena(const Sim8Args& p, uint32_t a, uint32_t decisions) {
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
    p.result[(size_t)d * p.n + a] = run_decision<UW>(s, w, c, TICKS_PER_DECISION, NoTrace{});
  }
  store_tick(p.buf, a, s);
}
__global__ void __launch_bounds__(MAX_BLOCK) k_run(Sim8Args p, uint32_t decisions) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= p.n) return;
  if (p.variant) run_arena<false>(p, a, decisions);
  else run_arena<true>(p, a, decisions);
}

// k_run with dynamic scheduling: persistent work-groups whose waves each take the next AoSoA block of 64 arenas from a
// queue (*next counts the blocks handed out; it is zero when the launch starts) and run all `decisions` decisions of it,
// until the queue is empty. The queue holds the blocks in the order order[0..], a permutation of the block indices
// (heaviest first when the order comes from predicted costs), or in index order without one. A wave holds one block at a
// time and no wave waits on another, so the launch drains whatever the number of resident work-groups. The block size is
// a multiple of 64 (whole waves).
__global__ void __launch_bounds__(MAX_BLOCK) k_run_queue(Sim8Args p, uint32_t decisions, const uint32_t* order, uint32_t* next) {
  const uint32_t lane = threadIdx.x & (AOSOA_LANES - 1);
  const uint32_t nblocks = (p.n + AOSOA_LANES - 1) / AOSOA_LANES;
  for (;;) {
    uint32_t u = 0;
    if (lane == 0) u = atomicAdd(next, 1u);
    u = __builtin_amdgcn_readfirstlane(u);  // lane 0 is the wave's first active lane here: every lane loops to this point
    if (u >= nblocks) break;
    const uint32_t blk = order ? order[u] : u;
    const uint32_t a = blk * AOSOA_LANES + lane;
    if (blk < nblocks && a < p.n) {
      if (p.variant) run_arena<false>(p, a, decisions);
      else run_arena<true>(p, a, decisions);
    }
  }
}
__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
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
hipError_t launch_sim8_ticks(hipStream_t st, const Sim8Args& p, uint32_t block, uint32_t ticks_per_launch) {
  const size_t lds = trav_lds_bytes(block);
  const dim3 g = grid_of(p.n, block);
  for (uint32_t first = 1; first <= TICKS_PER_DECISION; first += ticks_per_launch) {
    const uint32_t count = first + ticks_per_launch - 1 <= TICKS_PER_DECISION ? ticks_per_launch : TICKS_PER_DECISION - first + 1;
    if (p.variant) k_sim8_ticks<true><<<g, block, lds, st>>>(p, first, count);
    else k_sim8_ticks<false><<<g, block, lds, st>>>(p, first, count);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}
hipError_t stepping_static_lds(size_t* bytes) {
  const void* kernels[] = {(const void*)k_phase,        (const void*)k_step_tick,           (const void*)k_sim8<false, false>,
                           (const void*)k_sim8<false, true>, (const void*)k_sim8<true, false>, (const void*)k_sim8<true, true>,
                           (const void*)k_sim8_ticks<false>, (const void*)k_sim8_ticks<true>, (const void*)k_run,
                           (const void*)k_run_queue};
  size_t most = 0;
  for (const void* k : kernels) {
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, k);
    if (e != hipSuccess) return e;
    most = a.sharedSizeBytes > most ? a.sharedSizeBytes : most;
  }
  *bytes = most;
  return hipSuccess;
}
uint32_t run_queue_capacity(uint32_t block) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_run_queue, (int)block, trav_lds_bytes(block)) != hipSuccess)
    return 0;
  return (uint32_t)(cus > 0 ? cus : 0) * (uint32_t)(per_cu > 0 ? per_cu : 1);
}
hipError_t launch_run_queue(hipStream_t st, const Sim8Args& p, uint32_t decisions, uint32_t block, const uint32_t* order,
                            uint32_t* next, uint32_t workgroups) {
  k_fill_u32<<<1, 64, 0, st>>>(next, 1, 0u);
  k_run_queue<<<dim3(workgroups), block, trav_lds_bytes(block), st>>>(p, decisions, order, next);
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
