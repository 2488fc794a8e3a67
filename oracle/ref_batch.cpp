// TEST INFRASTRUCTURE ONLY. A batch driver over the reference's own CPU decision (rscuda_step_ticks,
// crates/rocketsim-cuda/csrc/phase.cpp:15-17, compiled in place into _ref/libref.so beside this file): arena ranges on
// host threads, each arena's result word formed as the reference's sim8 forms it (the scored tick, the unimplemented
// flag when the decision returned a non-OK status, the overflow flag when an overflow counter of the state is set;
// cuda/sim/arena_step.cuh:16-25). It times the reference's CPU path beside the GPU kernel and gives the words the
// GPU's results are compared with.
#include <stdint.h>

#include <thread>
#include <vector>

extern "C" int rscuda_step_ticks(void* s, const void* sd, const void* controls, uint32_t n, uint32_t* scored_tick, void* trace);
extern "C" uint64_t lf_arena_words();

namespace {
constexpr uint32_t ARENA_WORDS = 2784, OVERFLOW_WORD = 2560, OVERFLOW_COUNT = 6, CONTROL_WORDS = 16, TICKS = 8;

void run(uint32_t* states, const void* sd, const uint32_t* controls, uint32_t* result, uint32_t lo, uint32_t hi) {
  for (uint32_t a = lo; a < hi; ++a) {
    uint32_t* s = states + (size_t)a * ARENA_WORDS;
    uint32_t scored = 0;
    int status = rscuda_step_ticks(s, sd, controls + (size_t)a * CONTROL_WORDS, TICKS, &scored, nullptr);
    bool overflow = false;
    for (uint32_t k = 0; k < OVERFLOW_COUNT; ++k) overflow = overflow || s[OVERFLOW_WORD + k] != 0;
    result[a] = (scored & 0xFFu) | (status != 0 ? 0x80000000u : 0u) | (overflow ? 0x40000000u : 0u);
  }
}
}  // namespace

// One 8-tick decision of arenas 0..n-1 (ArenaState words, in place) over `threads` host threads; returns 0, or -1 when
// the reference's ArenaState is not the size this driver assumes.
extern "C" int refb_sim8_batch(uint32_t* states, const void* sd, const uint32_t* controls, uint32_t* result, uint32_t n,
                               uint32_t threads) {
  if (lf_arena_words() != ARENA_WORDS) return -1;
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < threads; ++t) {
    uint32_t lo = (uint32_t)((uint64_t)n * t / threads), hi = (uint32_t)((uint64_t)n * (t + 1) / threads);
    pool.emplace_back(run, states, sd, controls, result, lo, hi);
  }
  for (auto& th : pool) th.join();
  return 0;
}
