// TEST INFRASTRUCTURE ONLY — the CPU oracle. The product never links this file.
//
// The CPU build of the single-source tick (rocketsim-cuda_amd/csrc/core), behind the C ABI of the reference's host build of
// the same path (crates/rocketsim-cuda/csrc/phase.cpp:8-40: rscuda_phase, rscuda_step_tick, rscuda_step_ticks, lf_*). Tests
// pin it bit for bit to the reference's own C++ build (oracle/_ref) on golden vectors, then check the gfx950 kernels against it.
// Built by oracle/Makefile with g++ -O2 -ffp-contract=off (no FMA target).
#include <thread>
#include <vector>

#include "core/rs_tick.h"

using namespace rs;

extern "C" {
// phase.cpp:10 rscuda_phase
int rsim_cpu_phase(int k, ArenaState* s, const StaticData* w) { return run_phase(k, *s, *w); }
// phase.cpp:12 rscuda_step_tick
int rsim_cpu_step_tick(ArenaState* s, const StaticData* w) { return step_tick(*s, *w); }
// phase.cpp:15-17 rscuda_step_ticks: n ticks with controls[0..1] set before each; *s ends as the goal state (scored tick in
// *scored_tick, 0 for none) or the state after tick n; trace (null or n states): trace[t-1] = the state after tick t up to the
// goal. Returns 0 (PHASE_OK) or -1 (PHASE_UNIMPLEMENTED).
int rsim_cpu_step_ticks(ArenaState* s, const StaticData* w, const Controls* controls, uint32_t n, uint32_t* scored_tick,
                        ArenaState* trace) {
  uint32_t r = run_decision(*s, *w, controls, n, [trace](uint32_t t, const ArenaState& st) {
    if (trace) trace[t - 1] = st;
  });
  *scored_tick = r & SIM8_SCORED_MASK;
  return (r & SIM8_UNIMPLEMENTED) ? PHASE_UNIMPLEMENTED : PHASE_OK;
}
// The result word of one decision for each of n arenas (AoS, in place): the CPU baseline of the 8-tick decision.
void rsim_cpu_sim8_aos(ArenaState* states, const StaticData* w, const Controls* controls, uint32_t* result, uint32_t n) {
  for (uint32_t a = 0; a < n; ++a) result[a] = run_decision(states[a], *w, controls + 2 * (size_t)a, TICKS_PER_DECISION, NoTrace{});
}
// The same over `threads` host threads (arena ranges split evenly); the CPU baseline of the benchmark.
void rsim_cpu_sim8_batch(ArenaState* states, const StaticData* w, const Controls* controls, uint32_t* result, uint32_t n,
                         uint32_t threads) {
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < threads; ++t) {
    uint32_t lo = (uint32_t)((uint64_t)n * t / threads), hi = (uint32_t)((uint64_t)n * (t + 1) / threads);
    pool.emplace_back([=] { rsim_cpu_sim8_aos(states + lo, w, controls + 2 * (size_t)lo, result + lo, hi - lo); });
  }
  for (auto& th : pool) th.join();
}
// rsim_cpu_sim8_batch with arena a stepped in worlds[variant[a]] (the CPU side of rsim_sim8_variants / rsim_run with a
// variant table); an index past the table leaves the arena untouched with RSIM_SIM8_BAD_VARIANT, as on the device.
void rsim_cpu_sim8_batch_in(ArenaState* states, const StaticData* worlds, uint32_t count, const uint32_t* variant,
                            const Controls* controls, uint32_t* result, uint32_t n, uint32_t threads) {
  if (threads < 1) threads = 1;
  std::vector<std::thread> pool;
  for (uint32_t t = 0; t < threads; ++t) {
    uint32_t lo = (uint32_t)((uint64_t)n * t / threads), hi = (uint32_t)((uint64_t)n * (t + 1) / threads);
    pool.emplace_back([=] {
      for (uint32_t a = lo; a < hi; ++a)
        result[a] = variant[a] < count ? run_decision(states[a], worlds[variant[a]], controls + 2 * (size_t)a, TICKS_PER_DECISION,
                                                      NoTrace{})
                                       : 0x20000000u;
    });
  }
  for (auto& th : pool) th.join();
}
// Byte layout of the mesh records and BVH nodes in StaticData (for tests that relocate a mesh's nodes): offset of
// meshes, sizeof(MeshInfo), offsets of node_first and node_count in it, offset of nodes, sizeof(WideNode), NUM_MESHES,
// MAX_NODES.
void rsim_cpu_static_layout(uint64_t* out) {
  out[0] = offsetof(StaticData, meshes);
  out[1] = sizeof(MeshInfo);
  out[2] = offsetof(MeshInfo, node_first);
  out[3] = offsetof(MeshInfo, node_count);
  out[4] = offsetof(StaticData, nodes);
  out[5] = sizeof(WideNode);
  out[6] = NUM_MESHES;
  out[7] = MAX_NODES;
}
// Word offset of cars[car].controls in ArenaState (for setting controls before a per-phase run).
uint64_t rsim_cpu_controls_word(int car) { return (offsetof(ArenaState, cars) + car * sizeof(Car) + offsetof(Car, controls)) / 4; }
// Brute force: move_cells against the remove-then-push sequence it replaces, on `iters` random sorted lists (lengths up
// to the cap, so overflow is exercised) and random block moves in the given world's grid. Returns the mismatch count.
uint64_t rsim_cpu_check_move_cells(const StaticData* w, uint64_t seed, uint64_t iters) {
  uint64_t bad = 0, x = seed | 1;
  auto rnd = [&x]() {
    x ^= x << 13;
    x ^= x >> 7;
    x ^= x << 17;
    return x;
  };
  const uint32_t ncell = w->grid_num_cells[0] * w->grid_num_cells[1] * w->grid_num_cells[2];
  static ArenaState a, b;
  for (uint64_t it = 0; it < iters; ++it) {
    memset(&a, 0, sizeof a);
    uint32_t n = (uint32_t)(rnd() % (MAX_DYN_CELLS + 1));
    uint32_t cells[MAX_DYN_CELLS];
    for (uint32_t i = 0; i < n; ++i) cells[i] = (uint32_t)(rnd() % ncell);
    for (uint32_t i = 1; i < n; ++i)  // sort (insertion)
      for (uint32_t j = i; j > 0 && cells[j - 1] > cells[j]; --j) {
        uint32_t t = cells[j];
        cells[j] = cells[j - 1];
        cells[j - 1] = t;
      }
    uint32_t h = 21 + (uint32_t)(rnd() % 3);
    for (uint32_t i = 0; i < n; ++i) a.dyn_cells[i] = DynCell{cells[i], 21 + (uint32_t)(rnd() % 3)};
    a.n_dyn_cells = n;
    uint32_t from[3], to[3];
    for (int k = 0; k < 3; ++k) {
      from[k] = (uint32_t)(rnd() % w->grid_num_cells[k]);
      to[k] = (rnd() % 4 == 0) ? from[k] : (uint32_t)(rnd() % w->grid_num_cells[k]);
    }
    if (rnd() % 2) {  // make the handle's old block present, as in a real state
      a.n_dyn_cells = 0;
      for (uint32_t i = 0; i < n / 2; ++i) dyn_cell_push(a, cells[i], 21 + (uint32_t)(rnd() % 3));
      update_cells(a, *w, true, h, from);
    }
    a.overflow[OVF_DYN_CELLS] = 0;
    b = a;
    update_cells(a, *w, false, h, from);
    update_cells(a, *w, true, h, to);
    move_cells(b, *w, h, from, to);
    if (memcmp(&a, &b, sizeof a) != 0) ++bad;
  }
  return bad;
}
uint64_t rsim_cpu_arena_words() { return ARENA_WORDS; }
uint64_t rsim_cpu_tick_words() { return TICK_WORDS; }
uint64_t rsim_cpu_static_words() { return STATIC_WORDS; }

// The correctly rounded functions over inputs, for the libm checks: fn 0 sin, 1 cos, 2 tan over the bit patterns base..base+n;
// binary: op 0 atan2f(a[i], b[i]), 1 powf(a[i], b[i]); f64: op 0 atan2_64, 1 hypot64.
void rsim_cpu_libm_unary(int fn, uint64_t base, uint64_t n, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    float x = ffrom((uint32_t)(base + i));
    out[i] = fbits(fn == 0 ? crm::sinf_cr(x) : fn == 1 ? crm::cosf_cr(x) : crm::tanf_cr(x));
  }
}
// The same over a list of bit patterns.
void rsim_cpu_libm_unary_list(int fn, const uint32_t* x, uint64_t n, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    float v = ffrom(x[i]);
    out[i] = fbits(fn == 0 ? crm::sinf_cr(v) : fn == 1 ? crm::cosf_cr(v) : crm::tanf_cr(v));
  }
}
void rsim_cpu_libm_binary(int op, const float* a, const float* b, uint64_t n, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = fbits(op == 0 ? crm::atan2f_cr(a[i], b[i]) : crm::powf_cr(a[i], b[i]));
}
void rsim_cpu_libm_f64(int op, const double* a, const double* b, uint64_t n, double* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = op == 0 ? crm::atan2_64(a[i], b[i]) : crm::hypot64(a[i], b[i]);
}
}
