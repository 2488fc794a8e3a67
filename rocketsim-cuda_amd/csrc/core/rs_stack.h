// The traversal stack of the BVH4 walks (wheel rays: quad_ray_shapes.cuh:377-406; mesh jobs: bvh.cuh:31-56): capacity 128
// entries, and a push onto a full stack is dropped, as the reference's guard does.
//
// On the GPU the bottom TRAV_LDS entries of each lane's stack live in LDS (the kernel's dynamic shared memory, laid out
// entry-major with the work-group's lanes contiguous, so the lanes of a wave touching the same depth hit distinct banks);
// deeper entries spill to private memory. A pop is then an LDS read instead of a scratch round trip, which is what the
// walk's node-by-node loop waits on. Kernels that step arenas launch with trav_lds_bytes(block) of dynamic LDS.
#pragma once
#include "rs_vec.h"

namespace rs {

constexpr int TRAV_CAP = 128;
constexpr int TRAV_LDS = 32;
// The solver's bank (below): per solver body (up to 8) the mutable velocity and split-impulse deltas, its inverse mass and
// whether it is a dynamic body (16 words); then the contact rows' applied impulses (up to 32), which the friction rows read.
constexpr int BANK_BODIES = 8;
constexpr int BANK_BODY_WORDS = 16;
constexpr int BANK_ROWS = 32;
constexpr int BANK_WORDS = BANK_BODIES * BANK_BODY_WORDS + BANK_ROWS;
// Both live in the same per-lane LDS slice, as does the broadphase sweep's working copy (rs_broad.h): each is used in a
// phase of its own (the walks in P1 and P3b, the sweep in P3a, the bank in P5). Each wave of the work-group owns
// 64 * LANE_LDS_WORDS words, entry-major: word k of lane l of wave v is at v * 64 * LANE_LDS_WORDS + k * 64 + l. The layout
// does not depend on the work-group size (a kernel allocates whole waves), so the device functions never read it: they
// need no dispatch-packet pointer, which every call would otherwise carry.
constexpr int LANE_LDS_WORDS = TRAV_LDS > BANK_WORDS ? TRAV_LDS : BANK_WORDS;
constexpr uint32_t WAVE_LDS_WORDS = 64u * (uint32_t)LANE_LDS_WORDS;
RS_HD inline size_t trav_lds_bytes(uint32_t block) {
  return (size_t)WAVE_LDS_WORDS * ((block + 63u) / 64u) * sizeof(uint32_t);
}

#if defined(__HIP_DEVICE_COMPILE__)
// The first word of this wave's slices. The stepping kernels declare no static LDS (a group segment of fixed size 0, which
// rsim_open checks), so their dynamic LDS starts at LDS address 0 and the address is formed from that. Through an extern
// __shared__ array a non-kernel function would find the array's address per kernel, in a table indexed by a kernel id
// that every call then passes, and each caller keeps spilled across its phase calls.
__device__ inline uint32_t* wave_lds() {
  const uint32_t byte = (threadIdx.x >> 6) * WAVE_LDS_WORDS * (uint32_t)sizeof(uint32_t);
  return (uint32_t*)(__attribute__((address_space(3))) uint32_t*)byte;
}
// Word k of this lane's slice, and word k of lane `slot`'s slice of the same wave.
__device__ inline uint32_t* lane_word(uint32_t k) { return wave_lds() + (k * 64u + (threadIdx.x & 63u)); }
__device__ inline uint32_t* wave_lane_word(uint32_t slot, uint32_t k) { return wave_lds() + (k * 64u + slot); }
struct TravStack {
  uint32_t spill[TRAV_CAP - TRAV_LDS];
  int n;
  __device__ void push(uint32_t v) {
    if (n >= TRAV_CAP) return;
    if (n < TRAV_LDS) *lane_word((uint32_t)n) = v;
    else spill[n - TRAV_LDS] = v;
    ++n;
  }
  __device__ uint32_t pop() {
    --n;
    if (n < TRAV_LDS) return *lane_word((uint32_t)n);
    // The deep branch reads an LDS word as well and picks: with a single load in each branch the compiler merges the two
    // into one load through a generic pointer, a flat instruction on the common (LDS) path too.
    const uint32_t deep = spill[n - TRAV_LDS], top = *lane_word(TRAV_LDS - 1);
    return n < TRAV_LDS ? top : deep;
  }
};
// The solver's bank during the iterations: in LDS on the GPU, so that a row reading a body the previous row just updated,
// or the impulse of the contact row a friction row is tied to, waits on LDS rather than on a private-memory round trip.
// Fields 0-4 of a body are the deltas (dlin, dang, push, turn) and the inverse mass; word 15 is 1 for a dynamic body.
struct DeltaBank {
  __device__ static float* at(uint32_t word) { return reinterpret_cast<float*>(lane_word(word)); }
  __device__ V3 get(int body, int field) const {
    const float* p = at((uint32_t)(body * BANK_BODY_WORDS + field * 3));
    return V3{p[0], p[64], p[128]};
  }
  __device__ void set(int body, int field, V3 v) {
    float* p = at((uint32_t)(body * BANK_BODY_WORDS + field * 3));
    p[0] = v.x;
    p[64] = v.y;
    p[128] = v.z;
  }
  __device__ bool dynamic(int body) const { return *at((uint32_t)(body * BANK_BODY_WORDS + 15)) != 0.0f; }
  __device__ void set_dynamic(int body, bool d) { *at((uint32_t)(body * BANK_BODY_WORDS + 15)) = d ? 1.0f : 0.0f; }
  __device__ float applied(int row) const { return *at((uint32_t)(BANK_BODIES * BANK_BODY_WORDS + row)); }
  __device__ void set_applied(int row, float v) { *at((uint32_t)(BANK_BODIES * BANK_BODY_WORDS + row)) = v; }
};
// Wave-wide staging through the same slices (the narrowphase's batched separating-axis tests, between the BVH walks): slot
// r of a wave is the slice of the wave's lane r, so any lane of the wave can write or read any slot. LDS accesses of one
// wave execute in order; the fences only keep the compiler from moving them across the hand-off points.
__device__ inline float* wave_slot(uint32_t slot, uint32_t k) { return reinterpret_cast<float*>(wave_lane_word(slot, k)); }
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// A wave-pooled stage: this lane's `cnt` tasks (at most MAXT) and every other active lane's are run by all active lanes of
// the wave in turns of one task per lane. stage(t, rec) writes task t's record (REC words), exec(rec, out) computes its
// result (OUT words) from the record alone, collect(t, out) hands it back to the owner. Lane ranks and task offsets come
// from ballots (task order: lane order, then the owner's order); the records and results pass through the LDS slots.
template <int REC, int OUT, int MAXT, class Stage, class Exec, class Collect>
__device__ inline void wave_pool(uint32_t cnt, Stage stage, Exec exec, Collect collect) {
  static_assert(OUT <= REC && REC <= LANE_LDS_WORDS, "a slot holds a record, then its result");
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
  const uint64_t active = __ballot(1);
  const uint32_t nact = (uint32_t)__popcll(active), rank = (uint32_t)__popcll(active & below);
  uint32_t pre = 0, total = 0;
  for (uint32_t bit = 0; (1u << bit) <= (uint32_t)MAXT; ++bit) {
    const uint64_t m = __ballot((cnt >> bit) & 1u);
    pre += (uint32_t)__popcll(m & below) << bit;
    total += (uint32_t)__popcll(m) << bit;
  }
  float r[REC];
  for (uint32_t base = 0; base < total; base += nact) {  // total and nact are the same in every lane
    const uint32_t lo = (pre > base ? pre : base) - pre;
    const uint32_t end = pre + cnt < base + nact ? pre + cnt : base + nact;
    const uint32_t hi = end > pre ? end - pre : 0u;
    for (uint32_t t = lo; t < hi; ++t) {  // this lane's tasks in the window, into the slots their order gives
      stage(t, r);
      const uint32_t slot = pre + t - base;
      for (int k = 0; k < REC; ++k) *wave_slot(slot, (uint32_t)k) = r[k];
    }
    wave_lds_sync();
    if (base + rank < total) {
      for (int k = 0; k < REC; ++k) r[k] = *wave_slot(rank, (uint32_t)k);
      float o[OUT];
      exec(r, o);
      for (int k = 0; k < OUT; ++k) *wave_slot(rank, (uint32_t)k) = o[k];
    }
    wave_lds_sync();
    for (uint32_t t = lo; t < hi; ++t) {
      float o[OUT];
      const uint32_t slot = pre + t - base;
      for (int k = 0; k < OUT; ++k) o[k] = *wave_slot(slot, (uint32_t)k);
      collect(t, o);
    }
    wave_lds_sync();
  }
}
#else
struct DeltaBank {
  V3 v[BANK_BODIES][5];
  bool dyn[BANK_BODIES];
  float row_applied[BANK_ROWS];
  V3 get(int body, int field) const { return v[body][field]; }
  void set(int body, int field, V3 x) { v[body][field] = x; }
  bool dynamic(int body) const { return dyn[body]; }
  void set_dynamic(int body, bool d) { dyn[body] = d; }
  float applied(int row) const { return row_applied[row]; }
  void set_applied(int row, float x) { row_applied[row] = x; }
};
struct TravStack {
  uint32_t a[TRAV_CAP];
  int n;
  void push(uint32_t v) {
    if (n < TRAV_CAP) a[n++] = v;
  }
  uint32_t pop() { return a[--n]; }
};
// One lane: the tasks in order.
template <int REC, int OUT, int MAXT, class Stage, class Exec, class Collect>
inline void wave_pool(uint32_t cnt, Stage stage, Exec exec, Collect collect) {
  float r[REC], o[OUT];
  for (uint32_t t = 0; t < cnt; ++t) {
    stage(t, r);
    exec(r, o);
    collect(t, o);
  }
}
#endif

}  // namespace rs
