// The traversal stack of the BVH4 walks (wheel rays: quad_ray_shapes.cuh:377-406; mesh jobs: bvh.cuh:31-56): capacity 128
// entries, and a push onto a full stack is dropped, as the reference's guard does.
//
// On the GPU the bottom TRAV_LDS entries of each lane's stack live in LDS (the kernel's dynamic shared memory, laid out
// entry-major with the work-group's lanes contiguous, so the lanes of a wave touching the same depth hit distinct banks);
// deeper entries spill to private memory. A pop is then an LDS read instead of a scratch round trip, which is what the
// walk's node-by-node loop waits on. Kernels that step arenas launch with trav_lds_bytes(block) of dynamic LDS.
#pragma once
#include "rs_defs.h"

namespace rs {

constexpr int TRAV_CAP = 128;
constexpr int TRAV_LDS = 32;
RS_HD inline size_t trav_lds_bytes(uint32_t block) { return (size_t)TRAV_LDS * block * sizeof(uint32_t); }

#if defined(__HIP_DEVICE_COMPILE__)
extern __shared__ uint32_t rs_trav_lds[];
struct TravStack {
  uint32_t spill[TRAV_CAP - TRAV_LDS];
  int n;
  __device__ void push(uint32_t v) {
    if (n >= TRAV_CAP) return;
    if (n < TRAV_LDS) rs_trav_lds[(uint32_t)n * blockDim.x + threadIdx.x] = v;
    else spill[n - TRAV_LDS] = v;
    ++n;
  }
  __device__ uint32_t pop() {
    --n;
    return n < TRAV_LDS ? rs_trav_lds[(uint32_t)n * blockDim.x + threadIdx.x] : spill[n - TRAV_LDS];
  }
};
#else
struct TravStack {
  uint32_t a[TRAV_CAP];
  int n;
  void push(uint32_t v) {
    if (n < TRAV_CAP) a[n++] = v;
  }
  uint32_t pop() { return a[--n]; }
};
#endif

}  // namespace rs
