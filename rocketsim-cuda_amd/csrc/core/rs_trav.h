// Depth-first walks of an arena mesh's BVH4 for a query box (bvh.cuh:9-71): the per-lane cursor, and batches of walks that
// the GPU pools across the lanes of a wave. The narrowphase's mesh jobs (P3b) and the wheel rays' mesh candidates (P1)
// both take their leaves, in depth-first order, from a batch walked before they are consumed.
#pragma once
#include "rs_stack.h"

namespace rs {

constexpr uint32_t BVH_LEAF = 1u << 31;

// BVH4 node test (bvh.cuh:16-25): lane l set when child l's box overlaps `box` (touching counts)
RS_FN uint32_t node_mask(const WideNode& n, const Aabb& b) {
  uint32_t m = 0;
  for (int l = 0; l < 4; ++l) {
    bool o = n.min_x[l] <= b.max.x && n.max_x[l] >= b.min.x && n.min_y[l] <= b.max.y && n.max_y[l] >= b.min.y &&
             n.min_z[l] <= b.max.z && n.max_z[l] >= b.min.z;
    m |= (o ? 1u : 0u) << l;
  }
  return m & n.valid_mask;
}
// ---- the BVH4 cursor (bvh.cuh:31-56) -----------------------------------------------------------------------------------------
struct BvhCursor {
  TravStack st;
};
RS_FN void cursor_init(BvhCursor& cur, const MeshInfo& mesh, const Aabb& box) {
  cur.st.n = 0;
  if (!overlaps(box, mesh.root_aabb)) return;
  cur.st.push(0);
}
// The next overlapping leaf in depth-first order (its index in the mesh's leaves); false when the walk is done.
RS_FN bool cursor_next_leaf(BvhCursor& cur, const StaticData& w, const MeshInfo& mesh, const Aabb& box, uint32_t* leaf) {
  while (cur.st.n != 0) {
    uint32_t work = cur.st.pop();
    if (work & BVH_LEAF) {
      *leaf = work & ~BVH_LEAF;
      return true;
    }
    const WideNode& node = w.nodes[mesh.node_first + work];
    uint32_t mask = node_mask(node, box);
    for (int lane = (int)node.child_count - 1; lane >= 0; --lane)
      if (mask & (1u << lane)) cur.st.push(node.children[lane]);
  }
  return false;
}
// The same walk yielding each leaf's triangle.
RS_FN bool cursor_next(BvhCursor& cur, const StaticData& w, const MeshInfo& mesh, const Aabb& box, uint32_t* tri) {
  uint32_t leaf;
  if (!cursor_next_leaf(cur, w, mesh, box, &leaf)) return false;
  *tri = w.leaves[mesh.leaf_first + leaf].tri;
  return true;
}

// Each mesh job (a car box or the ball against an arena mesh in P3b, a wheel-ray packet's box against a candidate mesh in
// P1) walks the mesh's BVH4 depth first and visits the overlapping leaves in that order. The caller collects its jobs
// first and walks them all as one batch; the jobs then take their leaves from the list instead of walking. On the GPU the walks of all active lanes of a wave are pooled: a
// level-by-level walk over a frontier shared through LDS, every lane testing one node of any job at a time, each item
// carrying its path (the child slot at each level) as a key; the leaves of a job, ranked by key, are its depth-first
// order. Anything that does not fit (frontier, leaf list, depth, a partial wave) falls back to each lane's own walks.
constexpr int TJ_MAX = 4, TL_MAX = 48;
enum { TJ_CAR = 0, TJ_BALL = 1, TJ_RAYS = 2 };
struct TravJob {
  uint32_t kind, body, mesh_body;  // mesh_body: the mesh record (for the narrowphase's jobs also the mesh's body)
  Aabb box;              // the query box in mesh space
  uint32_t first, count; // its items in leaf_tri, in depth-first order
  bool in_place;         // did not fit: the job walks the BVH itself
};
struct TravBatch {
  uint32_t n_jobs, n_leaves, next;
  uint32_t leaves;  // 0: the list holds each leaf's triangle; 1: the leaf itself (its index in the mesh's leaves)
  TravJob job[TJ_MAX];
  uint32_t leaf_tri[TL_MAX];
};
RS_FN void trav_walk(const StaticData& w, TravBatch& tb) {  // each job's own depth-first walk
  tb.n_leaves = 0;
  for (uint32_t j = 0; j < tb.n_jobs; ++j) {
    TravJob& jb = tb.job[j];
    const MeshInfo& mesh = w.meshes[jb.mesh_body];
    jb.first = tb.n_leaves;
    jb.in_place = false;
    BvhCursor cur;
    cursor_init(cur, mesh, jb.box);
    uint32_t leaf;
    while (cursor_next_leaf(cur, w, mesh, jb.box, &leaf)) {
      if (tb.n_leaves >= (uint32_t)TL_MAX) {
        jb.in_place = true;
        tb.n_leaves = jb.first;
        break;
      }
      tb.leaf_tri[tb.n_leaves++] = tb.leaves ? leaf : w.leaves[mesh.leaf_first + leaf].tri;
    }
    jb.count = tb.n_leaves - jb.first;
  }
}
#if defined(__HIP_DEVICE_COMPILE__)
// The pooled walk's LDS area (words of the wave's 64 lane slices): job records, per-job leaf counts and first slots, the
// control words, two frontiers and the leaf list.
constexpr uint32_t PT_JOBS = 64 * TJ_MAX, PT_Q = 1024, PT_L = 1024, PT_DEPTH = 15;
constexpr uint32_t PT_REC = 10;  // a job record: box (6 words), the address of its mesh's first node (2), the mesh's node count
constexpr uint32_t PT_JOB_W = 0, PT_CNT_W = PT_JOB_W + PT_REC * PT_JOBS, PT_START_W = PT_CNT_W + PT_JOBS, PT_CTRL_W = PT_START_W + PT_JOBS,
                   PT_Q_W = PT_CTRL_W + 4, PT_L_W = PT_Q_W + 4 * PT_Q, PT_WORDS = PT_L_W + 2 * PT_L;
static_assert(PT_WORDS <= 64u * (uint32_t)LANE_LDS_WORDS, "the pooled walk fits the wave's LDS slices");
__device__ inline uint32_t* wave_word(uint32_t j) { return wave_lds() + j; }
__device__ inline uint32_t* pt_q(uint32_t which, uint32_t i, uint32_t k) { return wave_word(PT_Q_W + which * 2 * PT_Q + 2 * i + k); }

__device__ inline bool trav_pooled(const StaticData& w, TravBatch& tb) {
  const uint32_t lane = threadIdx.x & 63u;
  const uint64_t below = lane ? (~0ull >> (64u - lane)) : 0ull;
  const uint64_t active = __ballot(1);
  const uint32_t nact = (uint32_t)__popcll(active), rank = (uint32_t)__popcll(active & below);
  if (rank == 0)
    for (uint32_t k = 0; k < 4; ++k) *wave_word(PT_CTRL_W + k) = 0u;
  wave_lds_sync();
  for (uint32_t j = 0; j < tb.n_jobs; ++j) {  // job records and the roots whose box overlaps
    const TravJob& jb = tb.job[j];
    const uint32_t ref = lane * TJ_MAX + j;
    const MeshInfo& mesh = w.meshes[jb.mesh_body];
    const float bx[6] = {jb.box.min.x, jb.box.min.y, jb.box.min.z, jb.box.max.x, jb.box.max.y, jb.box.max.z};
    for (uint32_t k = 0; k < 6; ++k) *wave_word(PT_JOB_W + PT_REC * ref + k) = __builtin_bit_cast(uint32_t, bx[k]);
    // the owner's own world: with per-arena worlds another lane's StaticData may differ
    const uint64_t nodes = (uint64_t)(uintptr_t)(w.nodes + mesh.node_first);
    *wave_word(PT_JOB_W + PT_REC * ref + 6) = (uint32_t)nodes;
    *wave_word(PT_JOB_W + PT_REC * ref + 7) = (uint32_t)(nodes >> 32);
    *wave_word(PT_JOB_W + PT_REC * ref + 8) = mesh.node_count;
    *wave_word(PT_CNT_W + ref) = 0u;
    if (overlaps(jb.box, mesh.root_aabb)) {
      const uint32_t slot = atomicAdd(wave_word(PT_CTRL_W + 0), 1u);  // at most 256 roots
      *pt_q(0, slot, 0) = ref << 20;
      *pt_q(0, slot, 1) = 0u;
    }
  }
  wave_lds_sync();
  uint32_t cur = 0;
  for (uint32_t depth = 0;; ++depth) {  // one BVH level per pass; every lane sees the same counts
    const uint32_t n_all = *wave_word(PT_CTRL_W + cur);
    if (n_all == 0) break;
    // a level that overflowed the queue is flagged already (its walk is then redone per lane); only its first PT_Q
    // items were stored, so only those are read
    const uint32_t n = n_all < PT_Q ? n_all : PT_Q;
    if (depth >= PT_DEPTH) {
      if (rank == 0) *wave_word(PT_CTRL_W + 3) = 1u;
      break;
    }
    // each lane's items i = rank, rank + nact, ...; the next item's record and node are loaded before this item's
    // children are queued, so the node's load overlaps the queueing (the queues' order is free: leaves are ranked by key)
    uint32_t i = rank, w0 = 0u, key = 0u, ok = 0u;
    Aabb box = {};
    WideNode nd = {};
    auto fetch = [&](uint32_t it) {
      w0 = *pt_q(cur, it, 0);
      key = *pt_q(cur, it, 1);
      const uint32_t ref = w0 >> 20, node = w0 & 0xFFFFFu;
      ok = (ref < PT_JOBS && node < *wave_word(PT_JOB_W + PT_REC * ref + 8)) ? 1u : 0u;
      if (!ok) {  // never for a well-formed BVH
        *wave_word(PT_CTRL_W + 3) = 1u;
        return;
      }
      const uint32_t rec = PT_JOB_W + PT_REC * ref;
      box.min = v3(__builtin_bit_cast(float, *wave_word(rec)), __builtin_bit_cast(float, *wave_word(rec + 1)),
                   __builtin_bit_cast(float, *wave_word(rec + 2)));
      box.max = v3(__builtin_bit_cast(float, *wave_word(rec + 3)), __builtin_bit_cast(float, *wave_word(rec + 4)),
                   __builtin_bit_cast(float, *wave_word(rec + 5)));
      // the job's node array lives in the world (global memory): said so, its loads are global_ rather than flat ones
      const RS_GLOBAL_AS WideNode* nodes =
          (const RS_GLOBAL_AS WideNode*)(uintptr_t)(((uint64_t)*wave_word(rec + 7) << 32) | (uint64_t)*wave_word(rec + 6));
      nd = nodes[node];
    };
    if (i < n) fetch(i);
    while (i < n) {
      const uint32_t c_ok = ok, c_key = key, c_ref = w0 >> 20;
      uint32_t mask = 0u, children[4] = {0u, 0u, 0u, 0u}, child_count = 0u;
      if (c_ok) {
        mask = node_mask(nd, box);
        for (int l = 0; l < 4; ++l) children[l] = nd.children[l];
        child_count = nd.child_count;
      }
      const uint32_t next = i + nact;
      if (next < n) fetch(next);
      if (c_ok) {
        for (uint32_t l = 0; l < child_count && l < 4u; ++l) {
          if (!((mask >> l) & 1u)) continue;
          const uint32_t ch = children[l], k2 = c_key | (l << (28u - 2u * depth));
          const uint32_t idx = ch & ~BVH_LEAF;
          if (idx >= (1u << 20)) {
            *wave_word(PT_CTRL_W + 3) = 1u;
            continue;
          }
          if (ch & BVH_LEAF) {
            const uint32_t slot = atomicAdd(wave_word(PT_CTRL_W + 2), 1u);
            if (slot < PT_L) {
              *wave_word(PT_L_W + 2 * slot) = (c_ref << 20) | idx;
              *wave_word(PT_L_W + 2 * slot + 1) = k2;
              atomicAdd(wave_word(PT_CNT_W + c_ref), 1u);
            } else {
              *wave_word(PT_CTRL_W + 3) = 1u;
            }
          } else {
            const uint32_t slot = atomicAdd(wave_word(PT_CTRL_W + (cur ^ 1u)), 1u);
            if (slot < PT_Q) {
              *pt_q(cur ^ 1u, slot, 0) = (c_ref << 20) | idx;
              *pt_q(cur ^ 1u, slot, 1) = k2;
            } else {
              *wave_word(PT_CTRL_W + 3) = 1u;
            }
          }
        }
      }
      i = next;
    }
    wave_lds_sync();
    if (rank == 0) *wave_word(PT_CTRL_W + cur) = 0u;  // this level's frontier is consumed
    wave_lds_sync();
    cur ^= 1u;
  }
  wave_lds_sync();
  if (*wave_word(PT_CTRL_W + 3) != 0u) return false;
  const uint32_t nl = *wave_word(PT_CTRL_W + 2);
  // each job's slots: the lanes' leaf totals prefixed in lane order (ballots over the bits of the totals), then job order
  uint32_t total = 0;
  for (uint32_t j = 0; j < tb.n_jobs; ++j) total += *wave_word(PT_CNT_W + lane * TJ_MAX + j);
  uint32_t pre = 0;
  for (uint32_t bit = 0; (1u << bit) <= PT_L; ++bit) pre += (uint32_t)__popcll(__ballot((total >> bit) & 1u) & below) << bit;
  for (uint32_t j = 0; j < tb.n_jobs; ++j) {
    const uint32_t ref = lane * TJ_MAX + j;
    *wave_word(PT_START_W + ref) = pre;
    *wave_word(PT_JOB_W + PT_REC * ref) = 0u;  // the job record's first word becomes its fill counter
    pre += *wave_word(PT_CNT_W + ref);
  }
  wave_lds_sync();
  for (uint32_t i = rank; i < nl; i += nact) {  // the leaves into their job's segment (frontier 0 reused), any order
    const uint32_t w0 = *wave_word(PT_L_W + 2 * i), key = *wave_word(PT_L_W + 2 * i + 1), ref = w0 >> 20;
    const uint32_t pos = *wave_word(PT_START_W + ref) + atomicAdd(wave_word(PT_JOB_W + PT_REC * ref), 1u);
    *pt_q(0, pos, 0) = w0;
    *pt_q(0, pos, 1) = key;
  }
  wave_lds_sync();
  for (uint32_t i = rank; i < nl; i += nact) {  // each leaf's depth-first rank within its job (keys are distinct paths)
    const uint32_t w0 = *pt_q(0, i, 0), key = *pt_q(0, i, 1), ref = w0 >> 20;
    const uint32_t s0 = *wave_word(PT_START_W + ref), cnt = *wave_word(PT_CNT_W + ref);
    uint32_t r = 0;
    // unrolled so that several of the segment's key loads are in flight at once
#pragma unroll 4
    for (uint32_t j = s0; j < s0 + cnt; ++j) r += *pt_q(0, j, 1) < key ? 1u : 0u;
    *pt_q(1, s0 + r, 0) = w0 & 0xFFFFFu;
  }
  wave_lds_sync();
  tb.n_leaves = 0;
  for (uint32_t j = 0; j < tb.n_jobs; ++j) {
    TravJob& jb = tb.job[j];
    const uint32_t ref = lane * TJ_MAX + j, s0 = *wave_word(PT_START_W + ref), cnt = *wave_word(PT_CNT_W + ref);
    jb.first = tb.n_leaves;
    jb.in_place = tb.n_leaves + cnt > (uint32_t)TL_MAX;
    if (!jb.in_place) {
      const MeshInfo& mesh = w.meshes[jb.mesh_body];
      for (uint32_t k = 0; k < cnt; ++k) {
        const uint32_t leaf = *pt_q(1, s0 + k, 0), in_range = leaf < mesh.leaf_count ? leaf : 0u;
        tb.leaf_tri[tb.n_leaves + k] = tb.leaves ? in_range : w.leaves[mesh.leaf_first + in_range].tri;
      }
      tb.n_leaves += cnt;
    }
    jb.count = jb.in_place ? 0u : cnt;
  }
  wave_lds_sync();  // the slices are free again
  return true;
}
#endif
RS_FN void trav_run(const StaticData& w, TravBatch& tb) {
#if defined(__HIP_DEVICE_COMPILE__)
  if (trav_pooled(w, tb)) return;
#endif
  trav_walk(w, tb);
}

}  // namespace rs
