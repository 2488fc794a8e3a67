// Phase P3a, the broadphase: swept AABBs of the three dynamic bodies, Bullet's 3-axis sweep and prune with u16 edges and
// incremental edge sorts that add and remove overlapping pairs, the pair array (append; swap-remove), and the dynamic
// ray-grid cell lists. Reference: cuda/sim/collision_world.cuh, grid_broadphase.cuh, axis_sweep.cuh,
// overlapping_pair_cache.cuh, collision_dispatcher_release.cuh.
#pragma once
#include "rs_ray.h"

namespace rs {

// ---- shape AABBs (collision_world.cuh:10-31, narrowphase.cuh:75-118) ---------------------------------------------------------
RS_FN Aabb sphere_box(float radius, V3 c) {
  V3 e = splat(radius + k::SPHERE_RADIUS_MARGIN);
  return Aabb{c - e, c + e};
}
RS_FN Aabb compound_box(const BodyConst& b, const Aff& t) {  // compound_shape.rs get_aabb
  V3 half = 0.5f * (b.compound_local_aabb.max - b.compound_local_aabb.min);
  V3 ctr = 0.5f * (b.compound_local_aabb.max + b.compound_local_aabb.min);
  M3 ab = mabs(t.m);
  V3 c = xform_point(t, ctr), e = ab * half;
  return Aabb{c - e, c + e};
}
RS_FN Aabb box_from_half_extents(V3 half, float margin, const Aff& t) {  // shared/aabb.rs from_half_extents_transform
  V3 hm = add_s(half, margin);
  V3 e = mabs(t.m) * hm;
  return Aabb{t.t - e, t.t + e};
}

// ---- the pair array (overlapping_pair_cache.cuh) -----------------------------------------------------------------------------
// The first pair whose proxies carry the unique ids (id0, id1) (overlapping_pair_cache.cuh find_pair). The pair words
// are loaded for all slots at once and the first match selected, instead of one dependent load chain per slot.
RS_FN int find_pair(const ArenaState& s, const StaticData& w, uint32_t id0, uint32_t id1) {
  const uint32_t n = s.n_pairs < (uint32_t)MAX_PAIRS ? s.n_pairs : (uint32_t)MAX_PAIRS;
  Pair p[MAX_PAIRS];
  RS_UNROLL
  for (int i = 0; i < MAX_PAIRS; ++i) p[i] = (uint32_t)i < n ? s.pairs[i] : Pair{0, 0};
  int found = -1;
  RS_UNROLL
  for (int i = MAX_PAIRS - 1; i >= 0; --i)
    if ((uint32_t)i < n && w.proxies[p[i].proxy0].unique_id == id0 && w.proxies[p[i].proxy1].unique_id == id1) found = i;
  return found;
}
RS_FN void add_pair(ArenaState& s, const StaticData& w, uint32_t h0, uint32_t h1) {
  const ProxyConst &p0 = w.proxies[h0], &p1 = w.proxies[h1];
  if (!((p0.filter_group & p1.filter_mask) != 0 && (p1.filter_group & p0.filter_mask) != 0)) return;
  bool lo0 = p0.unique_id <= p1.unique_id;
  uint32_t id0 = lo0 ? p0.unique_id : p1.unique_id, id1 = lo0 ? p1.unique_id : p0.unique_id;
  uint32_t a = lo0 ? h0 : h1, b = lo0 ? h1 : h0;
  if (find_pair(s, w, id0, id1) >= 0) return;
  if (s.n_pairs >= (uint32_t)MAX_PAIRS) {
    s.overflow[OVF_PAIRS] += 1;
    return;
  }
  s.pairs[s.n_pairs++] = Pair{a, b};
}
// collision_dispatcher.rs release_manifold (collision_dispatcher_release.cuh:10-39): the manifold of the unordered body
// pair leaves the pool by swap-remove; the active list drops it and renumbers the moved one.
RS_FN int find_manifold(const ArenaState& s, uint32_t b0, uint32_t b1) {  // the first match, all slots loaded at once
  uint32_t lo = b0 < b1 ? b0 : b1, hi = b0 < b1 ? b1 : b0;
  const uint32_t n = s.n_manifolds < (uint32_t)MAX_MANIFOLDS ? s.n_manifolds : (uint32_t)MAX_MANIFOLDS;
  uint32_t a[MAX_MANIFOLDS], b[MAX_MANIFOLDS];
  RS_UNROLL
  for (int i = 0; i < MAX_MANIFOLDS; ++i) {
    a[i] = (uint32_t)i < n ? s.manifolds[i].body0 : 0u;
    b[i] = (uint32_t)i < n ? s.manifolds[i].body1 : 0u;
  }
  int found = -1;
  RS_UNROLL
  for (int i = MAX_MANIFOLDS - 1; i >= 0; --i)
    if ((uint32_t)i < n && (a[i] < b[i] ? a[i] : b[i]) == lo && (a[i] < b[i] ? b[i] : a[i]) == hi) found = i;
  return found;
}
RS_FN void release_manifold(ArenaState& s, uint32_t b0, uint32_t b1) {
  int f = find_manifold(s, b0, b1);
  if (f < 0) return;
  uint32_t idx = (uint32_t)f, last = s.n_manifolds - 1;
  uint32_t n = 0;
  for (uint32_t r = 0; r < s.n_active; ++r)
    if (s.active[r] != idx) s.active[n++] = s.active[r];
  for (uint32_t r = n; r < s.n_active; ++r) s.active[r] = 0;
  s.n_active = n;
  if (idx != last) {
    s.manifolds[idx] = s.manifolds[last];
    for (uint32_t r = 0; r < s.n_active; ++r)
      if (s.active[r] == last) s.active[r] = idx;
  }
  s.manifolds[last] = Manifold{};
  s.n_manifolds = last;
}
RS_FN void remove_pair(ArenaState& s, const StaticData& w, uint32_t ha, uint32_t hb) {  // axis_sweep.cuh:39-46
  uint32_t ua = w.proxies[ha].unique_id, ub = w.proxies[hb].unique_id;
  int i = find_pair(s, w, ua < ub ? ua : ub, ua < ub ? ub : ua);
  if (i < 0) return;
  Pair p = s.pairs[i];
  uint32_t last = s.n_pairs - 1;
  if ((uint32_t)i != last) s.pairs[i] = s.pairs[last];
  s.pairs[last] = Pair{0, 0};
  s.n_pairs -= 1;
  release_manifold(s, w.proxies[p.proxy0].client_body, w.proxies[p.proxy1].client_body);
}

// ---- sweep and prune (axis_sweep.cuh) --------------------------------------------------------------------------------------
// The sweep's working copy. During P3a the sorts read and write the handles' edge indices (six u16 per handle, 72 words)
// and one axis's edge list at a time (48 words) in a copy: on the GPU the lane's LDS slice, which no other phase uses
// between P1's wheel-ray walks and P3b's narrowphase walks; on the host a local array. A sort step is then a chain of LDS
// reads and writes, where in private memory it was a chain of dependent round trips that also waited on the step's own
// stores. The edge indices are copied in at the start of P3a and written back at its end for the handles whose indices
// changed; an axis's edges are copied in when that axis's sort starts and written back over the slots the sorts wrote.
// Indices are masked to the copy's extent (a malformed state stays within its own lane's words).
constexpr uint32_t SAP_HE_WORDS = 3u * NUM_HANDLES;  // then the axis's edges, at most 64 slots (NUM_EDGES of them used)
static_assert(sizeof(HandleEdges) == 12 && sizeof(Edge) == 4 && sizeof(((ArenaState*)nullptr)->handle_edges) == 4 * SAP_HE_WORDS,
              "HandleEdges is six u16 in three words, an Edge one word");
static_assert(NUM_HANDLES <= 32 && NUM_EDGES <= 64 && SAP_HE_WORDS + 64u <= (uint32_t)LANE_LDS_WORDS,
              "masked indices stay in the lane's slice");
struct SapCopy {
#if !defined(__HIP_DEVICE_COMPILE__)
  uint32_t words[SAP_HE_WORDS + 64u];
#endif
  uint32_t he_dirty;  // bit h: handle h's edge indices were written
  uint32_t lo, hi;    // the current axis's edge slots written, lo > hi for none
  RS_FN uint32_t* word(uint32_t k) {
#if defined(__HIP_DEVICE_COMPILE__)
    return lane_word(k);
#else
    return &words[k];
#endif
  }
  // u16 j of handle h's HandleEdges: min[a] is j = a, max[a] is j = 3 + a
  RS_FN uint16_t* he(uint32_t h, uint32_t j) { return reinterpret_cast<uint16_t*>(word((h & 31u) * 3u + (j >> 1))) + (j & 1u); }
  RS_FN uint32_t he_get(uint32_t h, uint32_t j) { return *he(h, j); }
  RS_FN void he_set(uint32_t h, uint32_t j, uint32_t v) {
    *he(h, j) = (uint16_t)v;
    he_dirty |= 1u << (h & 31u);
  }
  RS_FN Edge edge(uint32_t i) { return __builtin_bit_cast(Edge, *word(SAP_HE_WORDS + (i & 63u))); }
  RS_FN void set_edge(uint32_t i, Edge e) { *word(SAP_HE_WORDS + (i & 63u)) = __builtin_bit_cast(uint32_t, e); }
  RS_FN void wrote(uint32_t a, uint32_t b) {  // edge slots [a, b]
    lo = a < lo ? a : lo;
    hi = b > hi ? b : hi;
  }
};
RS_FN void sap_copy_in(const ArenaState& s, SapCopy& c) {
  const uint32_t* src = reinterpret_cast<const uint32_t*>(&s.handle_edges[0]);
  uint32_t v[SAP_HE_WORDS];
  RS_UNROLL
  for (uint32_t k = 0; k < SAP_HE_WORDS; ++k) v[k] = src[k];
  RS_UNROLL
  for (uint32_t k = 0; k < SAP_HE_WORDS; ++k) *c.word(k) = v[k];
  c.he_dirty = 0;
}
RS_FN void sap_copy_out(ArenaState& s, SapCopy& c) {
  uint32_t* dst = reinterpret_cast<uint32_t*>(&s.handle_edges[0]);
  RS_UNROLL
  for (uint32_t h = 0; h < (uint32_t)NUM_HANDLES; ++h)
    if (c.he_dirty & (1u << h)) {
      const uint32_t a = *c.word(3u * h), b = *c.word(3u * h + 1u), d = *c.word(3u * h + 2u);
      dst[3u * h] = a;
      dst[3u * h + 1u] = b;
      dst[3u * h + 2u] = d;
    }
}
RS_FN void sap_axis_in(const ArenaState& s, SapCopy& c, int ax) {
  Edge v[NUM_EDGES];
  RS_UNROLL
  for (int i = 0; i < NUM_EDGES; ++i) v[i] = s.edges[ax][i];
  RS_UNROLL
  for (int i = 0; i < NUM_EDGES; ++i) c.set_edge((uint32_t)i, v[i]);
  c.lo = (uint32_t)NUM_EDGES;
  c.hi = 0;
}
RS_FN void sap_axis_out(ArenaState& s, SapCopy& c, int ax) {
  const uint32_t hi = c.hi < (uint32_t)NUM_EDGES ? c.hi : (uint32_t)NUM_EDGES - 1u;
  for (uint32_t i = c.lo; i <= hi; ++i) s.edges[ax][i] = c.edge(i);
}

RS_FN bool edge_max(Edge e) { return (e.pos & 1u) != 0; }
RS_FN bool overlap_2d(const ArenaState& s, uint32_t ha, uint32_t hb, int axis) {  // :28-34, touching counts
  int a1 = (1 << axis) & 3, a2 = (1 << a1) & 3;
  const HandleEdges &a = s.handle_edges[ha], &b = s.handle_edges[hb];
  return !(a.max[a1] < b.min[a1] || b.max[a1] < a.min[a1] || a.max[a2] < b.min[a2] || b.max[a2] < a.min[a2]);
}
RS_FN bool overlap_2d(SapCopy& c, uint32_t ha, uint32_t hb, int axis) {  // the same test on the working copy
  uint32_t a1 = (1u << axis) & 3u, a2 = (1u << a1) & 3u;
  return !(c.he_get(ha, 3u + a1) < c.he_get(hb, a1) || c.he_get(hb, 3u + a1) < c.he_get(ha, a1) ||
           c.he_get(ha, 3u + a2) < c.he_get(hb, a2) || c.he_get(hb, 3u + a2) < c.he_get(ha, a2));
}
// The four incremental sorts (axis_sweep.cuh:49-127): the moved edge bubbles past its neighbours, each passed neighbour's
// edge index is adjusted, and a min edge passing a max edge (or the reverse) adds or removes the pair per direction.
// The moving edge is held in registers and written once at its final slot, and its own edge index is adjusted once by
// the number of slots passed: nothing inside the loop reads either (the overlap tests read the other two axes, the pair
// and manifold updates no edges), so this is the reference's sequence of swaps.
RS_FN void sort_min_down(ArenaState& s, const StaticData& w, SapCopy& c, int ax, uint32_t i) {
  const Edge cur = c.edge(i);
  const uint32_t mv = cur.handle, i0 = i;
  while (i > 0) {
    const Edge prev = c.edge(i - 1);
    if (!(cur.pos < prev.pos)) break;
    if (edge_max(prev)) {
      if (overlap_2d(c, mv, prev.handle, ax)) add_pair(s, w, mv, prev.handle);
      c.he_set(prev.handle, 3u + ax, c.he_get(prev.handle, 3u + ax) + 1u);
    } else {
      c.he_set(prev.handle, ax, c.he_get(prev.handle, ax) + 1u);
    }
    c.set_edge(i, prev);
    --i;
  }
  if (i == i0) return;
  c.set_edge(i, cur);
  c.wrote(i, i0);
  c.he_set(mv, ax, c.he_get(mv, ax) - (i0 - i));
}
RS_FN void sort_max_up(ArenaState& s, const StaticData& w, SapCopy& c, int ax, uint32_t i) {
  const Edge cur = c.edge(i);
  const uint32_t mv = cur.handle, i0 = i;
  while (i + 1 < (uint32_t)NUM_EDGES) {
    const Edge next = c.edge(i + 1);
    if (!(next.handle != 0 && cur.pos >= next.pos)) break;
    if (!edge_max(next)) {
      if (overlap_2d(c, mv, next.handle, ax)) add_pair(s, w, mv, next.handle);
      c.he_set(next.handle, ax, c.he_get(next.handle, ax) - 1u);
    } else {
      c.he_set(next.handle, 3u + ax, c.he_get(next.handle, 3u + ax) - 1u);
    }
    c.set_edge(i, next);
    ++i;
  }
  if (i == i0) return;
  c.set_edge(i, cur);
  c.wrote(i0, i);
  c.he_set(mv, 3u + ax, c.he_get(mv, 3u + ax) + (i - i0));
}
RS_FN void sort_min_up(ArenaState& s, const StaticData& w, SapCopy& c, int ax, uint32_t i) {
  const Edge cur = c.edge(i);
  const uint32_t mv = cur.handle, i0 = i;
  while (i + 1 < (uint32_t)NUM_EDGES) {
    const Edge next = c.edge(i + 1);
    if (!(next.handle != 0 && cur.pos >= next.pos)) break;
    if (edge_max(next)) {
      if (overlap_2d(c, mv, next.handle, ax)) remove_pair(s, w, mv, next.handle);
      c.he_set(next.handle, 3u + ax, c.he_get(next.handle, 3u + ax) - 1u);
    } else {
      c.he_set(next.handle, ax, c.he_get(next.handle, ax) - 1u);
    }
    c.set_edge(i, next);
    ++i;
  }
  if (i == i0) return;
  c.set_edge(i, cur);
  c.wrote(i0, i);
  c.he_set(mv, ax, c.he_get(mv, ax) + (i - i0));
}
RS_FN void sort_max_down(ArenaState& s, const StaticData& w, SapCopy& c, int ax, uint32_t i) {
  const Edge cur = c.edge(i);
  const uint32_t mv = cur.handle, i0 = i;
  while (i > 0) {
    const Edge prev = c.edge(i - 1);
    if (!(cur.pos < prev.pos)) break;
    if (!edge_max(prev)) {
      if (overlap_2d(c, mv, prev.handle, ax)) remove_pair(s, w, mv, prev.handle);
      c.he_set(prev.handle, ax, c.he_get(prev.handle, ax) + 1u);
    } else {
      c.he_set(prev.handle, 3u + ax, c.he_get(prev.handle, 3u + ax) + 1u);
    }
    c.set_edge(i, prev);
    --i;
  }
  if (i == i0) return;
  c.set_edge(i, cur);
  c.wrote(i, i0);
  c.he_set(mv, 3u + ax, c.he_get(mv, 3u + ax) - (i0 - i));
}
RS_FN void sap_quantize(const StaticData& w, V3 p, uint16_t is_max, uint16_t out[3]) {  // axis_sweep.cuh:16-24
  V3 sc = (p - w.sap_world_min) * w.sap_quantize;
  V3 cl = vselect(cmpge(sc, splat(0.0f)), vmin(sc, splat((float)w.sap_sentinel)), splat(0.0f));
  out[0] = (uint16_t)(((uint16_t)f2i_rs(cl.x) & w.sap_handle_mask) | is_max);
  out[1] = (uint16_t)(((uint16_t)f2i_rs(cl.y) & w.sap_handle_mask) | is_max);
  out[2] = (uint16_t)(((uint16_t)f2i_rs(cl.z) & w.sap_handle_mask) | is_max);
}
RS_FN void update_axis(ArenaState& s, const StaticData& w, SapCopy& c, int ax, uint32_t h, uint16_t qmin, uint16_t qmax) {
  sap_axis_in(s, c, ax);
  uint32_t mn = c.he_get(h, (uint32_t)ax), mx = c.he_get(h, 3u + ax);
  Edge emn = c.edge(mn), emx = c.edge(mx);
  int32_t dmin = (int32_t)qmin - (int32_t)emn.pos, dmax = (int32_t)qmax - (int32_t)emx.pos;
  if (dmin == 0 && dmax == 0) return;
  emn.pos = qmin;
  c.set_edge(mn, emn);
  emx = c.edge(mx);  // mn == mx is possible in a malformed state: the second write sees the first
  emx.pos = qmax;
  c.set_edge(mx, emx);
  c.wrote(mn & 63u, mn & 63u);
  c.wrote(mx & 63u, mx & 63u);
  if (dmin < 0) sort_min_down(s, w, c, ax, mn);  // expand first, then shrink
  if (dmax > 0) sort_max_up(s, w, c, ax, mx);
  if (dmin > 0) sort_min_up(s, w, c, ax, mn);
  if (dmax < 0) sort_max_down(s, w, c, ax, mx);
  sap_axis_out(s, c, ax);
}

// ---- dynamic ray-grid cells (grid_broadphase.cuh) ----------------------------------------------------------------------------
RS_FN void cell_indices(const StaticData& w, V3 p, uint32_t out[3]) {
  V3 f = (p - w.grid_min) / w.grid_cell_size;
  int32_t v[3] = {f2i_rs(f.x), f2i_rs(f.y), f2i_rs(f.z)};
  for (int a = 0; a < 3; ++a) {
    uint32_t m = (uint32_t)(v[a] > 0 ? v[a] : 0), top = w.grid_num_cells[a] - 1;
    out[a] = m < top ? m : top;
  }
}
RS_FN uint32_t cell_index(const StaticData& w, const uint32_t i[3]) {
  return i[0] * w.grid_num_cells[1] * w.grid_num_cells[2] + i[1] * w.grid_num_cells[2] + i[2];
}
RS_FN void dyn_cell_push(ArenaState& s, uint32_t cell, uint32_t h) {  // after the cell's last entry
  if (s.n_dyn_cells >= (uint32_t)MAX_DYN_CELLS) {
    s.overflow[OVF_DYN_CELLS] += 1;
    return;
  }
  uint32_t at = 0;
  while (at < s.n_dyn_cells && s.dyn_cells[at].cell <= cell) ++at;
  for (uint32_t i = s.n_dyn_cells; i > at; --i) s.dyn_cells[i] = s.dyn_cells[i - 1];
  s.dyn_cells[at] = DynCell{cell, h};
  s.n_dyn_cells += 1;
}
RS_FN void dyn_cell_remove(ArenaState& s, uint32_t cell, uint32_t h) {  // the cell's first match
  for (uint32_t at = 0; at < s.n_dyn_cells; ++at) {
    if (s.dyn_cells[at].cell == cell && s.dyn_cells[at].handle == h) {
      for (uint32_t i = at; i + 1 < s.n_dyn_cells; ++i) s.dyn_cells[i] = s.dyn_cells[i + 1];
      s.dyn_cells[s.n_dyn_cells - 1] = DynCell{0, 0};
      s.n_dyn_cells -= 1;
      return;
    }
  }
}
RS_FN void update_cells(ArenaState& s, const StaticData& w, bool add, uint32_t h, const uint32_t idx[3]) {  // the 3x3x3 block
  uint32_t lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = (idx[a] > 1 ? idx[a] : 1) - 1;
    uint32_t top = w.grid_num_cells[a] - 1, p = idx[a] + 1;
    hi[a] = top < p ? top : p;
  }
  for (uint32_t i = lo[0]; i <= hi[0]; ++i)
    for (uint32_t j = lo[1]; j <= hi[1]; ++j)
      for (uint32_t kk = lo[2]; kk <= hi[2]; ++kk) {
        uint32_t c3[3] = {i, j, kk};
        uint32_t cell = cell_index(w, c3);
        if (add) dyn_cell_push(s, cell, h);
        else dyn_cell_remove(s, cell, h);
      }
}
// The 3x3x3 block of cells around idx, clipped to the grid (update_cells' ranges).
RS_FN void cell_block(const StaticData& w, const uint32_t idx[3], uint32_t lo[3], uint32_t n[3]) {
  for (int a = 0; a < 3; ++a) {
    lo[a] = (idx[a] > 1 ? idx[a] : 1) - 1;
    uint32_t top = w.grid_num_cells[a] - 1, p = idx[a] + 1;
    n[a] = (top < p ? top : p) - lo[a] + 1;
  }
}
// Values v of [lo, lo + n) with v < x.
RS_FN uint32_t count_below(uint32_t x, uint32_t lo, uint32_t n) { return x <= lo ? 0u : (x - lo < n ? x - lo : n); }

// A dynamic handle's move from the cell block around `from` to the block around `to`: the same list as removing (cell, h)
// for every cell of the old block (first match each) and then pushing (cell, h) for every cell of the new block in
// ascending order (each after the last entry of a cell <= it; pushes past MAX_DYN_CELLS count as overflow), computed in
// one pass. The list is read once into registers and every entry written once to its final slot: the remove/push loops
// shift the list through memory once per cell, a chain of dependent accesses a wave pays whenever any of its lanes moves.
RS_FN void move_cells(ArenaState& s, const StaticData& w, uint32_t h, const uint32_t from[3], const uint32_t to[3]) {
  const uint32_t nz = w.grid_num_cells[2], nyz = w.grid_num_cells[1] * nz;
  uint32_t olo[3], on[3], nlo[3], nn[3];
  cell_block(w, from, olo, on);
  cell_block(w, to, nlo, nn);
  const uint32_t total = nn[0] * nn[1] * nn[2], plane = nn[1] * nn[2];
  const uint32_t n = s.n_dyn_cells < (uint32_t)MAX_DYN_CELLS ? s.n_dyn_cells : (uint32_t)MAX_DYN_CELLS;
  DynCell e[MAX_DYN_CELLS];
  bool keep[MAX_DYN_CELLS];
  uint32_t removed = 0, kept = 0;
RS_UNROLL
  for (int i = 0; i < MAX_DYN_CELLS; ++i) e[i] = (uint32_t)i < n ? s.dyn_cells[i] : DynCell{0, 0};
RS_UNROLL
  for (int i = 0; i < MAX_DYN_CELLS; ++i) {
    uint32_t c = e[i].cell, ci = c / nyz, cj = (c - ci * nyz) / nz, ck = c - ci * nyz - cj * nz;
    uint32_t di = ci - olo[0], dj = cj - olo[1], dk = ck - olo[2];  // unsigned: out of the block wraps high
    bool in_old = e[i].handle == h && di < on[0] && dj < on[1] && dk < on[2];
    uint32_t bit = in_old ? 1u << ((di * 3 + dj) * 3 + dk) : 0u;
    bool drop = in_old && (removed & bit) == 0;
    if (drop) removed |= bit;
    keep[i] = (uint32_t)i < n && !drop;
    kept += keep[i] ? 1u : 0u;
  }
  const uint32_t room = (uint32_t)MAX_DYN_CELLS - kept, placed = total < room ? total : room;
  uint32_t m = 0, done = 0;  // old entries written, new entries written
RS_UNROLL
  for (int i = 0; i < MAX_DYN_CELLS; ++i) {
    if (!keep[i]) continue;
    // new cells below this entry's cell (new cells are visited in ascending index order, which is (i, j, k) order)
    uint32_t c = e[i].cell, ci = c / nyz, cj = (c - ci * nyz) / nz, ck = c - ci * nyz - cj * nz;
    uint32_t below = count_below(ci, nlo[0], nn[0]) * plane;
    if (ci - nlo[0] < nn[0]) {
      below += count_below(cj, nlo[1], nn[1]) * nn[2];
      if (cj - nlo[1] < nn[1]) below += count_below(ck, nlo[2], nn[2]);
    }
    if (below > placed) below = placed;
    for (; done < below; ++done) {
      uint32_t a = done / plane, r = done - a * plane, b = r / nn[2], k2 = r - b * nn[2];
      s.dyn_cells[done + m] = DynCell{(nlo[0] + a) * nyz + (nlo[1] + b) * nz + nlo[2] + k2, h};
    }
    s.dyn_cells[done + m] = e[i];
    ++m;
  }
  for (; done < placed; ++done) {
    uint32_t a = done / plane, r = done - a * plane, b = r / nn[2], k2 = r - b * nn[2];
    s.dyn_cells[done + m] = DynCell{(nlo[0] + a) * nyz + (nlo[1] + b) * nz + nlo[2] + k2, h};
  }
  for (uint32_t i = kept + placed; i < n; ++i) s.dyn_cells[i] = DynCell{0, 0};
  s.n_dyn_cells = kept + placed;
  s.overflow[OVF_DYN_CELLS] += total - placed;
}
RS_FN void grid_set_aabb(ArenaState& s, const StaticData& w, SapCopy& c, uint32_t d, Aabb box) {  // grid_broadphase.cuh:68-84
  uint32_t h = w.dyn_handle[d];
  s.dyn_proxy[d].aabb = box;
  uint16_t qmin[3], qmax[3];
  sap_quantize(w, box.min, 0, qmin);
  sap_quantize(w, box.max, 1, qmax);
  // the handle's six edge positions read together; an axis whose edges keep their positions is left alone, as
  // update_axis would. Sorting one axis changes only that axis's edges and edge indices, so the other axes' reads stand.
  uint32_t hmin[3], hmax[3];
  for (int ax = 0; ax < 3; ++ax) {
    hmin[ax] = c.he_get(h, (uint32_t)ax);
    hmax[ax] = c.he_get(h, 3u + ax);
  }
  uint16_t pmin[3], pmax[3];
  for (int ax = 0; ax < 3; ++ax) {
    pmin[ax] = s.edges[ax][hmin[ax]].pos;
    pmax[ax] = s.edges[ax][hmax[ax]].pos;
  }
  for (int ax = 0; ax < 3; ++ax)
    if (qmin[ax] != pmin[ax] || qmax[ax] != pmax[ax]) update_axis(s, w, c, ax, h, qmin[ax], qmax[ax]);
  uint32_t old = s.dyn_proxy[d].cell_idx, ni[3];
  cell_indices(w, box.min, ni);
  uint32_t nidx = cell_index(w, ni);
  s.dyn_proxy[d].cell_idx = nidx;
  if (nidx != old) {
    uint32_t from[3] = {s.dyn_proxy[d].indices[0], s.dyn_proxy[d].indices[1], s.dyn_proxy[d].indices[2]};
    for (int a = 0; a < 3; ++a) s.dyn_proxy[d].indices[a] = ni[a];
    move_cells(s, w, h, from, ni);  // = update_cells(remove, from) then update_cells(add, ni)
  }
}

// P3a (collision_world.cuh:34-50): the dynamic bodies' swept boxes (current and predicted transform, each grown by the
// contact breaking threshold) into the broadphase, the sweep working on its copy (SapCopy).
RS_INLINE void phase_p3a(ArenaState& s, const StaticData& w) {
  SapCopy c;
  sap_copy_in(s, c);
  const V3 cbt = splat(k::CONTACT_BREAKING_THRESHOLD);
  for (uint32_t d = 0; d < (uint32_t)NUM_DYN; ++d) {
    const BodyConst& bc = w.bodies[w.dyn_body[d]];
    const Body& b = s.bodies[d];
    if (b.collision_flags & k::CF_STATIC) continue;
    Aabb a = bc.shape_kind == k::SHAPE_SPHERE ? sphere_box(bc.sphere_margin, b.world.t) : compound_box(bc, b.world);
    a.min -= cbt;
    a.max += cbt;
    Aabb a2 = bc.shape_kind == k::SHAPE_SPHERE ? sphere_box(bc.sphere_margin, b.interp_world.t) : compound_box(bc, b.interp_world);
    a2.min -= cbt;
    a2.max += cbt;
    a.min = vmin(a.min, a2.min);
    a.max = vmax(a.max, a2.max);
    grid_set_aabb(s, w, c, d, a);
  }
  sap_copy_out(s, c);
}

}  // namespace rs
