// P3b building blocks: body access by index, persistent manifolds (add, merge, refresh, the 4-point cache replacement),
// the arena's contact tracker (records and folded events) and Bullet's internal-edge normal fix. Reference:
// cuda/sim/narrowphase.cuh, persistent_manifold.cuh, arena_contact_tracker.cuh, internal_edge_utility.cuh, triangle_shape.cuh.
#pragma once
#include "rs_broad.h"

namespace rs {

struct SatBatch;
struct TravBatch;
// The narrowphase context: the arena, its world, a count of pair algorithms the tick reached that are not implemented, and
// the tick's precomputed mesh-job BVH walks and car-mesh separating-axis tests (null: every job does its own).
struct Np {
  ArenaState& s;
  const StaticData& w;
  uint32_t unimplemented;
  SatBatch* sb;
  TravBatch* tb;
  uint32_t fold;       // event bits the records fold in during the phase, or-ed into the state's fold_events at its end
  uint32_t n_records;  // the state's record count during the phase, written back at its end
};
RS_FN Aff body_trans(const Np& c, uint32_t b) {
  // Both candidates are loaded, each from its own memory (a dynamic body's transform from the arena state, a static
  // body's pose from the world), and one is picked. A load in each branch would be merged by the compiler into one load
  // through a conditional address, mixing the two memories into a generic pointer whose loads are flat instructions.
  // d < 0 reads slot 0, which exists, and discards it.
  const int d = dyn_slot(c.w, b);
  const Aff dyn = c.s.bodies[d >= 0 ? d : 0].world;
  const Aff fixed = c.w.bodies[b].pose;
  return d >= 0 ? dyn : fixed;
}
RS_FN uint32_t body_flags(const Np& c, uint32_t b) {
  int d = dyn_slot(c.w, b);
  return d >= 0 ? c.s.bodies[d].collision_flags : k::CF_STATIC;
}
RS_FN bool body_static(const Np& c, uint32_t b) { return (body_flags(c, b) & k::CF_STATIC) != 0; }
RS_FN bool body_responds(const Np& c, uint32_t b) { return (body_flags(c, b) & k::CF_NO_CONTACT_RESPONSE) == 0; }
RS_FN bool body_awake(const Np& c, uint32_t b) {  // a static body is added asleep
  int d = dyn_slot(c.w, b);
  return d >= 0 && c.s.bodies[d].activation != k::ACT_SLEEPING && c.s.bodies[d].activation != k::ACT_DISABLED;
}
RS_FN Aabb body_box(const Np& c, uint32_t b) {  // collision_shape.rs get_aabb (narrowphase.cuh:86-107)
  const BodyConst& bc = c.w.bodies[b];
  Aff t = body_trans(c, b);
  if (bc.shape_kind == k::SHAPE_SPHERE) return sphere_box(bc.sphere_margin, t.t);
  if (bc.shape_kind == k::SHAPE_PLANE) return c.w.planes[b - NUM_MESHES].aabb;
  if (bc.shape_kind == k::SHAPE_MESH) {
    const Aabb& id = c.w.meshes[b].aabb_ident;
    return Aabb{id.min + t.t, id.max + t.t};
  }
  return compound_box(bc, t);
}
RS_FN Aff child_world(const Np& c, uint32_t b) { return body_trans(c, b) * c.w.bodies[b].child_trans; }
RS_FN Aabb child_world_box(const Np& c, uint32_t b) {
  const BodyConst& bc = c.w.bodies[b];
  return box_from_half_extents(bc.box_half_extents, bc.box_margin, child_world(c, b));
}

// ---- persistent manifolds (persistent_manifold.cuh) ---------------------------------------------------------------------------
// A fresh manifold built in place: the header only. Its point slots past n_points are never read (points are added at
// n_points, the four-point replacement runs only when all four are written, a merge or an insert reads n_points of them),
// and an insert writes the unused slots of the pool's manifold as zeros, so no 79-word zero-fill and copy per pair.
RS_FN void manifold_init(Manifold& m, uint32_t b0, uint32_t b1) {
  m.body0 = b0;
  m.body1 = b1;
  m.n_points = 0;
}
RS_FN float breaking_threshold(const Np& c, const Manifold& m) {
  return fmin_rs(c.w.bodies[m.body0].contact_breaking_threshold, c.w.bodies[m.body1].contact_breaking_threshold);
}
RS_FN void point_swap_remove(Manifold& m, uint32_t i) {
  uint32_t last = m.n_points - 1;
  if (i != last) m.pts[i] = m.pts[last];
  m.pts[last] = ManifoldPoint{};
  m.n_points = last;
}
RS_FN float res_area(V3 p, V3 a, V3 b, V3 c) { return length_sq(cross(p - a, b - c)); }  // get_res
// A manifold's slot keys held in registers while points are added to it: each point's local_a and distance, the point
// count and the breaking threshold, which are all that choose an added point's slot. A run of additions then reads them
// once instead of reading back, for every addition, the points the previous ones stored.
struct PointKeys {
  V3 la[MAX_POINTS];
  float dist[MAX_POINTS];
  uint32_t n;
  float cbt;
};
RS_FN PointKeys point_keys(const Np& c, const Manifold& m) {
  PointKeys pk;
  pk.n = m.n_points;
  pk.cbt = breaking_threshold(c, m);
  RS_UNROLL
  for (uint32_t i = 0; i < (uint32_t)MAX_POINTS; ++i) {
    pk.la[i] = i < pk.n ? m.pts[i].local_a : V3{0.0f, 0.0f, 0.0f};
    pk.dist[i] = i < pk.n ? m.pts[i].distance : 0.0f;
  }
  return pk;
}
// persistent_manifold.rs add_point_no_callback (persistent_manifold.cuh:139-163) on the keys: the nearest cached point within
// the breaking threshold (get_cache_entry, :120-133), else the next free slot, else the replacement slot of
// sort_cached_points (:81-117); the point is stored there and the keys updated.
RS_FN uint32_t add_point_keyed(Manifold& m, PointKeys& pk, const ManifoldPoint& np) {
  int at = -1;  // cache_entry
  float shortest = pk.cbt * pk.cbt;
  RS_UNROLL
  for (uint32_t i = 0; i < (uint32_t)MAX_POINTS; ++i) {
    if (i >= pk.n) continue;
    float d2 = length_sq(pk.la[i] - np.local_a);
    if (d2 < shortest) {
      shortest = d2;
      at = (int)i;
    }
  }
  uint32_t slot;
  if (at >= 0) {
    slot = (uint32_t)at;
  } else if (pk.n == (uint32_t)MAX_POINTS) {  // replace_slot
    static_assert(MAX_POINTS == 4, "the replacement areas are written for four points");
    uint32_t deepest = MAX_POINTS;
    float maxpen = np.distance;
    RS_UNROLL
    for (uint32_t i = 0; i < (uint32_t)MAX_POINTS; ++i)
      if (pk.dist[i] < maxpen) {
        deepest = i;
        maxpen = pk.dist[i];
      }
    V3 p = np.local_a;
    float r0 = deepest == 0 ? 0.0f : res_area(p, pk.la[1], pk.la[3], pk.la[2]);
    float r1 = deepest == 1 ? 0.0f : res_area(p, pk.la[0], pk.la[3], pk.la[2]);
    float r2 = deepest == 2 ? 0.0f : res_area(p, pk.la[0], pk.la[3], pk.la[1]);
    float r3 = deepest == 3 ? 0.0f : res_area(p, pk.la[0], pk.la[2], pk.la[1]);
    uint32_t best = 0;
    float mx = r0;
    if (r1 > mx) { best = 1; mx = r1; }
    if (r2 > mx) { best = 2; mx = r2; }
    if (r3 > mx) best = 3;
    slot = best;
  } else {
    slot = pk.n++;
    m.n_points = pk.n;
  }
  m.pts[slot] = np;
  RS_UNROLL
  for (uint32_t i = 0; i < (uint32_t)MAX_POINTS; ++i)
    if (i == slot) {
      pk.la[i] = np.local_a;
      pk.dist[i] = np.distance;
    }
  return slot;
}
RS_FN float combined_friction(const Np& c, uint32_t b0, uint32_t b1) {
  float f0 = c.w.bodies[b0].friction, f1 = c.w.bodies[b1].friction;
  return (body_static(c, b0) || body_static(c, b1)) ? fmin_rs(f0, f1) : f0 * f1;
}
RS_FN float combined_restitution(const Np& c, uint32_t b0, uint32_t b1) {
  float r0 = c.w.bodies[b0].restitution, r1 = c.w.bodies[b1].restitution;
  return (body_static(c, b0) || body_static(c, b1)) ? fmax_rs(r0, r1) : r0 * r1;
}
RS_FN void merge_points(const Np& c, Manifold& m, const Manifold& o) {  // :196-205
  if (m.n_points == 0) {
    for (uint32_t i = 0; i < o.n_points; ++i) m.pts[m.n_points++] = o.pts[i];
  } else {
    PointKeys pk = point_keys(c, m);
    for (uint32_t i = 0; i < o.n_points; ++i) add_point_keyed(m, pk, o.pts[i]);
  }
}
RS_FN void refresh_points(const Np& c, Manifold& m, uint32_t b0, uint32_t b1) {  // :208-213 with persistent_manifold.rs:312
  if (m.n_points == 0) return;
  Aff ta = body_trans(c, b0), tb = body_trans(c, b1);
  const uint32_t n = m.n_points;
  // the refreshed values stay in registers for the removal pass: a removal moves the last point into a slot the pass has
  // already left, so point i is still the one refreshed here when the pass reaches it
  V3 wa[MAX_POINTS] = {}, wb[MAX_POINTS] = {}, nb[MAX_POINTS] = {};
  float dist[MAX_POINTS] = {};
  RS_UNROLL
  for (uint32_t i = 0; i < (uint32_t)MAX_POINTS; ++i) {
    if (i >= n) continue;
    ManifoldPoint& p = m.pts[i];
    wa[i] = xform_point(ta, p.local_a);
    wb[i] = xform_point(tb, p.local_b);
    nb[i] = p.normal_on_b;
    dist[i] = dot(wa[i] - wb[i], nb[i]);
    p.world_on_a = wa[i];
    p.world_on_b = wb[i];
    p.distance = dist[i];
  }
  float cbt = breaking_threshold(c, m), cbt2 = cbt * cbt;
  RS_UNROLL
  for (int i = MAX_POINTS - 1; i >= 0; --i) {  // from the last point down, over the points present at the start
    if (i >= (int)n) continue;
    if (dist[i] > cbt) {
      point_swap_remove(m, (uint32_t)i);
      continue;
    }
    V3 proj = wa[i] - nb[i] * dist[i];
    V3 diff = wb[i] - proj;
    if (dot(diff, diff) > cbt2) point_swap_remove(m, (uint32_t)i);
  }
}

// ---- the contact tracker (arena_contact_tracker.cuh) ----------------------------------------------------------------------------
RS_FN bool tracker_swap(const Np& c, uint32_t a, uint32_t b) {
  uint32_t ua = c.w.bodies[a].user_idx, ub = c.w.bodies[b].user_idx;
  return (ua != k::USER_NONE && ub != k::USER_NONE) ? ua > ub : ub != k::USER_NONE;
}
// A record: car-ball and car-car are kept in call order; car-world and ball-world fold into event bits (and, for a car, the
// normal of its last world record); others are dropped (:35-66).
RS_FN void push_record(Np& c, bool swapped, uint32_t a, uint32_t b, const ManifoldPoint& p) {
  ArenaState& s = c.s;
  uint32_t ua = c.w.bodies[a].user_idx, ub = c.w.bodies[b].user_idx;
  if (ua == k::USER_CAR) {
    if (ub != k::USER_BALL && ub != k::USER_CAR) {
      uint32_t car = c.w.bodies[a].user_pointer;
      c.fold |= 1u << (EV_CAR_HIT_WORLD + car);
      s.fold_car_normal[car] = p.normal_on_b;
      return;
    }
  } else if (ua == k::USER_BALL && ub == k::USER_NONE) {
    c.fold |= 1u << EV_BALL_HIT_WORLD;
    if (fabs_(p.normal_on_b.z) < 0.8f) c.fold |= 1u << EV_BALL_WALL;
    return;
  } else {
    return;
  }
  if (c.n_records >= (uint32_t)MAX_RECORDS) {
    s.overflow[OVF_RECORDS] += 1;
    return;
  }
  ContactRecord& r = s.records[c.n_records++];
  r.is_swap = swapped ? 1u : 0u;
  r.body_a = a;
  r.body_b = b;
  r.pt.world_on_b = p.world_on_b;
  r.pt.world_on_a = p.world_on_a;
  r.pt.normal_on_b = p.normal_on_b;
}
RS_FN void contact_processed(Np& c, const ManifoldPoint& p, uint32_t a, uint32_t b) {  // :96-108
  bool sw = tracker_swap(c, a, b);
  if (sw) {
    uint32_t t = a;
    a = b;
    b = t;
  }
  uint32_t ua = c.w.bodies[a].user_idx, ub = c.w.bodies[b].user_idx;
  if (ua != k::USER_CAR || ub == k::USER_BALL || ub == k::USER_CAR) return;
  push_record(c, sw, a, b, p);
}

// ---- triangle helpers (triangle_shape.cuh) ---------------------------------------------------------------------------------------
RS_FN V3 tri_edge(const Tri& t, int i) { return i == 0 ? t.p1 - t.p0 : i == 1 ? t.p2 - t.p1 : t.p0 - t.p2; }

// ---- internal edge fix (internal_edge_utility.cuh) ------------------------------------------------------------------------------
constexpr uint32_t TIF_V0V1_CONVEX = 1u, TIF_V1V2_CONVEX = 2u, TIF_V2V0_CONVEX = 4u;
constexpr uint32_t TIF_V0V1_SWAP = 8u, TIF_V1V2_SWAP = 16u, TIF_V2V0_SWAP = 32u;
RS_FN V3 nearest_on_segment(V3 p, V3 l0, V3 l1) {  // :37-46
  V3 d = l1 - l0;
  if (length_sq(d) < k::F32_EPS * k::F32_EPS) return l0;
  float t = dot(p - l0, d) / dot(d, d);
  return l0 + d * fclamp_rs(t, ffrom(0x3727c5acu), ffrom(0x3f7fff58u));  // EDGE_T_MIN 1e-5, EDGE_T_MAX 0.99999
}
RS_FN bool clamp_normal(V3 edge, V3 tri_n, V3 n_local, float corrected, V3* out) {  // :49-66
  V3 ec = normalize(cross(edge, tri_n));
  float cur = fatan2(dot(n_local, ec), dot(n_local, tri_n));
  if ((corrected < 0.0f && cur < corrected) || (corrected >= 0.0f && cur > corrected)) {
    float diff = corrected - cur;
    float s = fsin(diff * 0.5f), cc = fcos(diff * 0.5f);
    V3 v = normalize(edge) * splat(s);
    *out = qrotate(Quat{v.x, v.y, v.z, cc}, n_local);
    return true;
  }
  return false;
}
enum { EDGE_NONE = 0, EDGE_X = 1, EDGE_Y = 2, EDGE_Z = 3 };
RS_FN void select_edge(int edge, float dist, float angle, uint32_t flags, int* best, float* best_dist) {  // :72-87
  bool sel;
  if (edge == EDGE_X) sel = dist < *best_dist;
  else if (edge == EDGE_Y) sel = dist < *best_dist || (dist == *best_dist && angle != 0.0f && (flags & TIF_V1V2_CONVEX) == 0);
  else sel = dist < *best_dist || (dist == *best_dist && angle != 0.0f);
  if (sel) {
    *best = edge;
    if (edge == EDGE_X || edge == EDGE_Y) *best_dist = dist;
  }
}
RS_FN void edge_block(ManifoldPoint* cp, const Aff& mt, const Tri& tri, int edge_no, float angle, uint32_t flags, uint32_t convex_flag,
                      uint32_t swap_flag, V3 n_local, bool recompute, bool* concave_hit) {  // :93-121
  if (angle == 0.0f) {
    *concave_hit = true;
    return;
  }
  float swap_factor = ((flags & convex_flag) != 0 ? 1.0f : 0.0f) * 2.0f - 1.0f;
  V3 edge = -tri_edge(tri, edge_no);
  V3 na = splat(swap_factor) * tri.normal;
  float s = fsin(angle * 0.5f), cc = fcos(angle * 0.5f);
  V3 axis = normalize(edge) * splat(s);
  V3 nb_raw = qrotate(Quat{axis.x, axis.y, axis.z, cc}, tri.normal);
  if (flags & swap_flag) nb_raw *= -1.0f;
  V3 nb = splat(swap_factor) * nb_raw;
  if (dot(n_local, na) < 0.0f && dot(n_local, nb) < 0.0f) {  // TIM_CONVEX_EPSILON = 0
    *concave_hit = true;
    return;
  }
  V3 nl = recompute ? mul_t(mt.m, cp->normal_on_b) : n_local;
  V3 clamped;
  if (clamp_normal(edge, splat(swap_factor) * tri.normal, nl, angle, &clamped) && dot(clamped, tri.normal) > 0.0f) {
    V3 nn = mt.m * clamped;
    cp->normal_on_b = nn;
    cp->world_on_b = cp->world_on_a - nn * cp->distance;
    cp->local_b = inv_xform(mt, cp->world_on_b);
  }
}
RS_FN void fix_internal_edge(const Np& c, ManifoldPoint* cp, uint32_t mesh_body, uint32_t tri_idx) {  // :124-189
  const MeshInfo& mesh = c.w.meshes[mesh_body];
  const TriInfo& info = c.w.infos[mesh.tri_first + tri_idx];
  const Tri& tri = c.w.tris[mesh.tri_first + tri_idx];
  const Aff mt = c.w.bodies[mesh_body].pose;
  V3 nearest = nearest_on_segment(cp->local_b, tri.p0, tri.p1);
  V3 contact = cp->local_b;
  V3 n_local = mul_t(mt.m, cp->normal_on_b);
  int best = EDGE_NONE;
  float dx = k::F32_MAX, dy = k::F32_MAX, dz = k::F32_MAX, best_dist = k::F32_MAX;
  const float MAXA = k::TAU_F;  // TIM_MAX_EDGE_ANGLE_THRESHOLD
  if (fabs_(info.edge_v0_v1_angle) < MAXA) {
    dx = length(contact - nearest);
    select_edge(EDGE_X, dx, info.edge_v0_v1_angle, info.flags, &best, &best_dist);
  }
  if (fabs_(info.edge_v1_v2_angle) < MAXA) {
    dy = length(contact - nearest_on_segment(cp->local_b, tri.p1, tri.p2));
    select_edge(EDGE_Y, dy, info.edge_v1_v2_angle, info.flags, &best, &best_dist);
  }
  if (fabs_(info.edge_v2_v0_angle) < MAXA) {
    dz = length(contact - nearest_on_segment(cp->local_b, tri.p2, tri.p0));
    select_edge(EDGE_Z, dz, info.edge_v2_v0_angle, info.flags, &best, &best_dist);
  }
  bool processed = false, concave = false;
  if (fabs_(info.edge_v0_v1_angle) < MAXA && dx < 0.1f && (best == EDGE_X || (dx < k::F32_EPS && info.edge_v0_v1_angle != 0.0f))) {
    processed = true;
    edge_block(cp, mt, tri, 0, info.edge_v0_v1_angle, info.flags, TIF_V0V1_CONVEX, TIF_V0V1_SWAP, n_local, false, &concave);
  }
  if (fabs_(info.edge_v1_v2_angle) < MAXA && dy < 0.1f && (best == EDGE_Y || (dy < k::F32_EPS && info.edge_v1_v2_angle != 0.0f))) {
    processed = true;
    edge_block(cp, mt, tri, 1, info.edge_v1_v2_angle, info.flags, TIF_V1V2_CONVEX, TIF_V1V2_SWAP, n_local, true, &concave);
  }
  if (fabs_(info.edge_v2_v0_angle) < MAXA && dz < 0.1f && (best == EDGE_Z || (dz < k::F32_EPS && info.edge_v2_v0_angle != 0.0f))) {
    processed = true;
    edge_block(cp, mt, tri, 2, info.edge_v2_v0_angle, info.flags, TIF_V2V0_CONVEX, TIF_V2V0_SWAP, n_local, true, &concave);
  }
  if (!processed || !concave) return;
  if (dot(tri.normal, n_local) < 0.0f) return;
  cp->normal_on_b = mt.m * tri.normal;
  cp->world_on_b = cp->world_on_a - cp->normal_on_b * cp->distance;
  cp->local_b = inv_xform(mt, cp->world_on_b);
}

// The tracker's contact-added callback (arena_contact_tracker.cuh:69-93): material override, the record, the edge fix.
RS_FN void contact_added(Np& c, ManifoldPoint* mp, uint32_t a, uint32_t b, int32_t tri_idx) {
  bool sw = tracker_swap(c, a, b);
  if (sw) {
    uint32_t t = a;
    a = b;
    b = t;
  }
  uint32_t ua = c.w.bodies[a].user_idx, ub = c.w.bodies[b].user_idx;
  if (ua == k::USER_CAR) {
    if (ub == k::USER_BALL) {
      mp->friction = 2.0f;  // HIT_BALL_COEFS
      mp->restitution = 0.0f;
    } else if (ub == k::USER_CAR) {
      mp->friction = 0.09f;  // HIT_CAR_COEFS
      mp->restitution = 0.1f;
    } else {
      mp->friction = 0.3f;  // HIT_WORLD_COEFS
      mp->restitution = 0.3f;
    }
  } else if (ua == k::USER_BALL && ub == k::USER_NONE && body_static(c, b)) {
    mp->is_special = 1;
  }
  push_record(c, sw, a, b, *mp);
  if (tri_idx >= 0) fix_internal_edge(c, mp, b, (uint32_t)tri_idx);
}
// persistent_manifold.rs add_contact_point (persistent_manifold.cuh:166-192). tri_idx is the mesh triangle, or -1.
RS_FN void add_contact(Np& c, Manifold& m, PointKeys& pk, uint32_t b0, uint32_t b1, V3 normal_on_b, V3 point_world, float depth,
                       int32_t tri_idx) {
  if (depth > pk.cbt) return;
  V3 pa = point_world + normal_on_b * depth;
  ManifoldPoint np = ManifoldPoint{};
  np.local_a = inv_xform(body_trans(c, b0), pa);
  np.local_b = inv_xform(body_trans(c, b1), point_world);
  np.normal_on_b = normal_on_b;
  np.distance = depth;
  np.world_on_a = pa;
  np.world_on_b = point_world;
  np.friction = combined_friction(c, b0, b1);
  np.restitution = combined_restitution(c, b0, b1);
  // The callback runs on the new point before it is stored rather than on its slot after: it reads and writes only the
  // point (never its local_a or distance, which choose the slot) and the tracker's records, so the stored point and the
  // records are the same, and the point is written to the manifold once instead of written, read back and rewritten.
  if ((body_flags(c, b0) & k::CF_CUSTOM_MATERIAL) != 0 || (body_flags(c, b1) & k::CF_CUSTOM_MATERIAL) != 0)
    contact_added(c, &np, b0, b1, tri_idx);
  add_point_keyed(m, pk, np);
}
RS_FN void add_contact(Np& c, Manifold& m, uint32_t b0, uint32_t b1, V3 normal_on_b, V3 point_world, float depth, int32_t tri_idx) {
  PointKeys pk = point_keys(c, m);
  add_contact(c, m, pk, b0, b1, normal_on_b, point_world, depth, tri_idx);
}

}  // namespace rs
