// The wheel ray query of P1: four rays per car through the ray grid, each candidate body tested in its own frame, the
// closest hit per ray kept (reference cuda/sim/raycaster.cuh, quad_ray.cuh, quad_ray_callbacks.cuh, quad_ray_shapes.cuh,
// bvh.cuh). Candidate order and tie rules decide which body wins a tie, so they follow the reference exactly.
#pragma once
#include "rs_body.h"
#include "rs_stack.h"
#include "rs_trav.h"

namespace rs {

constexpr uint32_t NO_BODY = 0xFFFFFFFFu;

RS_FN int dyn_slot(const StaticData& w, uint32_t body) {  // state slot of a body, -1 for a static one
  for (int d = 0; d < NUM_DYN; ++d)
    if (w.dyn_body[d] == body) return d;
  return -1;
}

// What the vehicle reads of a body (rigid_body_p1.cuh:26-58): a static body's pose, zero velocity and inertia.
struct BodyView {
  Aff world;
  V3 lin_vel, ang_vel;
  M3 inv_inertia_world;
  float inv_mass;
  uint32_t flags;
  uint32_t user_idx;
};
RS_FN BodyView body_view(const ArenaState& s, const StaticData& w, uint32_t body) {
  BodyView v;
  const BodyConst& c = w.bodies[body];
  int d = dyn_slot(w, body);
  v.user_idx = c.user_idx;
  v.inv_mass = c.inv_mass;
  // both transforms loaded and one picked (see body_trans: no load through a conditional address)
  const Aff dyn_world = s.bodies[d >= 0 ? d : 0].world, pose = c.pose;
  v.world = d >= 0 ? dyn_world : pose;
  if (d >= 0) {
    const Body& b = s.bodies[d];
    v.lin_vel = b.lin_vel;
    v.ang_vel = b.ang_vel;
    v.inv_inertia_world = b.inv_inertia_world;
    v.flags = b.collision_flags;
  } else {
    v.lin_vel = splat(0.0f);
    v.ang_vel = splat(0.0f);
    v.inv_inertia_world = M3{splat(0.0f), splat(0.0f), splat(0.0f)};
    v.flags = k::CF_STATIC;
  }
  return v;
}
RS_FN bool view_static(const BodyView& v) { return (v.flags & k::CF_STATIC) != 0; }
RS_FN V3 vel_at(const BodyView& v, V3 rel) { return v.lin_vel + cross(v.ang_vel, rel); }
RS_FN float impulse_denom(const BodyView& v, V3 pos, V3 n) {  // rigid_body_p1.cuh:66-71
  V3 r0 = pos - v.world.t;
  V3 c0 = cross(r0, n);
  V3 vec = cross(mul_t(v.inv_inertia_world, c0), r0);
  return v.inv_mass + dot(n, vec);
}

// ---- the ray packet (quad_ray.cuh) ------------------------------------------------------------------------------------------
struct RayPacket {
  V3 src[4], dst[4];
  float lambda_max[4];
  Aabb box;
};
RS_FN void packet_init(RayPacket& r, const V3* src, const V3* dst) {
  for (int i = 0; i < 4; ++i) {
    r.src[i] = src[i];
    r.dst[i] = dst[i];
    r.lambda_max[i] = 1.0f;
  }
  r.box.min = vmin(src[0], dst[0]);
  r.box.max = vmax(src[0], dst[0]);
  for (int i = 1; i < 4; ++i) {
    r.box.min = vmin(r.box.min, vmin(src[i], dst[i]));
    r.box.max = vmax(r.box.max, vmax(src[i], dst[i]));
  }
}
// origins[axis][lane], inverse directions with non-finite lanes replaced by FLT_MAX (quad_ray.cuh:40-55)
RS_FN void packet_pos_dir(const RayPacket& r, float org[3][4], float inv[3][4]) {
  for (int l = 0; l < 4; ++l) {
    org[0][l] = r.src[l].x;
    org[1][l] = r.src[l].y;
    org[2][l] = r.src[l].z;
    float i0 = 1.0f / (r.dst[l].x - r.src[l].x), i1 = 1.0f / (r.dst[l].y - r.src[l].y), i2 = 1.0f / (r.dst[l].z - r.src[l].z);
    inv[0][l] = ffinite(i0) ? i0 : k::F32_MAX;
    inv[1][l] = ffinite(i1) ? i1 : k::F32_MAX;
    inv[2][l] = ffinite(i2) ? i2 : k::F32_MAX;
  }
}
// slab test, bit l set when ray l overlaps the box within [0, lambda_max[l]] (quad_ray.cuh:58-85)
RS_FN uint32_t slab_mask(const float org[3][4], const float inv[3][4], const Aabb& b, const float lambda_max[4]) {
  uint32_t mask = 0;
  for (int l = 0; l < 4; ++l) {
    float t0x = (b.min.x - org[0][l]) * inv[0][l], t1x = (b.max.x - org[0][l]) * inv[0][l];
    float t0y = (b.min.y - org[1][l]) * inv[1][l], t1y = (b.max.y - org[1][l]) * inv[1][l];
    float t0z = (b.min.z - org[2][l]) * inv[2][l], t1z = (b.max.z - org[2][l]) * inv[2][l];
    float enter = mm_max(mm_max(mm_min(t0x, t1x), mm_min(t0y, t1y)), mm_min(t0z, t1z));
    float exit = mm_min(mm_min(mm_max(t0x, t1x), mm_max(t0y, t1y)), mm_max(t0z, t1z));
    if (enter <= exit && exit >= 0.0f && enter <= lambda_max[l]) mask |= 1u << l;
  }
  return mask;
}

// ---- the closest-hit callback (quad_ray_callbacks.cuh) ----------------------------------------------------------------------
struct ClosestHits {
  float frac[4];
  uint32_t body[4];
  uint32_t ignore_body, group, mask;
  float min_hit_dist_sq;
  V3 from[4], to[4];
  V3 normal[4], point[4];
};
RS_FN void hits_init(ClosestHits& h, const V3* from, const V3* to, uint32_t ignore) {
  for (int i = 0; i < 4; ++i) {
    h.frac[i] = 1.0f;
    h.body[i] = NO_BODY;
    h.from[i] = from[i];
    h.to[i] = to[i];
    h.normal[i] = splat(0.0f);
    h.point[i] = splat(0.0f);
  }
  h.ignore_body = ignore;
  h.group = k::FILTER_DEFAULT;
  h.mask = k::FILTER_ALL;
  h.min_hit_dist_sq = 0.0f;
}
RS_FN void hits_add(ClosestHits& h, uint32_t body, V3 normal_world, float frac, int i) {  // :92-101
  if (frac >= h.frac[i]) return;
  float s = 1.0f - frac;
  V3 cand = s * h.from[i] + frac * h.to[i];  // interpolate_3
  if (h.min_hit_dist_sq > 0.0f && length_sq(cand - h.from[i]) < h.min_hit_dist_sq) return;
  h.frac[i] = frac;
  h.normal[i] = normal_world;
  h.body[i] = body;
  h.point[i] = cand;
}
// One body's test: its local rays, the fractions as the callback held them when the body's test began, its rotation.
struct BodyRays {
  V3 to[4], from[4];
  float frac[4];
  M3 rot;
  uint32_t body;
  ClosestHits* hits;
};
RS_FN void body_report(BodyRays& br, V3 n_local, float frac, int i) {  // report_hit (:179-182) + internal_report_hit
  if (frac >= br.frac[i]) return;
  hits_add(*br.hits, br.body, br.rot * n_local, frac, i);
}
RS_FN void ray_triangle(BodyRays& br, const Tri& t, float* lambda_max, int i) {  // process_triangle (:187-226)
  float dist = dot(t.p0, t.normal);
  float da = dot(t.normal, br.from[i]) - dist;
  float db = dot(t.normal, br.to[i]) - dist;
  if (da * db >= 0.0f) return;
  float d = da / (da - db);
  float hf = br.frac[i];
  if (d >= hf) {
    *lambda_max = fmin_rs(*lambda_max, hf);
    return;
  }
  V3 p = lerp(br.from[i], br.to[i], d);
  const float tol = -0.0001f;
  V3 v0 = t.p0 - p, v1 = t.p1 - p;
  if (dot(cross(v0, v1), t.normal) < tol) return;
  V3 v2 = t.p2 - p;
  if (dot(cross(v1, v2), t.normal) < tol) return;
  if (dot(cross(v2, v0), t.normal) < tol) return;
  float prev = br.hits->frac[i];
  hits_add(*br.hits, br.body, br.rot * (da <= 0.0f ? -t.normal : t.normal), d, i);
  if (br.hits->frac[i] < prev) *lambda_max = br.hits->frac[i];
}

// ---- shapes (quad_ray_shapes.cuh) -------------------------------------------------------------------------------------------
// compound_shape.rs solve_rounded_box (:15-125): entry of a ray into the box with rounded edges, or false
RS_FN bool rounded_box_hit(V3 src, V3 dir, float dist, V3 half, float margin, float* t_out, V3* n_out) {
  float pts[8];
  for (int i = 0; i < 8; ++i) pts[i] = 0.0f;
  pts[1] = dist;
  int cnt = 2;
  for (int ax = 0; ax < 3; ++ax) {
    float d = comp(dir, ax);
    if (d == 0.0f) continue;
    float s = comp(src, ax), h = comp(half, ax);
    for (int bi = 0; bi < 2; ++bi) {
      float t = ((bi == 0 ? fneg(h) : h) - s) / d;
      if (t > 0.0f && t < dist && cnt < 8) pts[cnt++] = t;
    }
  }
  for (int i = 1; i < cnt; ++i) {  // insertion sort
    float key = pts[i];
    int j = i;
    while (j > 0 && pts[j - 1] > key) {
      pts[j] = pts[j - 1];
      --j;
    }
    pts[j] = key;
  }
  float m2 = margin * margin;
  for (int i = 0; i + 1 < cnt; ++i) {
    float t0 = pts[i], t1 = pts[i + 1];
    if (t1 <= t0) continue;
    float mid = t0 + (t1 - t0) * 0.5f;
    V3 pm = src + dir * mid;
    float a = 0.0f, b = 0.0f, c = 0.0f;
    uint32_t outside = 0;
    for (int ax = 0; ax < 3; ++ax) {
      float p = comp(pm, ax), h = comp(half, ax), bound;
      if (p < fneg(h)) bound = fneg(h);
      else if (p > h) bound = h;
      else continue;
      ++outside;
      float o = comp(src, ax) - bound, d = comp(dir, ax);
      a = a + d * d;
      b = b + (2.0f * d) * o;
      c = c + o * o;
    }
    if (outside == 0) continue;
    float cc = c - m2;
    float disc = b * b - (4.0f * a) * cc;
    if (disc < 0.0f) {
      float scale = b * b + fabs_((4.0f * a) * cc);
      if (disc > fneg(k::F32_EPS) * scale) disc = 0.0f;
      else continue;
    }
    float sq = __builtin_sqrtf(disc);
    float th;
    if (disc == 0.0f) {
      th = fneg(b) / (2.0f * a);
    } else {
      float q = -0.5f * (b >= 0.0f ? b + sq : b - sq);
      if (q == 0.0f) {
        th = fneg(b) / (2.0f * a);
      } else {
        float r0 = q / a, r1 = cc / q;
        th = r0 <= r1 ? r0 : r1;
      }
    }
    if (th <= 0.0f || th > dist || th < t0 || th > t1) continue;
    V3 hit = src + dir * th;
    V3 n = hit - vclamp(hit, -half, half);
    float l2 = length_sq(n);
    if (l2 <= k::F32_EPS * k::F32_EPS) continue;
    V3 nn = n / __builtin_sqrtf(l2);
    if (dot(nn, dir) > 0.0f) continue;
    *t_out = th;
    *n_out = nn;
    return true;
  }
  return false;
}
// compound_shape.rs internal_perform_raycast (:149-240): the car's box, one ray
RS_FN void compound_ray(const BodyConst& bc, BodyRays& br, V3 src, V3 dst, int i) {
  const Aabb& box = bc.compound_local_aabb;
  V3 delta = dst - src;
  float dist = length(delta);
  V3 dir = delta / dist;
  float tenter = 0.0f, texit = dist;
  int hit_axis = 0;
  V3 inv = 1.0f / dir;
  V3 t1 = (box.min - src) * inv, t2 = (box.max - src) * inv;
  V3 tmin = vmin(t1, t2), tmax = vmax(t1, t2);
  uint32_t neg = (fbits(tmax.x) >> 31) | ((fbits(tmax.y) >> 31) << 1) | ((fbits(tmax.z) >> 31) << 2);
  B3 fin = finite_mask(inv);
  uint32_t finite[3] = {fin.x, fin.y, fin.z};
  for (int ax = 0; ax < 3; ++ax) {
    if (!finite[ax]) {
      float o = comp(src, ax);
      if (comp(box.min, ax) > o || o > comp(box.max, ax)) return;
      continue;
    }
    if (neg & (1u << ax)) return;
    texit = fmin_rs(texit, comp(tmax, ax));
    if (comp(tmin, ax) > tenter) {
      tenter = comp(tmin, ax);
      hit_axis = ax;
    }
    if (tenter > texit) return;
  }
  if (0.0f >= tenter || tenter > dist) return;
  V3 half = bc.box_half_extents;
  {  // a face entry (the child transform is a translation)
    V3 local = (src + dir * tenter) - bc.child_trans.t;
    bool inside = true;
    for (int ax = 0; ax < 3; ++ax) {
      if (ax == hit_axis) continue;
      if (fabs_(comp(local, ax)) > comp(half, ax)) {
        inside = false;
        break;
      }
    }
    if (inside) {
      V3 n = splat(0.0f);
      set_comp(n, hit_axis, fneg(fsignum_rs(comp(dir, hit_axis))));
      if (length_sq(n) > 0.0f) body_report(br, n, fmax_rs(tenter, 0.0f) / dist, i);
      return;
    }
  }
  // an edge or corner entry: analytic_rounded_hit (:193-211)
  float th;
  V3 n;
  if (!rounded_box_hit(src - bc.child_trans.t, dir, dist, half, bc.box_margin, &th, &n)) return;
  n = n / length(n);
  if (dot(n, dir) > 0.0f) return;
  if (0.0f >= th || th > dist) return;
  body_report(br, n, th / dist, i);
}
RS_FN void compound_rays(const BodyConst& bc, BodyRays& br, const RayPacket& rp) {  // :244-256
  const Aabb& box = bc.compound_local_aabb;
  if (!overlaps(rp.box, box)) return;
  float org[3][4], inv[3][4];
  packet_pos_dir(rp, org, inv);
  uint32_t mask = slab_mask(org, inv, box, br.frac);
  RS_UNROLL for (int i = 0; i < 4; ++i)
    if (mask & (1u << i)) compound_ray(bc, br, rp.src[i], rp.dst[i], i);
}
RS_FN void sphere_rays(float radius, BodyRays& br, const RayPacket& rp) {  // sphere_shape.rs (:261-337)
  float sx[4], sy[4], sz[4], tx[4], ty[4], tz[4];
  for (int l = 0; l < 4; ++l) {
    sx[l] = rp.src[l].x; sy[l] = rp.src[l].y; sz[l] = rp.src[l].z;
    tx[l] = rp.dst[l].x; ty[l] = rp.dst[l].y; tz[l] = rp.dst[l].z;
  }
  float r2 = radius * radius;
  bool within[4], any = false;
  for (int l = 0; l < 4; ++l) {
    float cx = mm_min(mm_max(0.0f, mm_min(sx[l], tx[l])), mm_max(sx[l], tx[l]));
    float cy = mm_min(mm_max(0.0f, mm_min(sy[l], ty[l])), mm_max(sy[l], ty[l]));
    float cz = mm_min(mm_max(0.0f, mm_min(sz[l], tz[l])), mm_max(sz[l], tz[l]));
    within[l] = ((cx * cx + cy * cy) + cz * cz) <= r2;
    any = any || within[l];
  }
  if (!any) return;
  float dist[4], dx[4], dy[4], dz[4], disc[4];
  bool valid[4];
  any = false;
  for (int l = 0; l < 4; ++l) {
    float ex = tx[l] - sx[l], ey = ty[l] - sy[l], ez = tz[l] - sz[l];
    dist[l] = __builtin_sqrtf((ex * ex + ey * ey) + ez * ez);
    dx[l] = ex / dist[l];
    dy[l] = ey / dist[l];
    dz[l] = ez / dist[l];
    float b = 2.0f * ((sx[l] * dx[l] + sy[l] * dy[l]) + sz[l] * dz[l]);
    float c = ((sx[l] * sx[l] + sy[l] * sy[l]) + sz[l] * sz[l]) - r2;
    disc[l] = b * b - 4.0f * c;
    valid[l] = disc[l] >= 0.0f && within[l];
    any = any || valid[l];
  }
  if (!any) return;
  float t[4];
  bool nearest[4];
  any = false;
  for (int l = 0; l < 4; ++l) {
    float b = 2.0f * ((sx[l] * dx[l] + sy[l] * dy[l]) + sz[l] * dz[l]);
    float sq = __builtin_sqrtf(disc[l]);
    float t0 = (fneg(b) - sq) / 2.0f, t1 = (fneg(b) + sq) / 2.0f;
    bool sw = t0 > t1;
    t0 = sw ? t1 : t0;
    t1 = sw ? t0 : t1;  // as V3: the second select reads the updated t0
    bool n0 = t0 >= 0.0f && t0 <= dist[l];
    bool n1 = t1 >= 0.0f && t1 <= dist[l] && !n0;
    nearest[l] = valid[l] && (n0 || n1);
    any = any || nearest[l];
    t[l] = n0 ? t0 : t1;
  }
  if (!any) return;
  for (int l = 0; l < 4; ++l) {
    if (!nearest[l]) continue;
    V3 hp = V3{sx[l] + dx[l] * t[l], sy[l] + dy[l] * t[l], sz[l] + dz[l] * t[l]};
    body_report(br, V3{hp.x / radius, hp.y / radius, hp.z / radius}, t[l] / dist[l], l);
  }
}
RS_FN void plane_rays(const Plane& pl, BodyRays& br, const RayPacket& rp) {  // static_plane_shape.rs (:340-386)
  if (!overlaps(rp.box, pl.aabb_ident)) return;
  float sx[4], sy[4], sz[4], dx[4], dy[4], dz[4], org[3][4], inv[3][4];
  for (int l = 0; l < 4; ++l) {
    sx[l] = rp.src[l].x; sy[l] = rp.src[l].y; sz[l] = rp.src[l].z;
    dx[l] = rp.dst[l].x - sx[l];
    dy[l] = rp.dst[l].y - sy[l];
    dz[l] = rp.dst[l].z - sz[l];
    org[0][l] = sx[l]; org[1][l] = sy[l]; org[2][l] = sz[l];
    inv[0][l] = 1.0f / dx[l];
    inv[1][l] = 1.0f / dy[l];
    inv[2][l] = 1.0f / dz[l];
  }
  uint32_t rmask = slab_mask(org, inv, pl.aabb_ident, br.frac);
  if (rmask == 0) return;
  V3 n = pl.normal;
  float dist[4], align[4], t[4];
  uint32_t amask = 0, tmask = 0;
  for (int l = 0; l < 4; ++l) {
    dist[l] = __builtin_sqrtf((dx[l] * dx[l] + dy[l] * dy[l]) + dz[l] * dz[l]);
    float id = 1.0f / dist[l];
    float ux = dx[l] * id, uy = dy[l] * id, uz = dz[l] * id;
    align[l] = (ux * n.x + uy * n.y) + uz * n.z;
    if (fabs_(align[l]) >= k::F32_EPS) amask |= 1u << l;
  }
  if (amask == 0) return;
  for (int l = 0; l < 4; ++l) {
    float ns = (sx[l] * n.x + sy[l] * n.y) + sz[l] * n.z;
    t[l] = fneg(ns) / align[l];
    if (t[l] >= 0.0f && t[l] < dist[l]) tmask |= 1u << l;
  }
  uint32_t hm = rmask & (amask & tmask);
  RS_UNROLL for (int l = 0; l < 4; ++l)
    if (hm & (1u << l)) body_report(br, n, t[l] / dist[l], l);
}

// mesh rays: depth-first BVH4 walk, overlapping children pushed in reverse lane order (quad_ray_shapes.cuh:377-406)
RS_FN void mesh_rays(const StaticData& w, const MeshInfo& mesh, BodyRays& br, RayPacket& rp) {
  if (!overlaps(rp.box, mesh.root_aabb)) return;
  float org[3][4], inv[3][4];
  packet_pos_dir(rp, org, inv);
  TravStack st;
  st.n = 0;
  st.push(0);
  while (st.n != 0) {
    uint32_t work = st.pop();
    if (work & BVH_LEAF) {
      const Leaf& leaf = w.leaves[mesh.leaf_first + (work & ~BVH_LEAF)];
      uint32_t mask = slab_mask(org, inv, leaf.bounds, rp.lambda_max);
      if (mask != 0) {
        const Tri& tri = w.tris[mesh.tri_first + leaf.tri];
        RS_UNROLL for (int i = 0; i < 4; ++i)
          if (mask & (1u << i)) ray_triangle(br, tri, &rp.lambda_max[i], i);
      }
      continue;
    }
    const WideNode& node = w.nodes[mesh.node_first + work];
    uint32_t mask = node_mask(node, rp.box);
    for (int lane = (int)node.child_count - 1; lane >= 0; --lane)
      if (mask & (1u << lane)) st.push(node.children[lane]);
  }
}
// mesh rays over a leaf list walked beforehand: the leaves mesh_rays' own walk reaches (its node test reads only the
// packet's box), in the same depth-first order, each tested against the rays as they stand when it is reached
RS_FN void mesh_rays_leaves(const StaticData& w, const MeshInfo& mesh, BodyRays& br, RayPacket& rp, const uint32_t* leaves,
                            uint32_t count) {
  if (count == 0) return;
  float org[3][4], inv[3][4];
  packet_pos_dir(rp, org, inv);
  // the next leaf is loaded before this one's tests, and a leaf's triangle before its slab test, so the triangle's load
  // overlaps the test instead of following it (loads only: the same sequence of tests)
  Leaf next = w.leaves[mesh.leaf_first + leaves[0]];
  for (uint32_t k = 0; k < count; ++k) {
    const Leaf leaf = next;
    if (k + 1 < count) next = w.leaves[mesh.leaf_first + leaves[k + 1]];
    const Tri tri = w.tris[mesh.tri_first + leaf.tri];
    uint32_t mask = slab_mask(org, inv, leaf.bounds, rp.lambda_max);
    if (mask != 0) {
      RS_UNROLL for (int i = 0; i < 4; ++i)
        if (mask & (1u << i)) ray_triangle(br, tri, &rp.lambda_max[i], i);
    }
  }
}

// The first mesh (plane) record at or after `from` that belongs to `body`; NUM_MESHES (NUM_PLANES) when none does.
RS_FN int next_mesh_of(const StaticData& w, uint32_t body, int from) {
  int found = NUM_MESHES;
  RS_UNROLL
  for (int mi = NUM_MESHES - 1; mi >= 0; --mi)
    if (mi >= from && w.meshes[mi].body == body) found = mi;
  return found;
}
RS_FN int next_plane_of(const StaticData& w, uint32_t body, int from) {
  int found = NUM_PLANES;
  RS_UNROLL
  for (int p = NUM_PLANES - 1; p >= 0; --p)
    if (p >= from && w.planes[p].body == body) found = p;
  return found;
}

// collision_world.rs quad_ray_test (raycaster.cuh:20-70): the rays into the body's frame, then its shape
RS_FN void body_ray_setup(const ArenaState& s, const StaticData& w, uint32_t body, ClosestHits& h, BodyRays& br,
                          RayPacket& rp) {
  BodyView co = body_view(s, w, body);
  const Aff& tr = co.world;
  V3 from[4], to[4];
  const M3& m = tr.m;
  bool identity = m.c0.x == 1.0f && m.c0.y == 0.0f && m.c0.z == 0.0f && m.c1.x == 0.0f && m.c1.y == 1.0f && m.c1.z == 0.0f &&
                  m.c2.x == 0.0f && m.c2.y == 0.0f && m.c2.z == 1.0f;
  if (identity) {
    for (int i = 0; i < 4; ++i) {
      from[i] = h.from[i] - tr.t;
      to[i] = h.to[i] - tr.t;
    }
  } else {
    Aff inv = aff_transpose(tr);
    for (int i = 0; i < 4; ++i) {
      from[i] = xform_point(inv, h.from[i]);
      to[i] = xform_point(inv, h.to[i]);
    }
  }
  for (int i = 0; i < 4; ++i) {
    br.from[i] = from[i];
    br.to[i] = to[i];
    br.frac[i] = h.frac[i];
  }
  br.rot = tr.m;
  br.body = body;
  br.hits = &h;
  packet_init(rp, from, to);
  for (int i = 0; i < 4; ++i) rp.lambda_max[i] = br.frac[i];
}
RS_FN bool rays_walk_mesh(uint32_t shape_kind) {
  return shape_kind != k::SHAPE_COMPOUND && shape_kind != k::SHAPE_SPHERE && shape_kind != k::SHAPE_PLANE;
}
// `tb` (may be null): mesh walks done beforehand, taken in order through *next_job up to end_job (ray_jobs collects them)
RS_FN void body_rays(const ArenaState& s, const StaticData& w, uint32_t body, ClosestHits& h, const TravBatch* tb,
                     uint32_t* next_job, uint32_t end_job) {
  const BodyConst& bc = w.bodies[body];
  BodyRays br;
  RayPacket rp;
  body_ray_setup(s, w, body, h, br, rp);
  if (bc.shape_kind == k::SHAPE_COMPOUND) {
    compound_rays(bc, br, rp);
  } else if (bc.shape_kind == k::SHAPE_SPHERE) {
    sphere_rays(bc.sphere_margin, br, rp);
  } else if (bc.shape_kind == k::SHAPE_PLANE) {
    for (int p = next_plane_of(w, body, 0); p < NUM_PLANES; p = next_plane_of(w, body, p + 1)) plane_rays(w.planes[p], br, rp);
  } else {
    // every mesh record of the body, in record order, from its walked leaf list when the batch holds one
    for (int mi = next_mesh_of(w, body, 0); mi < NUM_MESHES; mi = next_mesh_of(w, body, mi + 1)) {
      const TravJob* jb = nullptr;
      if (tb && *next_job < end_job && tb->job[*next_job].body == body && tb->job[*next_job].mesh_body == (uint32_t)mi)
        jb = &tb->job[(*next_job)++];
      if (jb && !jb->in_place) mesh_rays_leaves(w, w.meshes[mi], br, rp, tb->leaf_tri + jb->first, jb->count);
      else mesh_rays(w, w.meshes[mi], br, rp);
    }
  }
}

// grid_broadphase.rs get_cell_idx (raycaster.cuh:73-86)
RS_FN uint32_t grid_cell_of(const StaticData& w, V3 pos) {
  V3 f = (pos - w.grid_min) / w.grid_cell_size;
  int32_t ix = f2i_rs(f.x), iy = f2i_rs(f.y), iz = f2i_rs(f.z);
  uint32_t x = (uint32_t)(ix > 0 ? ix : 0), y = (uint32_t)(iy > 0 ? iy : 0), z = (uint32_t)(iz > 0 ? iz : 0);
  uint32_t nx = w.grid_num_cells[0], ny = w.grid_num_cells[1], nz = w.grid_num_cells[2];
  x = x < nx - 1 ? x : nx - 1;
  y = y < ny - 1 ? y : ny - 1;
  z = z < nz - 1 ? z : nz - 1;
  return x * ny * nz + y * nz + z;
}
RS_FN bool ray_candidate_passes(const ProxyConst& p, const ClosestHits& hits) {  // :89-95
  return hits.ignore_body != p.client_body && (p.filter_group & hits.mask) != 0 && (hits.group & p.filter_mask) != 0;
}
RS_FN void ray_candidate(const ArenaState& s, const StaticData& w, uint32_t h, ClosestHits& hits, const TravBatch* tb,
                         uint32_t* next_job, uint32_t end_job) {
  const ProxyConst& p = w.proxies[h];
  if (ray_candidate_passes(p, hits)) body_rays(s, w, p.client_body, hits, tb, next_job, end_job);
}
// grid_broadphase.rs ray_test (raycaster.cuh:100-148), in three steps. The screen: the cells of the four ray sources in
// ray order; in each cell its static handles, then its dynamic ones; every handle at most once, kept if its box overlaps
// the rays' box. The screen runs over all cells before any candidate's ray test: a ray test changes only the hits, which
// the screen does not read, so this is the sequence of screening and testing entry by entry, with the screen's loads no
// longer waiting behind the tests. Returns the number of candidates written to cand (at most 32).
// A handle's box test does not depend on the cell it is met in, so every handle is tested once, up front, into a mask (the
// world's box for a static handle, the proxy's for a dynamic one); the cells' ranges, their runs in the dynamic list
// (bisections side by side) and their static entries (eight at a time) are then read as independent loads, and only the
// walk in entry order, with its at-most-once rule, stays sequential.
RS_FN int ray_screen(const ArenaState& s, const StaticData& w, const ClosestHits& hits, uint32_t* cand) {
  static_assert(NUM_HANDLES <= 32, "one bit per handle");
  Aabb rb;
  rb.min = vmin(hits.from[0], hits.to[0]);
  rb.max = vmax(hits.from[0], hits.to[0]);
  for (int i = 1; i < 4; ++i) {
    rb.min = vmin(rb.min, vmin(hits.from[i], hits.to[i]));
    rb.max = vmax(rb.max, vmax(hits.from[i], hits.to[i]));
  }
  uint32_t cells[4] = {0, 0, 0, 0};
  int nc = 0;
  for (int r = 0; r < 4; ++r) {
    uint32_t c = grid_cell_of(w, hits.from[r]);
    bool found = false;
    for (int q = 0; q < nc; ++q) found = found || cells[q] == c;
    if (!found) cells[nc++] = c;
  }
  uint32_t over = 0;
  RS_UNROLL
  for (int h = 0; h < NUM_HANDLES; ++h) over |= overlaps(rb, w.proxies[h].aabb) ? 1u << h : 0u;
  RS_UNROLL
  for (int d = NUM_DYN - 1; d >= 0; --d) {  // a handle's first dynamic slot gives its box
    const uint32_t bit = 1u << w.dyn_handle[d];
    over = overlaps(rb, s.dyn_proxy[d].aabb) ? over | bit : over & ~bit;
  }
  // the cell's dynamic entries: the list is sorted by cell (an invariant of the format, dump.rs:683-685), so its run
  // is found by bisection and walked in list order, the entries a full scan would visit, in the same order
  const uint32_t nd = s.n_dyn_cells < (uint32_t)MAX_DYN_CELLS ? s.n_dyn_cells : (uint32_t)MAX_DYN_CELLS;
  CellRange cr[4];
  uint32_t lo[4], hi[4];
  RS_UNROLL
  for (int q = 0; q < 4; ++q) {
    cr[q] = q < nc ? w.cells[cells[q]] : CellRange{0, 0};
    lo[q] = 0;
    hi[q] = q < nc ? nd : 0;
  }
  for (bool busy = true; busy;) {
    busy = false;
    RS_UNROLL
    for (int q = 0; q < 4; ++q) {
      if (lo[q] < hi[q]) {
        uint32_t mid = (lo[q] + hi[q]) >> 1;
        if (s.dyn_cells[mid].cell < cells[q]) lo[q] = mid + 1;
        else hi[q] = mid;
        busy = true;
      }
    }
  }
  uint32_t seen = 0;  // distinct handles: at most one per bit of `seen`, so at most 32 candidates
  int ncand = 0;
  RS_UNROLL
  for (int q = 0; q < 4; ++q) {
    if (q >= nc) break;
    const uint32_t cnt = cr[q].count;
    for (uint32_t e0 = 0; e0 < cnt; e0 += 8) {
      uint32_t hs[8];
      RS_UNROLL
      for (int j = 0; j < 8; ++j) hs[j] = e0 + j < cnt ? w.cell_handles[cr[q].first + e0 + j] : 0u;
      RS_UNROLL
      for (int j = 0; j < 8; ++j) {
        if (e0 + j >= cnt) break;
        const uint32_t bit = 1u << hs[j];
        if (seen & bit) continue;
        seen |= bit;
        if (over & bit) cand[ncand++] = hs[j];
      }
    }
    for (uint32_t e = lo[q]; e < nd && s.dyn_cells[e].cell == cells[q]; ++e) {
      const uint32_t h = s.dyn_cells[e].handle, bit = 1u << h;
      if (seen & bit) continue;
      seen |= bit;
      if (over & bit) cand[ncand++] = h;
    }
  }
  return ncand;
}
// The candidates' mesh walks, in the order ray_tests reaches them, appended to tb (as many as fit); a walk's node test reads
// only the packet's box in the mesh frame, which the hits do not change, so the walks can run before the tests (on the GPU
// pooled across the wave's lanes).
RS_FN void ray_jobs(const ArenaState& s, const StaticData& w, ClosestHits& hits, const uint32_t* cand, int ncand,
                    TravBatch& tb) {
  for (int k = 0; k < ncand && tb.n_jobs < (uint32_t)TJ_MAX; ++k) {
    const ProxyConst& p = w.proxies[cand[k]];
    const uint32_t body = p.client_body;
    if (!ray_candidate_passes(p, hits) || !rays_walk_mesh(w.bodies[body].shape_kind)) continue;
    BodyRays br;
    RayPacket rp;
    body_ray_setup(s, w, body, hits, br, rp);
    for (int mi = next_mesh_of(w, body, 0); mi < NUM_MESHES && tb.n_jobs < (uint32_t)TJ_MAX; mi = next_mesh_of(w, body, mi + 1)) {
      TravJob& jb = tb.job[tb.n_jobs++];
      jb.kind = TJ_RAYS;
      jb.body = body;
      jb.mesh_body = (uint32_t)mi;
      jb.box = rp.box;
    }
  }
}
// The candidates' ray tests in screen order, mesh candidates taking their walked leaves from tb[first_job, end_job).
RS_FN void ray_tests(const ArenaState& s, const StaticData& w, ClosestHits& hits, const uint32_t* cand, int ncand,
                     const TravBatch& tb, uint32_t first_job, uint32_t end_job) {
  uint32_t next_job = first_job;
  for (int k = 0; k < ncand; ++k) ray_candidate(s, w, cand[k], hits, &tb, &next_job, end_job);
}

struct WheelHit {
  uint32_t some;
  V3 point, normal;
  uint32_t body;
};
RS_FN void wheel_hits_init(ClosestHits& hits, const V3* from, const V3* to, uint32_t ignore) {
  hits_init(hits, from, to, ignore);
  hits.group |= k::FILTER_DROPSHOT_FLOOR;
  hits.min_hit_dist_sq = 0.116684f;  // WHEEL_RAY_MIN_HIT_DIST_SQ
}
// Both cars' wheel-ray queries of a tick, screened and walked together at the start of P1 (one pooled walk per tick). A
// car's packets are fixed by its chassis transform and wheel geometry, its candidates by the grid and the dynamic proxies
// and cells: P1 writes none of these before the car's rays are cast (a car's transform moves only when it respawns, and a
// respawning car casts no rays that tick), so the screen and the walks done early are the ones done at the cast.
struct WheelRayPrep {
  TravBatch tb;
  uint32_t cand[NUM_CARS][32];
  int ncand[NUM_CARS];
  uint32_t first_job[NUM_CARS], end_job[NUM_CARS];
};
// vehicle/raycaster.rs cast_rays (raycaster.cuh:153-176), the screen and walks taken from the car's share of `prep`
RS_FN void cast_wheel_rays(const ArenaState& s, const StaticData& w, const V3* from, const V3* to, uint32_t ignore,
                           const WheelRayPrep& prep, int ci, WheelHit out[4]) {
  ClosestHits hits;
  wheel_hits_init(hits, from, to, ignore);
  ray_tests(s, w, hits, prep.cand[ci], prep.ncand[ci], prep.tb, prep.first_job[ci], prep.end_job[ci]);
  RS_UNROLL for (int i = 0; i < 4; ++i) {
    out[i].some = 0;
    out[i].point = splat(0.0f);
    out[i].normal = splat(0.0f);
    out[i].body = 0;
    if (hits.body[i] != NO_BODY) {
      BodyView v = body_view(s, w, hits.body[i]);
      if ((v.flags & k::CF_NO_CONTACT_RESPONSE) == 0) {
        out[i].some = 1;
        out[i].body = hits.body[i];
        out[i].point = hits.point[i];
        out[i].normal = normalize_or_zero(hits.normal[i]);
      }
    }
  }
}

}  // namespace rs
