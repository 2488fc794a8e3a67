// Phase P3b, the narrowphase: for each retained broadphase pair in pair order, the near callback's gates, then the pair's
// algorithm (ball-car sphere-OBB, car-car box-box, ball/car against a plane, ball/car against an arena mesh through its
// BVH4 one leaf at a time), the manifold merge or insert, and the active list. Reference: cuda/sim/collision_dispatcher.cuh,
// obb_obb_collision_alg.cuh, sphere_obb_collision_alg.cuh, convex_plane_collision_alg.cuh, compound_collision_alg.cuh,
// sphere_concave_collision_alg.cuh, triangle_shape.cuh, bvh.cuh.
#pragma once
#include "rs_sat.h"

namespace rs {

// ---- box-box (obb_obb_collision_alg.cuh, V3 box_box_detector.rs) -----------------------------------------------------------------
struct Obb { V3 center; M3 axis; V3 extent; };
struct BoxHit { float depth; V3 normal; uint32_t code; };
struct Poly16 { float p[16][2]; uint32_t n; };

RS_FN float line_closest_approach(V3 pa, V3 ua, V3 pb, V3 ub) {  // :20-28
  V3 p = pb - pa;
  float uaub = dot(ua, ub), q1 = dot(ua, p), q2 = fneg(dot(ub, p));
  float d = 1.0f - uaub * uaub;
  if (d <= 0.0001f) return 0.0f;
  return (uaub * q1 + q2) / d;
}
RS_FN Poly16 clip_rect_quad(const float h[2], const float quad[4][2]) {  // :31-70 intersect_rect_quad2
  Poly16 q;
  q.n = 4;
  for (int i = 0; i < 4; ++i) {
    q.p[i][0] = quad[i][0];
    q.p[i][1] = quad[i][1];
  }
  for (int dir = 0; dir < 2; ++dir)
    for (int si = 0; si < 2; ++si) {
      float sign = si == 0 ? -1.0f : 1.0f;
      Poly16 r;
      r.n = 0;
      float clip = sign * h[dir];
      uint32_t len = q.n;
      for (uint32_t i = 0; i < len; ++i) {
        const float* cur = q.p[i];
        const float* nxt = q.p[(i + 1) % len];
        bool in_cur = sign * cur[dir] < h[dir];
        if (in_cur) {
          r.p[r.n][0] = cur[0];
          r.p[r.n][1] = cur[1];
          r.n++;
        }
        bool in_nxt = sign * nxt[dir] < h[dir];
        if (in_cur != in_nxt) {
          float pv = cur[1 - dir] + ((nxt[1 - dir] - cur[1 - dir]) / (nxt[dir] - cur[dir])) * (clip - cur[dir]);
          r.p[r.n][0] = dir == 0 ? clip : pv;
          r.p[r.n][1] = dir == 0 ? pv : clip;
          r.n++;
        }
      }
      q = r;
    }
  return q;
}
RS_FN void cull_points(const float p[][2], uint32_t n, uint32_t i0, uint32_t m, uint32_t* res) {  // :73-124 cull_points2
  float cx, cy;
  if (n == 1) {
    cx = p[0][0];
    cy = p[0][1];
  } else if (n == 2) {
    cx = (p[0][0] + p[1][0]) * 0.5f;
    cy = (p[0][1] + p[1][1]) * 0.5f;
  } else {
    float a = 0.0f, sx = 0.0f, sy = 0.0f;
    for (uint32_t i = 0; i + 1 < n; ++i) {
      float x0 = p[i][0], y0 = p[i][1], x1 = p[i + 1][0], y1 = p[i + 1][1];
      float q = x0 * y1 - x1 * y0;
      a = a + q;
      sx = sx + q * (x0 + x1);
      sy = sy + q * (y0 + y1);
    }
    float xl = p[n - 1][0], yl = p[n - 1][1], x0 = p[0][0], y0 = p[0][1];
    float ql = xl * y0 - x0 * yl;
    float area = a + ql;
    float inv = fabs_(area) > k::F32_EPS ? 1.0f / (3.0f * area) : 1e18f;
    cx = inv * (sx + ql * (xl + x0));
    cy = inv * (sy + ql * (yl + y0));
  }
  float ang[8];
  for (uint32_t i = 0; i < n; ++i) ang[i] = fatan2(p[i][1] - cy, p[i][0] - cx);
  bool avail[8] = {false, false, false, false, false, false, false, false};
  for (uint32_t i = 0; i < n; ++i) avail[i] = true;
  avail[i0] = false;
  res[0] = i0;
  for (uint32_t j = 1; j < m; ++j) {
    float target = (float)j * (k::TAU_F / (float)m) + ang[i0];
    if (target > k::PI_F) target = target - k::TAU_F;
    uint32_t best = i0;
    float bd = k::F32_MAX;
    for (uint32_t i = 0; i < n; ++i) {
      if (!avail[i]) continue;
      float d = fabs_(ang[i] - target);
      if (d > k::PI_F) d = k::TAU_F - d;
      if (d < bd) {
        bd = d;
        best = i;
      }
    }
    avail[best] = false;
    res[j] = best;
  }
}
struct BoxSat {  // the running best axis of box_box_sat (:127-160)
  float s, s2;
  bool normal_is_face;
  V3 nr, nc;
  bool invert;
  uint32_t code;
  RS_FN bool face(float e1, float e2, V3 n, uint32_t cc) {
    s2 = fabs_(e1) - e2;
    if (s2 > 0.0f) return false;
    if (s2 > s) {
      s = s2;
      normal_is_face = true;
      nr = n;
      invert = e1 < 0.0f;
      code = cc;
    }
    return true;
  }
  RS_FN bool edge(float e1, float e2, V3 n, uint32_t cc) {
    s2 = fabs_(e1) - e2;
    if (s2 > k::F32_EPS) return false;
    float l = length(n);
    if (l > k::F32_EPS) {
      s2 = s2 / l;
      if (s2 * 1.05f > s) {
        s = s2;
        normal_is_face = false;
        nc = n / l;
        invert = e1 < 0.0f;
        code = cc;
      }
    }
    return true;
  }
};
RS_FN bool box_box_sat(const Obb& o1, const M3& r1t, const Obb& o2, BoxHit* hit) {  // :163-216
  V3 p = o2.center - o1.center;
  V3 pp = r1t * p;
  M3 rij = r1t * o2.axis;
  V3 rc[3] = {rij.c0, rij.c1, rij.c2};
  M3 qm = mabs(rij);
  V3 qc[3] = {qm.c0, qm.c1, qm.c2};
  M3 qt = transpose(qm);
  V3 qr[3] = {qt.c0, qt.c1, qt.c2};
  BoxSat st;
  st.s = ffrom(0xFF800000u);
  st.s2 = 0.0f;
  st.normal_is_face = false;
  st.nr = splat(0.0f);
  st.nc = splat(0.0f);
  st.invert = false;
  st.code = 0;
  V3 a1[3] = {o1.axis.c0, o1.axis.c1, o1.axis.c2};
  for (int i = 0; i < 3; ++i)
    if (!st.face(comp(pp, i), comp(o1.extent, i) + dot(o2.extent, qr[i]), a1[i], (uint32_t)(i + 1))) return false;
  V3 a2[3] = {o2.axis.c0, o2.axis.c1, o2.axis.c2};
  for (int i = 0; i < 3; ++i)
    if (!st.face(dot(a2[i], p), dot(o1.extent, qc[i]) + comp(o2.extent, i), a2[i], (uint32_t)(i + 4))) return false;
  for (int i = 0; i < 3; ++i) qc[i] = qc[i] + splat(1e-5f);
  qt = transpose(M3{qc[0], qc[1], qc[2]});
  qr[0] = qt.c0;
  qr[1] = qt.c1;
  qr[2] = qt.c2;
  for (int i = 0; i < 3; ++i) {
    int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
    for (int j = 0; j < 3; ++j) {
      int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
      V3 n = cross(unit_axis(i), rc[j]);
      float e2 = ((comp(o1.extent, i1) * comp(qc[j], i2) + comp(o1.extent, i2) * comp(qc[j], i1)) + comp(o2.extent, j1) * comp(qr[i], j2)) +
                 comp(o2.extent, j2) * comp(qr[i], j1);
      if (!st.edge(dot(pp, n), e2, n, (uint32_t)(7 + i * 3 + j))) return false;
    }
  }
  if (st.code == 0) return false;
  V3 normal = st.normal_is_face ? st.nr : o1.axis * st.nc;
  if (st.invert) normal = -normal;
  hit->depth = fneg(st.s);
  hit->normal = normal;
  hit->code = st.code;
  return true;
}
RS_FN void box_box_contacts(Np& c, Obb o1, M3 r1t, Obb o2, M3 r2t, const BoxHit& hit, Manifold& m, uint32_t b1, uint32_t b2) {  // :219-376
  V3 a1[3] = {r1t.c0, r1t.c1, r1t.c2}, a2[3] = {r2t.c0, r2t.c1, r2t.c2};
  if (hit.code > 6) {  // edge-edge: one contact between the two edges
    V3 pa = o1.center;
    for (int i = 0; i < 3; ++i) {
      float sg = 2.0f * (dot(hit.normal, a1[i]) > 0.0f ? 1.0f : 0.0f) - 1.0f;
      pa = pa + (sg * comp(o1.extent, i)) * a1[i];
    }
    V3 pb = o2.center;
    for (int i = 0; i < 3; ++i) {
      float sg = 1.0f - 2.0f * (dot(hit.normal, a2[i]) > 0.0f ? 1.0f : 0.0f);
      pb = pb + (sg * comp(o2.extent, i)) * a2[i];
    }
    V3 ua = a1[(hit.code - 7) / 3], ub = a2[(hit.code - 7) % 3];
    float beta = line_closest_approach(pa, ua, pb, ub);
    add_contact(c, m, b1, b2, -hit.normal, pb + ub * beta, fneg(hit.depth), -1);
    return;
  }
  if (hit.code > 3) {  // a face of box 2: swap the roles
    Obb to = o1;
    o1 = o2;
    o2 = to;
    M3 tm = r1t;
    r1t = r2t;
    r2t = tm;
    for (int i = 0; i < 3; ++i) {
      V3 t = a1[i];
      a1[i] = a2[i];
      a2[i] = t;
    }
  }
  V3 n2 = hit.code <= 3 ? hit.normal : -hit.normal;
  V3 nr = mul_t(r2t, n2);
  V3 anr = vabs(nr);
  int lanr, ax1, ax2;
  if (anr.y > anr.x) {
    if (anr.y > anr.z) { lanr = 1; ax1 = 0; ax2 = 2; }
    else { lanr = 2; ax1 = 0; ax2 = 1; }
  } else if (anr.x > anr.z) {
    lanr = 0; ax1 = 1; ax2 = 2;
  } else {
    lanr = 2; ax1 = 0; ax2 = 1;
  }
  float inc = comp(nr, lanr) < 0.0f ? comp(o2.extent, lanr) : fneg(comp(o2.extent, lanr));
  V3 center = (o2.center - o1.center) + inc * a2[lanr];
  uint32_t cn = hit.code <= 3 ? hit.code - 1 : hit.code - 4;
  int c1i = cn == 0 ? 1 : 0, c2i = cn == 2 ? 1 : 2;
  float c1 = dot(center, a1[c1i]), c2 = dot(center, a1[c2i]);
  V4 mm = V4{dot(a1[c1i], a2[ax1]), dot(a1[c1i], a2[ax2]), dot(a1[c2i], a2[ax1]), dot(a1[c2i], a2[ax2])};
  V4 kk = mm * V4{comp(o2.extent, ax1), comp(o2.extent, ax2), comp(o2.extent, ax1), comp(o2.extent, ax2)};
  float quad[4][2] = {{(c1 - kk.x) - kk.y, (c2 - kk.z) - kk.w},
                      {(c1 - kk.x) + kk.y, (c2 - kk.z) + kk.w},
                      {(c1 + kk.x) + kk.y, (c2 + kk.z) + kk.w},
                      {(c1 + kk.x) - kk.y, (c2 + kk.z) - kk.w}};
  float rect[2] = {comp(o1.extent, c1i), comp(o1.extent, c2i)};
  Poly16 ret = clip_rect_quad(rect, quad);
  if (ret.n == 0) return;
  float det1 = 1.0f / (mm.x * mm.w - mm.y * mm.z);
  mm = mm * V4{det1, det1, det1, det1};
  uint32_t cnum = 0;
  V3 point[16];
  float dep[16];
  for (uint32_t j = 0; j < ret.n; ++j) {
    float k1 = mm.w * (ret.p[j][0] - c1) - mm.y * (ret.p[j][1] - c2);
    float k2 = fneg(mm.z) * (ret.p[j][0] - c1) + mm.x * (ret.p[j][1] - c2);
    point[cnum] = (center + k1 * a2[ax1]) + k2 * a2[ax2];
    dep[cnum] = comp(o1.extent, (int)cn) - dot(n2, point[cnum]);
    if (dep[cnum] >= 0.0f) {
      ret.p[cnum][0] = ret.p[j][0];
      ret.p[cnum][1] = ret.p[j][1];
      ++cnum;
    }
  }
  if (cnum == 0) return;
  uint32_t maxc = cnum < (uint32_t)MAX_POINTS ? cnum : (uint32_t)MAX_POINTS;
  PointKeys pk = point_keys(c, m);
  if (cnum <= maxc) {
    for (uint32_t i = 0; i < cnum; ++i) {
      V3 pw = point[i] + o1.center;
      add_contact(c, m, pk, b1, b2, -hit.normal, hit.code < 4 ? pw : pw - hit.normal * dep[i], fneg(dep[i]), -1);
    }
    return;
  }
  uint32_t i1 = 0;
  for (uint32_t i = 0; i < cnum; ++i)
    if (dep[i] > dep[i1]) i1 = i;
  uint32_t keep[MAX_POINTS];
  cull_points(ret.p, cnum, i1, maxc, keep);
  for (uint32_t t = 0; t < maxc; ++t) {
    uint32_t i = keep[t];
    V3 pw = point[i] + o1.center;
    add_contact(c, m, pk, b1, b2, -hit.normal, hit.code < 4 ? pw : pw - hit.normal * dep[i], fneg(dep[i]), -1);
  }
}
RS_FN bool alg_box_box(Np& c, uint32_t ba, uint32_t bb, Manifold* out) {  // :379-409
  if (!overlaps(body_box(c, ba), body_box(c, bb))) return false;
  Aff ta = child_world(c, ba), tb = child_world(c, bb);
  const BodyConst &c1 = c.w.bodies[ba], &c2 = c.w.bodies[bb];
  Obb o1 = Obb{ta.t, ta.m, add_s(c1.box_half_extents, c1.box_margin)};
  Obb o2 = Obb{tb.t, tb.m, add_s(c2.box_half_extents, c2.box_margin)};
  BoxHit hit;
  if (!box_box_sat(o1, transpose(ta.m), o2, &hit)) return false;
  Manifold& m = *out;
  manifold_init(m, ba, bb);
  box_box_contacts(c, o1, ta.m, o2, tb.m, hit, m, ba, bb);
  if (m.n_points == 0) return false;
  refresh_points(c, m, ba, bb);
  return true;
}

// ---- sphere-OBB (sphere_obb_collision_alg.cuh) -----------------------------------------------------------------------------------
RS_FN bool alg_sphere_obb(Np& c, uint32_t sphere, uint32_t box, Manifold* out) {
  Aff st = body_trans(c, sphere);
  float radius = c.w.bodies[sphere].sphere_margin;
  if (!overlaps(sphere_box(radius, st.t), body_box(c, box))) return false;
  Aff ct = child_world(c, box);
  const BodyConst& bc = c.w.bodies[box];
  V3 ext = bc.box_half_extents;
  V3 local = inv_xform(ct, st.t);
  V3 closest = vclamp(local, -ext, ext);
  V3 delta = local - closest;
  float d2 = length_sq(delta);
  Manifold& m = *out;
  manifold_init(m, sphere, box);
  float isect = radius + bc.box_margin;
  float cdist = isect + breaking_threshold(c, m);
  if (d2 > cdist * cdist) return false;
  float dist;
  if (d2 > k::F32_EPS) {
    dist = __builtin_sqrtf(d2);
    delta = delta / dist;
  } else {  // the centre inside the box: the nearest face (get_sphere_penetration, :16-36)
    float md = ext.x - local.x;
    V3 cl = local;
    cl.x = ext.x;
    V3 nn = v3(1.0f, 0.0f, 0.0f);
    const int axes[5] = {0, 1, 1, 2, 2};
    const float sgn[5] = {-1.0f, 1.0f, -1.0f, 1.0f, -1.0f};
    for (int t = 0; t < 5; ++t) {
      int a = axes[t];
      float fd = sgn[t] < 0.0f ? comp(ext, a) + comp(local, a) : comp(ext, a) - comp(local, a);
      if (fd < md) {
        md = fd;
        cl = local;
        set_comp(cl, a, comp(ext, a) * sgn[t]);
        nn = splat(0.0f);
        set_comp(nn, a, sgn[t]);
      }
    }
    dist = md * -1.0f;
    closest = cl;
    delta = nn;
  }
  V3 n_box = ct.m * delta;
  V3 p_box = xform_point(ct, closest);
  add_contact(c, m, sphere, box, n_box, p_box + n_box * bc.box_margin, dist - isect, -1);
  refresh_points(c, m, sphere, box);
  return m.n_points != 0;
}

// ---- convex-plane (convex_plane_collision_alg.cuh) ------------------------------------------------------------------------------
RS_FN bool convex_plane(Np& c, uint32_t obj, uint32_t plane_obj, const Aff& wt, bool box_child, Manifold* out) {
  const Plane& pl = c.w.planes[plane_obj - NUM_MESHES];
  const BodyConst& bc = c.w.bodies[obj];
  Aabb cb = box_child ? box_from_half_extents(bc.box_half_extents, bc.box_margin, wt) : sphere_box(bc.sphere_margin, wt.t);
  if (!overlaps(cb, pl.aabb)) return false;
  V3 pn = pl.normal;
  Aff pt = body_trans(c, plane_obj);
  M3 plane_in_convex = transpose(wt.m);
  Aff convex_in_plane = Aff{wt.m, wt.t - pt.t};
  V3 dir = plane_in_convex * (-pn);
  V3 vtx;
  if (box_child) {
    V3 sg = vselect(cmplt(dir, splat(0.0f)), splat(-1.0f), splat(1.0f));
    vtx = add_s(bc.box_half_extents, bc.box_margin) * sg;
  } else {
    V3 n = splat(0.0f);
    try_normalize(dir, &n);
    vtx = splat(bc.sphere_margin) * n;
  }
  V3 vip = xform_point(convex_in_plane, vtx);
  float distance = dot(pn, vip);
  Manifold& m = *out;
  manifold_init(m, obj, plane_obj);
  if (distance >= breaking_threshold(c, m)) return false;
  V3 proj = vip - splat(distance) * pn;
  add_contact(c, m, obj, plane_obj, pn, proj + pt.t, distance, -1);
  refresh_points(c, m, obj, plane_obj);
  return true;
}

RS_FN const MeshInfo* mesh_of(const StaticData& w, uint32_t body) {
  return (body < (uint32_t)NUM_MESHES && w.meshes[body].body == body) ? &w.meshes[body] : nullptr;
}

// ---- the ball against an arena mesh (sphere_concave_collision_alg.cuh, triangle_shape.cuh) ------------------------------------------
RS_FN float seg_dist2(V3 from, V3 to, V3 p, V3* nearest) {  // triangle_shape.cuh:11-31
  V3 diff = p - from, v = to - from;
  float t = dot(v, diff);
  if (t > 0.0f) {
    float vv = dot(v, v);
    if (t < vv) {
      t = t / vv;
      diff -= splat(t) * v;
    } else {
      t = 1.0f;
      diff -= v;
    }
  } else {
    t = 0.0f;
  }
  *nearest = from + splat(t) * v;
  return dot(diff, diff);
}
// intersect_sphere_front_precomputed + intersect_sphere_from_plane (triangle_shape.cuh:65-117)
RS_FN bool sphere_triangle(const Tri& t, V3 ctr, float radius, float rwt, float rwt2, V3* normal, V3* point, float* depth) {
  V3 to_center = ctr - t.p0;
  float dfp = dot(to_center, t.normal);
  if (dfp < 0.0f) return false;
  if (dfp >= rwt) return false;
  V3 otp[3] = {to_center, ctr - t.p1, ctr - t.p2};
  V3 cp;
  bool inside = !(dot(cross(tri_edge(t, 0), otp[0]), t.normal) < 0.0f) && !(dot(cross(tri_edge(t, 1), otp[1]), t.normal) < 0.0f) &&
                dot(cross(tri_edge(t, 2), otp[2]), t.normal) >= 0.0f;  // face_contains (:45-54)
  if (inside) {
    cp = ctr - t.normal * dfp;
  } else {
    float best = rwt2;
    cp = splat(0.0f);
    for (int e = 0; e < 3; ++e) {
      V3 from = e == 0 ? t.p0 : e == 1 ? t.p1 : t.p2, to = e == 0 ? t.p1 : e == 1 ? t.p2 : t.p0, ne = splat(0.0f);
      float d2 = seg_dist2(from, to, ctr, &ne);
      if (d2 < best) {
        best = d2;
        cp = ne;
      }
    }
    if (!(best < rwt2)) return false;
  }
  V3 ctc = ctr - cp;
  float d2 = length_sq(ctc);
  if (d2 >= rwt2) return false;
  if (d2 > k::F32_EPS) {
    float d = __builtin_sqrtf(d2);
    *normal = ctc * (1.0f / d);
    *depth = fneg(radius - d);
  } else {
    *normal = t.normal;
    *depth = fneg(radius);
  }
  *point = cp;
  return true;
}
// ---- the tick's mesh-job BVH walks, ahead of the pair loop (rs_trav.h) -----------------------------------------------------
// The job record of (kind, body, mesh body), searched from the last one taken (jobs are taken in collection order).
RS_FN const TravJob* trav_find(const Np& c, uint32_t kind, uint32_t body, uint32_t mesh_body) {
  if (!c.tb) return nullptr;
  for (uint32_t k = c.tb->next; k < c.tb->n_jobs; ++k) {
    const TravJob& jb = c.tb->job[k];
    if (jb.kind == kind && jb.body == body && jb.mesh_body == mesh_body) {
      c.tb->next = k + 1;
      return &jb;
    }
  }
  return nullptr;
}

// sphere_concave_collision_alg.rs process_collision_into: every BVH leaf near the ball, in traversal order, adds its contact to
// the pair's retained manifold; then the manifold is refreshed.
RS_FN void sphere_mesh(Np& c, uint32_t ball, uint32_t mesh_body, Manifold& m) {
  Aff bt = body_trans(c, ball), mt = body_trans(c, mesh_body);
  const MeshInfo* mesh = mesh_of(c.w, mesh_body);
  if (mesh == nullptr) {
    c.unimplemented += 1;
    return;
  }
  V3 ctr = bt.t - mt.t;
  float radius = c.w.bodies[ball].sphere_margin;
  float rwt = radius + breaking_threshold(c, m), rwt2 = rwt * rwt;
  Aabb box = sphere_box(radius, ctr);
  PointKeys pk;  // read at the first contact
  bool keyed = false;
  const TravJob* jb = trav_find(c, TJ_BALL, ball, mesh_body);
  if (jb && !jb->in_place) {  // the walk ran ahead: its leaves in depth-first order
    // the next leaf's triangle is loaded before this one's test and contact (loads only: the same sequence of tests)
    const uint32_t* leaves = c.tb->leaf_tri + jb->first;
    const uint32_t count = jb->count;
    uint32_t tri_next = 0u;
    Tri t_next = {};
    if (count) {
      tri_next = leaves[0];
      t_next = c.w.tris[mesh->tri_first + tri_next];
    }
    for (uint32_t k = 0; k < count; ++k) {
      const uint32_t tri = tri_next;
      const Tri t = t_next;
      if (k + 1 < count) {
        tri_next = leaves[k + 1];
        t_next = c.w.tris[mesh->tri_first + tri_next];
      }
      V3 n, p;
      float depth;
      if (!sphere_triangle(t, ctr, radius, rwt, rwt2, &n, &p, &depth)) continue;
      if (!keyed) pk = point_keys(c, m);
      keyed = true;
      add_contact(c, m, pk, ball, mesh_body, n, p + mt.t, depth, (int32_t)tri);
    }
  } else {
    BvhCursor cur;
    cursor_init(cur, *mesh, box);
    uint32_t tri;
    while (cursor_next(cur, c.w, *mesh, box, &tri)) {
      V3 n, p;
      float depth;
      if (!sphere_triangle(c.w.tris[mesh->tri_first + tri], ctr, radius, rwt, rwt2, &n, &p, &depth)) continue;
      if (!keyed) pk = point_keys(c, m);
      keyed = true;
      add_contact(c, m, pk, ball, mesh_body, n, p + mt.t, depth, (int32_t)tri);
    }
  }
  if (m.n_points != 0) refresh_points(c, m, ball, mesh_body);
}
// The ball job's query box, as sphere_mesh computes it.
RS_FN Aabb ball_mesh_box(const Np& c, uint32_t ball, uint32_t mesh_body) {
  return sphere_box(c.w.bodies[ball].sphere_margin, body_trans(c, ball).t - body_trans(c, mesh_body).t);
}

// ---- a car against an arena mesh (compound_collision_alg.cuh:63-160) --------------------------------------------------------------
// The job's frames: the car box in world space, the mesh body's transform, the box's AABB in mesh space, the mesh-to-box
// transform the triangles are tested in, and the contact distance limit.
struct CarMeshFrame {
  Aff box_t, mt, mesh_to_box;
  Aabb local_box;
  float maxd;
};
RS_FN CarMeshFrame car_mesh_frame(const Np& c, uint32_t car, uint32_t mesh_body) {
  const BodyConst& bc = c.w.bodies[car];
  CarMeshFrame f;
  f.box_t = child_world(c, car);
  f.mt = body_trans(c, mesh_body);
  Aff x1 = aff_transpose(f.mt);
  Aff in_tri = Aff{x1.m * f.box_t.m, xform_point(x1, f.box_t.t)};
  f.local_box = box_from_half_extents(bc.box_half_extents, bc.box_margin, in_tri);
  f.mesh_to_box = aff_transpose(f.box_t) * f.mt;
  float breaking = fmin_rs(bc.contact_breaking_threshold, c.w.bodies[mesh_body].contact_breaking_threshold);
  f.maxd = bc.box_margin + breaking;
  return f;
}
RS_FN bool tri_in_box(const Tri& t, const Aabb& b) {
  V3 tmin = vmin(vmin(t.p0, t.p1), t.p2), tmax = vmax(vmax(t.p0, t.p1), t.p2);
  return all3(cmple(tmin, b.max)) && all3(cmpge(tmax, b.min));
}

// The tick's car-mesh separating-axis tests, run ahead of the pair loop. P3b first walks the pairs that will reach
// car_mesh (the same gate as near_callback) and collects each job's candidate triangles in BVH order; the tests then run
// over all collected triangles at once (on the GPU spread over every active lane of the wave, so a lane with 30
// triangles no longer holds the other 63 for 30 tests); car_mesh then adds the contacts in the original order. Results
// are the same: a test depends only on its box, triangle and limits, and nothing in P3b changes those. A job that does
// not fit (more than SB_JOBS jobs or SB_TASKS triangles in the tick) is left to car_mesh's own loop.
constexpr int SB_JOBS = 4, SB_TASKS = 32;
struct SatJob {
  uint32_t car, mesh_body;
  uint32_t start, count;
  bool in_place;  // the triangles did not fit: car_mesh walks and tests them itself
  Aff box_t, mesh_to_box;
  V3 half;
  float margin, maxd;
  uint32_t tri_first;
};
struct SatBatch {
  uint32_t n_jobs, n_tasks, next_job, ok;  // ok: bit t set when task t's test found a contact
  SatJob jobs[SB_JOBS];
  uint32_t tri[SB_TASKS];
  uint8_t job_of[SB_TASKS];
  SatContact res[SB_TASKS];
};
constexpr int SB_RECORD = 26;  // a staged test: box transform (12), half extents (3), margin, limit, triangle (9)

RS_FN void sat_collect(const Np& c, const TravJob& tj, SatBatch& sb) {
  if (sb.n_jobs >= (uint32_t)SB_JOBS || tj.in_place) return;  // the job then walks and tests in car_mesh
  const uint32_t car = tj.body, mesh_body = tj.mesh_body;
  const MeshInfo* mesh = mesh_of(c.w, mesh_body);
  if (mesh == nullptr) return;
  const BodyConst& bc = c.w.bodies[car];
  CarMeshFrame f = car_mesh_frame(c, car, mesh_body);
  SatJob& j = sb.jobs[sb.n_jobs];
  j.car = car;
  j.mesh_body = mesh_body;
  j.start = sb.n_tasks;
  j.in_place = false;
  j.box_t = f.box_t;
  j.mesh_to_box = f.mesh_to_box;
  j.half = bc.box_half_extents;
  j.margin = bc.box_margin;
  j.maxd = f.maxd;
  j.tri_first = mesh->tri_first;
  const Tri* tris = c.w.tris + mesh->tri_first;
  for (uint32_t k = 0; k < tj.count; ++k) {  // the walk's leaves in depth-first order, as car_mesh visits them
    const uint32_t tri = c.tb->leaf_tri[tj.first + k];
    if (!tri_in_box(tris[tri], f.local_box)) continue;
    if (sb.n_tasks >= (uint32_t)SB_TASKS) {
      j.in_place = true;
      sb.n_tasks = j.start;
      break;
    }
    sb.tri[sb.n_tasks] = tri;
    sb.job_of[sb.n_tasks] = (uint8_t)sb.n_jobs;
    ++sb.n_tasks;
  }
  j.count = sb.n_tasks - j.start;
  ++sb.n_jobs;
}
// The staged record of task t (its triangle transformed into the box frame, as car_mesh does).
RS_FN void sat_record(const Np& c, const SatBatch& sb, uint32_t t, float* r) {
  const SatJob& j = sb.jobs[sb.job_of[t]];
  const Tri& tr = c.w.tris[j.tri_first + sb.tri[t]];
  V3 q[3] = {xform_point(j.mesh_to_box, tr.p0), xform_point(j.mesh_to_box, tr.p1), xform_point(j.mesh_to_box, tr.p2)};
  const V3 v[8] = {j.box_t.m.c0, j.box_t.m.c1, j.box_t.m.c2, j.box_t.t, j.half, q[0], q[1], q[2]};
  for (int i = 0; i < 5; ++i) {
    r[3 * i] = v[i].x;
    r[3 * i + 1] = v[i].y;
    r[3 * i + 2] = v[i].z;
  }
  r[15] = j.margin;
  r[16] = j.maxd;
  for (int i = 0; i < 3; ++i) {
    r[17 + 3 * i] = v[5 + i].x;
    r[18 + 3 * i] = v[5 + i].y;
    r[19 + 3 * i] = v[5 + i].z;
  }
}
RS_FN bool sat_run_record(const float* r, SatContact* out) {
  Aff box_t = Aff{M3{v3(r[0], r[1], r[2]), v3(r[3], r[4], r[5]), v3(r[6], r[7], r[8])}, v3(r[9], r[10], r[11])};
  V3 half = v3(r[12], r[13], r[14]);
  V3 q[3] = {v3(r[17], r[18], r[19]), v3(r[20], r[21], r[22]), v3(r[23], r[24], r[25])};
  return box_triangle_sat(box_t, half, r[15], q, r[16], out);
}
// All active lanes of the wave take the collected tests of all of them in turns of one test each (wave_pool); on the
// host, one lane in order.
RS_FN void sat_batch_run(const Np& c, SatBatch& sb) {
  sb.ok = 0;
  wave_pool<SB_RECORD, 5, SB_TASKS>(
      sb.n_tasks, [&](uint32_t t, float* r) { sat_record(c, sb, t, r); },
      [](const float* r, float* o) {
        SatContact ct = SatContact{};
        o[0] = sat_run_record(r, &ct) ? 1.0f : 0.0f;
        o[1] = ct.point_world.x, o[2] = ct.point_world.y, o[3] = ct.point_world.z, o[4] = ct.distance;
      },
      [&](uint32_t t, const float* o) {
        if (o[0] != 0.0f) sb.ok |= 1u << t;
        sb.res[t] = SatContact{v3(o[1], o[2], o[3]), o[4]};
      });
}

RS_FN bool car_mesh(Np& c, uint32_t car, uint32_t mesh_body, Manifold* out) {
  const MeshInfo* mesh = mesh_of(c.w, mesh_body);
  if (mesh == nullptr) {
    c.unimplemented += 1;
    return false;
  }
  const BodyConst& bc = c.w.bodies[car];
  CarMeshFrame f = car_mesh_frame(c, car, mesh_body);
  const Tri* tris = c.w.tris + mesh->tri_first;
  Manifold& m = *out;
  manifold_init(m, car, mesh_body);
  PointKeys pk = point_keys(c, m);
  bool has_m = false;
  const SatJob* job = nullptr;
  if (c.sb)
    for (uint32_t k = c.sb->next_job; k < c.sb->n_jobs; ++k)
      if (c.sb->jobs[k].car == car && c.sb->jobs[k].mesh_body == mesh_body) {
        job = &c.sb->jobs[k];
        c.sb->next_job = k + 1;
        break;
      }
  if (job && !job->in_place) {  // the tests ran ahead: add the contacts in BVH order
    for (uint32_t t = job->start; t < job->start + job->count; ++t) {
      if (!((c.sb->ok >> t) & 1u)) continue;
      const uint32_t tri = c.sb->tri[t];
      has_m = true;  // the manifold exists from the first contact on
      add_contact(c, m, pk, car, mesh_body, f.mt.m * tris[tri].normal, c.sb->res[t].point_world, c.sb->res[t].distance, (int32_t)tri);
    }
  } else {
    BvhCursor cur;
    cursor_init(cur, *mesh, f.local_box);
    uint32_t tri;
    while (cursor_next(cur, c.w, *mesh, f.local_box, &tri)) {
      const Tri& t = tris[tri];
      if (!tri_in_box(t, f.local_box)) continue;
      V3 q[3] = {xform_point(f.mesh_to_box, t.p0), xform_point(f.mesh_to_box, t.p1), xform_point(f.mesh_to_box, t.p2)};
      SatContact ct;
      if (!box_triangle_sat(f.box_t, bc.box_half_extents, bc.box_margin, q, f.maxd, &ct)) continue;
      has_m = true;
      add_contact(c, m, pk, car, mesh_body, f.mt.m * t.normal, ct.point_world, ct.distance, (int32_t)tri);
    }
  }
  if (!has_m) return false;
  refresh_points(c, m, car, mesh_body);
  return m.n_points != 0;
}

// ---- the dispatcher (collision_dispatcher.cuh) -----------------------------------------------------------------------------------
// Appends a fresh manifold to the pool: its header and points, the unused point slots written as zeros (the whole
// manifold, as a copy of a zero-filled one would write it).
RS_FN int insert_manifold(ArenaState& s, const Manifold& m) {
  if (s.n_manifolds >= (uint32_t)MAX_MANIFOLDS) {
    s.overflow[OVF_MANIFOLDS] += 1;
    return -1;
  }
  Manifold& d = s.manifolds[s.n_manifolds];
  d.body0 = m.body0;
  d.body1 = m.body1;
  d.n_points = m.n_points;
  for (uint32_t i = 0; i < (uint32_t)MAX_POINTS; ++i) {
    if (i < m.n_points) d.pts[i] = m.pts[i];
    else d.pts[i] = ManifoldPoint{};
  }
  return (int)s.n_manifolds++;
}
// Appends an empty manifold of bodies b0, b1 (all points zero).
RS_FN int insert_manifold_empty(ArenaState& s, uint32_t b0, uint32_t b1) {
  if (s.n_manifolds >= (uint32_t)MAX_MANIFOLDS) {
    s.overflow[OVF_MANIFOLDS] += 1;
    return -1;
  }
  Manifold& d = s.manifolds[s.n_manifolds];
  d.body0 = b0;
  d.body1 = b1;
  d.n_points = 0;
  for (uint32_t i = 0; i < (uint32_t)MAX_POINTS; ++i) d.pts[i] = ManifoldPoint{};
  return (int)s.n_manifolds++;
}
RS_FN void push_active(ArenaState& s, uint32_t idx) {
  if (s.n_active >= (uint32_t)MAX_ACTIVE) {
    s.overflow[OVF_ACTIVE] += 1;
    return;
  }
  s.active[s.n_active++] = idx;
}
// the body order of the empty manifold a gated-in pair with no contact keeps (:80-93)
RS_FN void empty_manifold_order(const Np& c, uint32_t a, uint32_t b, uint32_t* f, uint32_t* g) {
  uint32_t ka = c.w.bodies[a].shape_kind, kb = c.w.bodies[b].shape_kind;
  bool sw = ka == k::SHAPE_PLANE || (ka == k::SHAPE_MESH && (kb == k::SHAPE_SPHERE || kb == k::SHAPE_COMPOUND)) ||
            (ka == k::SHAPE_COMPOUND && kb == k::SHAPE_SPHERE);
  *f = sw ? b : a;
  *g = sw ? a : b;
}
// the pairs that need no mesh job (:97-105)
RS_FN bool process_inner(Np& c, uint32_t a, uint32_t b, Manifold* out) {
  uint32_t ka = c.w.bodies[a].shape_kind, kb = c.w.bodies[b].shape_kind;
  if (ka == k::SHAPE_PLANE && kb == k::SHAPE_SPHERE) return convex_plane(c, b, a, body_trans(c, b), false, out);
  if (ka == k::SHAPE_SPHERE && kb == k::SHAPE_PLANE) return convex_plane(c, a, b, body_trans(c, a), false, out);
  if (ka == k::SHAPE_SPHERE && kb == k::SHAPE_COMPOUND) return alg_sphere_obb(c, a, b, out);
  if (ka == k::SHAPE_COMPOUND && kb == k::SHAPE_SPHERE) return alg_sphere_obb(c, b, a, out);
  if (ka == k::SHAPE_COMPOUND && kb == k::SHAPE_COMPOUND) return alg_box_box(c, a, b, out);
  c.unimplemented += 1;
  return false;
}
// the end of the near callback (:122-156): merge into the retained manifold or insert the fresh one, then activate it
RS_FN void near_tail(Np& c, uint32_t a, uint32_t b, int cached, bool has_fresh, const Manifold& fresh) {
  ArenaState& s = c.s;
  int act;
  if (cached >= 0) {
    Manifold& m = s.manifolds[cached];
    uint32_t b0 = m.body0, b1 = m.body1;
    if (m.n_points != 0) refresh_points(c, m, b0, b1);
    if (has_fresh) {
      merge_points(c, m, fresh);
    } else {
      for (uint32_t i = 0; i < m.n_points; ++i) {
        ManifoldPoint p = m.pts[i];
        contact_processed(c, p, b0, b1);
      }
    }
    act = cached;
  } else if (has_fresh) {
    act = insert_manifold(s, fresh);
  } else {
    uint32_t f, g;
    empty_manifold_order(c, a, b, &f, &g);
    act = insert_manifold_empty(s, f, g);
  }
  if (act >= 0) push_active(s, (uint32_t)act);
}
// the near callback (:160-217 with the mesh jobs run in place)
RS_FN void near_callback(Np& c, uint32_t a, uint32_t b) {
  ArenaState& s = c.s;
  if ((!body_awake(c, a) && !body_awake(c, b)) || !body_responds(c, a) || !body_responds(c, b)) return;
  int cached = find_manifold(s, a, b);
  uint32_t ka = c.w.bodies[a].shape_kind, kb = c.w.bodies[b].shape_kind;
  bool compound_pair = (ka == k::SHAPE_COMPOUND) != (kb == k::SHAPE_COMPOUND);
  uint32_t comp_b = ka == k::SHAPE_COMPOUND ? a : b, other = ka == k::SHAPE_COMPOUND ? b : a;
  if (compound_pair && !overlaps(child_world_box(c, comp_b), body_box(c, other))) {
    release_manifold(s, a, b);
    return;
  }
  if ((ka == k::SHAPE_SPHERE && kb == k::SHAPE_MESH) || (ka == k::SHAPE_MESH && kb == k::SHAPE_SPHERE)) {
    uint32_t ball = ka == k::SHAPE_SPHERE ? a : b, mesh_b = ka == k::SHAPE_SPHERE ? b : a;
    int idx = cached >= 0 ? cached : insert_manifold_empty(s, ball, mesh_b);
    if (idx < 0) return;
    sphere_mesh(c, ball, mesh_b, s.manifolds[idx]);
    push_active(s, (uint32_t)idx);
    return;
  }
  Manifold fresh;  // built in place by the algorithm that reports a contact (manifold_init); read only then
  fresh.n_points = 0;
  bool has_fresh = false;
  uint32_t ko = c.w.bodies[other].shape_kind;
  if (compound_pair && ko == k::SHAPE_PLANE) has_fresh = convex_plane(c, comp_b, other, child_world(c, comp_b), true, &fresh);
  else if (compound_pair && ko == k::SHAPE_MESH) has_fresh = car_mesh(c, comp_b, other, &fresh);
  else has_fresh = process_inner(c, a, b, &fresh);
  near_tail(c, a, b, cached, has_fresh, fresh);
}
// Whether near_callback(a, b) reaches car_mesh or sphere_mesh, and with which bodies: its checks before those calls, none
// of which depends on anything P3b writes (activation, collision flags, transforms, shapes).
RS_FN int mesh_job_of(const Np& c, uint32_t a, uint32_t b, uint32_t* body, uint32_t* mesh_b) {
  if ((!body_awake(c, a) && !body_awake(c, b)) || !body_responds(c, a) || !body_responds(c, b)) return -1;
  uint32_t ka = c.w.bodies[a].shape_kind, kb = c.w.bodies[b].shape_kind;
  if ((ka == k::SHAPE_SPHERE && kb == k::SHAPE_MESH) || (ka == k::SHAPE_MESH && kb == k::SHAPE_SPHERE)) {
    *body = ka == k::SHAPE_SPHERE ? a : b;
    *mesh_b = ka == k::SHAPE_SPHERE ? b : a;
    return mesh_of(c.w, *mesh_b) ? TJ_BALL : -1;
  }
  if ((ka == k::SHAPE_COMPOUND) == (kb == k::SHAPE_COMPOUND)) return -1;
  uint32_t comp_b = ka == k::SHAPE_COMPOUND ? a : b, other = ka == k::SHAPE_COMPOUND ? b : a;
  if (c.w.bodies[other].shape_kind != k::SHAPE_MESH) return -1;
  if (!overlaps(child_world_box(c, comp_b), body_box(c, other))) return -1;
  *body = comp_b;
  *mesh_b = other;
  return mesh_of(c.w, other) ? TJ_CAR : -1;
}
// P3b (:230-282): the retained pairs in pair-array order. Returns the number of unimplemented pair algorithms reached.
RS_INLINE uint32_t phase_p3b_body(ArenaState& s, const StaticData& w) {
  // the records' event bits and count are kept in registers and written once (nothing else reads them before P7)
  const uint32_t n_records0 = s.n_records;
  Np c = {s, w, 0, nullptr, nullptr, 0u, n_records0};
  uint32_t n = s.n_pairs < (uint32_t)MAX_PAIRS ? s.n_pairs : (uint32_t)MAX_PAIRS;
  TravBatch tb;
  tb.n_jobs = tb.n_leaves = tb.next = tb.leaves = 0;
  for (uint32_t i = 0; i < n && tb.n_jobs < (uint32_t)TJ_MAX; ++i) {  // the mesh jobs, in pair order
    Pair p = s.pairs[i];
    uint32_t body, mesh_b;
    int kind = mesh_job_of(c, w.proxies[p.proxy0].client_body, w.proxies[p.proxy1].client_body, &body, &mesh_b);
    if (kind < 0) continue;
    TravJob& jb = tb.job[tb.n_jobs++];
    jb.kind = (uint32_t)kind;
    jb.body = body;
    jb.mesh_body = mesh_b;
    jb.box = kind == TJ_CAR ? car_mesh_frame(c, body, mesh_b).local_box : ball_mesh_box(c, body, mesh_b);
  }
  trav_run(w, tb);
  c.tb = &tb;
  SatBatch sb;
  sb.n_jobs = sb.n_tasks = sb.next_job = sb.ok = 0;
  for (uint32_t j = 0; j < tb.n_jobs; ++j)
    if (tb.job[j].kind == TJ_CAR) sat_collect(c, tb.job[j], sb);
  sat_batch_run(c, sb);
  c.sb = &sb;
  for (uint32_t i = 0; i < n; ++i) {
    Pair p = s.pairs[i];
    near_callback(c, w.proxies[p.proxy0].client_body, w.proxies[p.proxy1].client_body);
  }
  if (c.fold != 0) s.fold_events |= c.fold;
  if (c.n_records != n_records0) s.n_records = c.n_records;
  return c.unimplemented;
}
// The phase as a function of its own whose state and world pointers carry their address spaces: its accesses compile to
// scratch_ and global_ instructions (through generic pointers they would be flat ones, which wait on both the vector-memory
// and the LDS counters and need 64-bit addresses).
// UW: every lane steps in the same world (its pointer is then made wave-uniform).
template <bool UW>
RS_NOINL uint32_t phase_p3b_in(RS_PRIVATE_AS ArenaState* sp, const RS_CONST_AS StaticData* wp) {
  if (UW) wp = rs_uniform_ptr(wp);
  return phase_p3b_body(*(ArenaState*)sp, *(const StaticData*)wp);
}
template <bool UW = false>
RS_FN uint32_t phase_p3b(ArenaState& s, const StaticData& w) {
  return phase_p3b_in<UW>((RS_PRIVATE_AS ArenaState*)&s, (const RS_CONST_AS StaticData*)&w);
}

}  // namespace rs
