// The car box against one arena triangle, in the box's frame (V3 box_triangle_sat.rs as restated in
// cuda/sim/box_triangle_sat.cuh): a pre-screen, a separating-axis sweep over 3 box faces, the triangle face and 9 edge pairs,
// then a contact point from 47 candidate point pairs (separated) or from the penetration axis's witness. The axis order and
// the tie rules decide which contact is made, so they follow the reference.
#pragma once
#include "rs_contact.h"

namespace rs {

constexpr float SAT_TIE_EPS = 1e-4f;
constexpr int SAT_NPAIRS = 47;

struct SatContact { V3 point_world; float distance; };
struct SatPair { V3 pbox, ptri; float d2; };
enum { SAT_NONE = 0, SAT_TRI_FACE = 1, SAT_BOX_FACE = 2, SAT_EDGE_EDGE = 3 };
struct SatKind { int tag, box_axis, tri_edge; };
struct SatSweep {
  bool has_sep;
  float best_gap;
  V3 best_sep_axis;
  float best_overlap;
  V3 best_pen_axis, best_pen_raw;
  SatKind pen_kind;
};

RS_FN V3 unit_axis(int i) { return i == 0 ? v3(1.0f, 0.0f, 0.0f) : i == 1 ? v3(0.0f, 1.0f, 0.0f) : v3(0.0f, 0.0f, 1.0f); }

// :331 orient_pen_axis
RS_FN V3 orient_pen(V3 a, V3 to_box, V3 face, bool has_face) {
  float s = dot(a, to_box);
  if (s < 0.0f) return -a;
  if (s > 1e-12f) return a;
  if (has_face && dot(a, face) > 0.0f) return -a;
  return a;
}
// :525 axis_gap_overlap: the box's radius on axis a against the triangle's projection
RS_FN void gap_overlap(V3 a, const V3 q[3], V3 half, float* gap, float* overlap, float* side) {
  RS_PRIVATE(q); RS_PRIVATE(gap); RS_PRIVATE(overlap); RS_PRIVATE(side);  // locals of the caller
  float r = (half.x * fabs_(a.x) + half.y * fabs_(a.y)) + half.z * fabs_(a.z);
  float d0 = dot(q[0], a), d1 = dot(q[1], a), d2 = dot(q[2], a);
  float mn = fmin_rs(fmin_rs(d0, d1), d2), mx = fmax_rs(fmax_rs(d0, d1), d2);
  if (mn > r) {
    *gap = mn - r; *overlap = 0.0f; *side = 1.0f;
  } else if (mx < fneg(r)) {
    *gap = fneg(r) - mx; *overlap = 0.0f; *side = -1.0f;
  } else {
    *gap = 0.0f; *overlap = fmax_rs(fmin_rs(mx + r, r - mn), 0.0f); *side = 0.0f;
  }
}
// :341 consider_axis: true when the axis rejects the pair
RS_FN bool consider_axis(SatSweep* o, V3 a, SatKind kind, const V3 q[3], V3 half, float margin, float maxd) {
  RS_PRIVATE(o); RS_PRIVATE(q);  // locals of the caller
  float gap, overlap, side;
  gap_overlap(a, q, half, &gap, &overlap, &side);
  if (gap > 0.0f) {
    if (gap - margin >= maxd) return true;
    if (!o->has_sep || gap > o->best_gap) {
      o->has_sep = true;
      o->best_gap = gap;
      o->best_sep_axis = side > 0.0f ? -a : a;
    }
  } else if (!o->has_sep && overlap < o->best_overlap) {
    o->best_overlap = overlap;
    o->pen_kind = kind;
    o->best_pen_raw = a;
  }
  return false;
}
// one box face of the sweep (:373-427)
RS_FN bool box_face(SatSweep* o, float tmin, float tmax, float h, int axis, float margin, float maxd) {
  V3 u = unit_axis(axis);
  if (tmin > h) {
    float gap = tmin - h;
    if (gap - margin >= maxd) return true;
    if (!o->has_sep || gap > o->best_gap) {
      o->has_sep = true;
      o->best_gap = gap;
      o->best_sep_axis = -u;
    }
  } else if (tmax < fneg(h)) {
    float gap = fneg(h) - tmax;
    if (gap - margin >= maxd) return true;
    if (!o->has_sep || gap > o->best_gap) {
      o->has_sep = true;
      o->best_gap = gap;
      o->best_sep_axis = u;
    }
  } else if (!o->has_sep) {
    float overlap = fmax_rs(fmin_rs(tmax + h, h - tmin), 0.0f);
    if (overlap < o->best_overlap) {
      o->best_overlap = overlap;
      o->pen_kind = SatKind{SAT_BOX_FACE, axis, 0};
      o->best_pen_raw = u;
    }
  }
  return false;
}
// :359 sat_sweep; false for V3's None (a rejecting axis)
RS_FN bool sat_sweep(V3 half, const V3 q[3], V3 tmin, V3 tmax, const V3 edges[3], V3 face, bool has_face, V3 centroid, float margin,
                        float maxd, SatSweep* o) {
  RS_PRIVATE(q); RS_PRIVATE(edges); RS_PRIVATE(o);  // locals of the caller
  o->has_sep = false;
  o->best_gap = 0.0f;
  o->best_sep_axis = splat(0.0f);
  o->best_overlap = k::F32_MAX;
  o->best_pen_axis = splat(0.0f);
  o->best_pen_raw = splat(0.0f);
  o->pen_kind = SatKind{SAT_NONE, 0, 0};
  V3 to_box = -centroid;
  if (box_face(o, tmin.x, tmax.x, half.x, 0, margin, maxd)) return false;
  if (box_face(o, tmin.y, tmax.y, half.y, 1, margin, maxd)) return false;
  if (box_face(o, tmin.z, tmax.z, half.z, 2, margin, maxd)) return false;
  if (has_face && consider_axis(o, face, SatKind{SAT_TRI_FACE, 0, 0}, q, half, margin, maxd)) return false;
  for (int bi = 0; bi < 3; ++bi) {
    V3 ba = unit_axis(bi);
    for (int ei = 0; ei < 3; ++ei) {
      V3 c = cross(ba, edges[ei]);
      float l2 = length_sq(c);
      if (l2 < 1e-18f) continue;
      if (consider_axis(o, c / __builtin_sqrtf(l2), SatKind{SAT_EDGE_EDGE, bi, ei}, q, half, margin, maxd)) return false;
    }
  }
  if (!o->has_sep && o->pen_kind.tag != SAT_NONE) o->best_pen_axis = orient_pen(o->best_pen_raw, to_box, face, has_face);
  return true;
}
RS_FN V3 tri_support(const V3 q[3], V3 dir) {  // :554 first maximum
  V3 best = q[0];
  float bd = dot(q[0], dir);
  for (int i = 1; i < 3; ++i) {
    float d = dot(q[i], dir);
    if (d > bd) {
      bd = d;
      best = q[i];
    }
  }
  return best;
}
RS_FN int clip_halfspace(const V3* in, int n, V3* out, int axis, float bound, bool keep_above) {  // :607
  RS_PRIVATE(in); RS_PRIVATE(out);  // locals of the caller
  if (n == 0) return 0;
  int m = 0;
  V3 prev = in[n - 1];
  float pd = keep_above ? comp(prev, axis) - bound : bound - comp(prev, axis);
  for (int i = 0; i < n; ++i) {
    V3 v = in[i];
    float d = keep_above ? comp(v, axis) - bound : bound - comp(v, axis);
    if (d >= 0.0f) {
      if (pd < 0.0f) out[m++] = prev + (v - prev) * (pd / (pd - d));
      out[m++] = v;
    } else if (pd >= 0.0f) {
      out[m++] = prev + (v - prev) * (pd / (pd - d));
    }
    prev = v;
    pd = d;
  }
  return m;
}
RS_INLINE bool clipped_centroid(V3 half, const V3 q[3], V3* res) {  // :576
  RS_PRIVATE(q); RS_PRIVATE(res);  // locals of the caller
  V3 poly[9], tmp[9];
  for (int i = 0; i < 9; ++i) {
    poly[i] = splat(0.0f);
    tmp[i] = splat(0.0f);
  }
  poly[0] = q[0];
  poly[1] = q[1];
  poly[2] = q[2];
  int n = 3;
  for (int ax = 0; ax < 3; ++ax) {
    float h = comp(half, ax);
    n = clip_halfspace(poly, n, tmp, ax, fneg(h), true);
    if (n == 0) return false;
    n = clip_halfspace(tmp, n, poly, ax, h, false);
    if (n == 0) return false;
  }
  V3 sum = splat(0.0f);
  for (int i = 0; i < n; ++i) sum += poly[i];
  *res = sum / (float)n;
  return true;
}
RS_FN V3 clamp_to_box(V3 p, V3 h) {
  return v3(fclamp_rs(p.x, fneg(h.x), h.x), fclamp_rs(p.y, fneg(h.y), h.y), fclamp_rs(p.z, fneg(h.z), h.z));
}
RS_FN V3 closest_on_segment(V3 p, V3 a, V3 b) {  // :657
  V3 ab = b - a;
  float den = dot(ab, ab);
  if (den < 1e-30f) return a;
  return a + ab * fclamp_rs(dot(p - a, ab) / den, 0.0f, 1.0f);
}
RS_FN V3 closest_on_degenerate(V3 p, const V3 q[3]) {  // :669
  V3 c0 = closest_on_segment(p, q[0], q[1]), c1 = closest_on_segment(p, q[1], q[2]), c2 = closest_on_segment(p, q[2], q[0]);
  float d0 = length_sq(p - c0), d1 = length_sq(p - c1), d2 = length_sq(p - c2);
  if (d0 <= d1 && d0 <= d2) return c0;
  if (d1 <= d2) return c1;
  return c2;
}
// segment-segment closest points (:690), with the near-parallel overlap midpoint rule
RS_FN void segseg_closest(V3 p1, V3 q1, V3 p2, V3 q2, V3* o1, V3* o2) {
  RS_PRIVATE(o1); RS_PRIVATE(o2);  // locals of the caller
  V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
  float a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  const float eps = 1e-30f;
  float s, t;
  if (a <= eps && e <= eps) {
    *o1 = p1;
    *o2 = p2;
    return;
  }
  if (a <= eps) {
    s = 0.0f;
    t = fclamp_rs(f / e, 0.0f, 1.0f);
  } else {
    float c = dot(d1, r);
    if (e <= eps) {
      t = 0.0f;
      s = fclamp_rs(fneg(c) / a, 0.0f, 1.0f);
    } else {
      float b = dot(d1, d2);
      if (b * b > (0.99900025f * a) * e) {
        float t0 = dot(p2 - p1, d1) / a, t1 = dot(q2 - p1, d1) / a;
        float lo = fmax_rs(0.0f, fmin_rs(t0, t1)), hi = fmin_rs(1.0f, fmax_rs(t0, t1));
        if (lo <= hi) {
          float sm = (lo + hi) * 0.5f;
          V3 c1 = p1 + d1 * sm;
          float tm = fclamp_rs(dot(c1 - p2, d2) / e, 0.0f, 1.0f);
          *o1 = c1;
          *o2 = p2 + d2 * tm;
          return;
        }
      }
      float den = a * e - b * b;
      s = den > eps ? fclamp_rs((b * f - c * e) / den, 0.0f, 1.0f) : 0.0f;
      t = (b * s + f) / e;
      if (t < 0.0f) {
        t = 0.0f;
        s = fclamp_rs(fneg(c) / a, 0.0f, 1.0f);
      } else if (t > 1.0f) {
        t = 1.0f;
        s = fclamp_rs((b - c) / a, 0.0f, 1.0f);
      }
    }
  }
  *o1 = p1 + d1 * s;
  *o2 = p2 + d2 * t;
}
RS_FN V3 box_corner_toward(V3 h, V3 d) {  // :749
  return v3(d.x < 0.0f ? fneg(h.x) : h.x, d.y < 0.0f ? fneg(h.y) : h.y, d.z < 0.0f ? fneg(h.z) : h.z);
}
RS_FN bool point_in_tri_eps(V3 p, const V3 q[3], V3 n) {  // :762
  bool pos = false, neg = false;
  for (int i = 0; i < 3; ++i) {
    float s = dot(cross(q[(i + 1) % 3] - q[i], p - q[i]), n);
    if (s > 1e-6f) pos = true;
    else if (s < -1e-6f) neg = true;
    if (pos && neg) return false;
  }
  return true;
}
RS_FN bool clip_segment_box(V3 a, V3 b, V3 half, V3* o0, V3* o1) {  // :783
  float t0 = 0.0f, t1 = 1.0f;
  V3 d = b - a;
  for (int ax = 0; ax < 3; ++ax) {
    float h = comp(half, ax);
    for (int kk = 0; kk < 2; ++kk) {
      float p = kk == 0 ? comp(d, ax) : fneg(comp(d, ax));
      float qv = kk == 0 ? comp(a, ax) + h : h - comp(a, ax);
      if (fabs_(p) < k::F32_EPS) {
        if (qv < 0.0f) return false;
      } else {
        float r = qv / p;
        if (p < 0.0f) {
          if (r > t1) return false;
          if (r > t0) t0 = r;
        } else {
          if (r < t0) return false;
          if (r < t1) t1 = r;
        }
      }
    }
  }
  if (t0 > t1) return false;
  *o0 = a + d * t0;
  *o1 = a + d * t1;
  return true;
}
RS_INLINE bool pen_witness(V3 half, const V3 q[3], SatKind kind, V3 n, V3* res) {  // :735
  RS_PRIVATE(q); RS_PRIVATE(res);  // locals of the caller
  if (kind.tag == SAT_TRI_FACE) {
    float pd = dot(n, q[0]);
    V3 corner = box_corner_toward(half, -n);
    V3 p = corner - n * (dot(n, corner) - pd);
    if (point_in_tri_eps(p, q, n)) {
      *res = p;
      return true;
    }
    return false;
  }
  if (kind.tag == SAT_BOX_FACE) {
    float deepest = ffrom(0xFF800000u);
    for (int i = 0; i < 3; ++i) deepest = fmax_rs(deepest, dot(n, q[i]));
    V3 tied[3];
    int nt = 0;
    for (int i = 0; i < 3; ++i)
      if (fabs_(dot(n, q[i]) - deepest) <= SAT_TIE_EPS) tied[nt++] = q[i];
    if (nt == 0) return false;
    if (nt == 1) {
      *res = tied[0];
      return true;
    }
    if (nt == 2) {
      V3 a, b;
      if (!clip_segment_box(tied[0], tied[1], half, &a, &b)) return false;
      *res = (a + b) * 0.5f;
      return true;
    }
    return clipped_centroid(half, q, res);
  }
  if (kind.tag == SAT_EDGE_EDGE) {
    V3 sg = vsignum(-n);
    V3 st = half * sg, en = half * sg;
    set_comp(st, kind.box_axis, comp(half, kind.box_axis));
    set_comp(en, kind.box_axis, fneg(comp(half, kind.box_axis)));
    V3 c1, tp;
    segseg_closest(st, en, q[kind.tri_edge], q[(kind.tri_edge + 1) % 3], &c1, &tp);
    *res = tp;
    return true;
  }
  return false;
}
RS_FN V3 closest_on_tri(V3 p, const V3 q[3], V3 ab, V3 ac) {  // :848 (Ericson's region test)
  RS_PRIVATE(q);  // locals of the caller
  V3 ap = p - q[0];
  float d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0f && d2 <= 0.0f) return q[0];
  V3 bp = p - q[1];
  float d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0f && d4 <= d3) return q[1];
  float vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0f && d1 >= 0.0f && d3 <= 0.0f) return q[0] + ab * (d1 / (d1 - d3));
  V3 cp = p - q[2];
  float d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0f && d5 <= d6) return q[2];
  float vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0f && d2 >= 0.0f && d6 <= 0.0f) return q[0] + ac * (d2 / (d2 - d6));
  float va = d3 * d6 - d5 * d4;
  if (va <= 0.0f && (d4 - d3) >= 0.0f && (d5 - d6) >= 0.0f) return q[1] + (q[2] - q[1]) * ((d4 - d3) / ((d4 - d3) + (d5 - d6)));
  float den = (va + vb) + vc;
  if (fabs_(den) < 1e-24f) return closest_on_degenerate(p, q);
  float v = vb / den, ww = vc / den;
  return q[0] + ab * v + ac * ww;
}
RS_FN void tri_edge_box_edge(V3 tp, V3 td, float ta, V3 b0, V3 b1, V3* o_tri, V3* o_box) {  // :899
  RS_PRIVATE(o_tri); RS_PRIVATE(o_box);  // locals of the caller
  V3 d2 = b1 - b0, r = tp - b0;
  float e = dot(d2, d2), f = dot(d2, r);
  const float eps = 1e-30f;
  if (ta <= eps) {
    if (e <= eps) {
      *o_tri = tp;
      *o_box = b0;
      return;
    }
    *o_tri = tp;
    *o_box = b0 + d2 * fclamp_rs(f / e, 0.0f, 1.0f);
    return;
  }
  if (e <= eps) {
    float c = dot(td, r);
    *o_tri = tp + td * fclamp_rs(fneg(c) / ta, 0.0f, 1.0f);
    *o_box = b0;
    return;
  }
  float c = dot(td, r), b = dot(td, d2);
  if (b * b > (0.99900025f * ta) * e) {
    float t0 = dot(b0 - tp, td) / ta, t1 = dot(b1 - tp, td) / ta;
    float lo = fmax_rs(0.0f, fmin_rs(t0, t1)), hi = fmin_rs(1.0f, fmax_rs(t0, t1));
    if (lo <= hi) {
      float sm = (lo + hi) * 0.5f;
      V3 c1 = tp + td * sm;
      float tm = fclamp_rs(dot(c1 - b0, d2) / e, 0.0f, 1.0f);
      *o_tri = c1;
      *o_box = b0 + d2 * tm;
      return;
    }
  }
  float den = ta * e - b * b;
  float s = den > eps ? fclamp_rs((b * f - c * e) / den, 0.0f, 1.0f) : 0.0f;
  float t = (b * s + f) / e;
  if (t < 0.0f) {
    t = 0.0f;
    s = fclamp_rs(fneg(c) / ta, 0.0f, 1.0f);
  } else if (t > 1.0f) {
    t = 1.0f;
    s = fclamp_rs((b - c) / ta, 0.0f, 1.0f);
  }
  *o_tri = tp + td * s;
  *o_box = b0 + d2 * t;
}
// :960 the 47 candidate pairs, handed to emit in order: 3 clamped vertices, 8 box corners against the triangle, 3 x 12
// edge pairs
template <class Emit>
RS_INLINE void enum_pairs_each(V3 half, const V3 q[3], const V3 edges[3], Emit emit) {
  RS_PRIVATE(q); RS_PRIVATE(edges);  // locals of the caller
  for (int i = 0; i < 3; ++i) {
    V3 pb = clamp_to_box(q[i], half);
    emit(SatPair{pb, q[i], length_sq(pb - q[i])});
  }
  V3 ab = edges[0], ac = q[2] - q[0];
  for (int ix = 0; ix < 2; ++ix)
    for (int iy = 0; iy < 2; ++iy)
      for (int iz = 0; iz < 2; ++iz) {
        V3 corner = v3((ix ? 1.0f : -1.0f) * half.x, (iy ? 1.0f : -1.0f) * half.y, (iz ? 1.0f : -1.0f) * half.z);
        V3 pt = closest_on_tri(corner, q, ab, ac);
        emit(SatPair{corner, pt, length_sq(corner - pt)});
      }
  const V3 c[8] = {v3(fneg(half.x), fneg(half.y), fneg(half.z)), v3(half.x, fneg(half.y), fneg(half.z)),
                   v3(half.x, half.y, fneg(half.z)),             v3(fneg(half.x), half.y, fneg(half.z)),
                   v3(fneg(half.x), fneg(half.y), half.z),       v3(half.x, fneg(half.y), half.z),
                   v3(half.x, half.y, half.z),                   v3(fneg(half.x), half.y, half.z)};
  const int be[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6}, {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
  for (int ti = 0; ti < 3; ++ti) {
    float ta = dot(edges[ti], edges[ti]);
    for (int kk = 0; kk < 12; ++kk) {
      V3 pt, pb;
      tri_edge_box_edge(q[ti], edges[ti], ta, c[be[kk][0]], c[be[kk][1]], &pt, &pb);
      emit(SatPair{pb, pt, length_sq(pb - pt)});
    }
  }
}
RS_INLINE void enum_pairs(V3 half, const V3 q[3], const V3 edges[3], SatPair* pr) {  // all 47, stored
  RS_PRIVATE(pr);  // a local of the caller
  int n = 0;
  enum_pairs_each(half, q, edges, [&](const SatPair& p) { pr[n++] = p; });
}
// The first pair of least distance (strict <: the first pair wins ties), found while the pairs are generated, so the 47
// pairs need not be stored for it.
RS_INLINE SatPair first_min_pair(V3 half, const V3 q[3], const V3 edges[3]) {
  SatPair best;
  bool have = false;
  enum_pairs_each(half, q, edges, [&](const SatPair& p) {
    if (!have || p.d2 < best.d2) best = p;
    have = true;
  });
  return best;
}
// :1096 flat_tie_segment: the triangle-side endpoints of a flat contact interval, or false
RS_INLINE bool flat_tie_segment(V3 half, const SatPair* pr, V3* r0, V3* r1) {
  RS_PRIVATE(pr); RS_PRIVATE(r0); RS_PRIVATE(r1);  // locals of the caller
  float link = 0.5f * max_elem(half);
  if (link <= 0.0f || !ffinite(link)) return false;
  float slope0 = SAT_TIE_EPS / link;
  float best_d2 = k::F32_MAX;
  for (int i = 0; i < SAT_NPAIRS; ++i)
    if (pr[i].d2 < best_d2) best_d2 = pr[i].d2;
  float best_d = __builtin_sqrtf(best_d2);
  float link2 = link * link;
  float tie_d = best_d + SAT_TIE_EPS, tie_d2 = tie_d * tie_d;
  SatPair tied[SAT_NPAIRS];
  int nt = 0;
  for (int i = 0; i < SAT_NPAIRS; ++i)
    if (pr[i].d2 <= tie_d2) tied[nt++] = pr[i];
  if (nt == 0) return false;
  V3 ma = splat(0.0f), mb = splat(0.0f);
  for (int i = 0; i < nt; ++i) {
    ma += tied[i].pbox;
    mb += tied[i].ptri;
  }
  float inv = 1.0f / (float)nt;
  ma *= inv;
  mb *= inv;
  float d0 = length(ma - mb);
  bool has_flat = false;
  SatPair flat = SatPair{};
  float best_slope = slope0;
  for (int i = 0; i < SAT_NPAIRS; ++i) {
    const SatPair& p = pr[i];
    float sep2 = length_sq(p.ptri - mb);
    if (sep2 <= link2) continue;
    float slope = (__builtin_sqrtf(p.d2) - d0) / __builtin_sqrtf(sep2);
    if (slope <= best_slope) {
      V3 mida = (ma + p.pbox) * 0.5f, midb = (mb + p.ptri) * 0.5f;
      if (length(mida - midb) <= d0 + SAT_TIE_EPS) {
        best_slope = slope;
        flat = p;
        has_flat = true;
      }
    }
  }
  SatPair e0 = tied[0], e1 = tied[0];
  float span2 = 0.0f;
  int nc = nt + (has_flat ? 1 : 0);
  for (int ci = 0; ci < nc; ++ci) {
    SatPair cand = ci < nt ? tied[ci] : flat;
    for (int pi = 0; pi < nc; ++pi) {
      const SatPair& p = pi < nt ? tied[pi] : flat;
      float s2 = length_sq(cand.ptri - p.ptri);
      if (s2 > span2) {
        span2 = s2;
        e0 = cand;
        e1 = p;
      }
    }
  }
  if (span2 <= link2) return false;
  V3 mida = (e0.pbox + e1.pbox) * 0.5f, midb = (e0.ptri + e1.ptri) * 0.5f;
  if (length(mida - midb) > best_d + SAT_TIE_EPS) return false;
  *r0 = e0.ptri;
  *r1 = e1.ptri;
  return true;
}
// :52 box_triangle_sat: q is the triangle in the box frame, half the unmargined half extents. True for V3's Some.
// One call per test; its four stages (the pair enumeration, the flat-tie segment, the penetration witness and the clipped
// centroid) are inlined into it, each having this single caller: as calls they saved and restored the caller's
// registers through private memory on every test.
RS_INLINE bool box_triangle_sat_body(const Aff& box_trans, V3 half, float margin, const V3 q[3], float maxd, SatContact* out) {
  V3 tmin = vmin(vmin(q[0], q[1]), q[2]), tmax = vmax(vmax(q[0], q[1]), q[2]);
  for (int ax = 0; ax < 3; ++ax) {  // pre-screen, X then Y then Z
    float lo = comp(tmin, ax), hi = comp(tmax, ax), h = comp(half, ax);
    if (lo > h) {
      if ((lo - h) - margin >= maxd) return false;
    } else if (hi < fneg(h) && (fneg(h) - hi) - margin >= maxd) {
      return false;
    }
  }
  V3 e0 = q[1] - q[0], e1 = q[2] - q[1], e2 = q[0] - q[2];
  V3 fc = cross(e0, -e2);
  float fl2 = length_sq(fc);
  bool has_face = fl2 > 1e-20f;
  V3 face = has_face ? fc / __builtin_sqrtf(fl2) : splat(0.0f);
  V3 centroid = (q[0] + q[1] + q[2]) * 0.33333334f;
  V3 edges[3] = {e0, e1, e2};
  SatSweep sw;
  if (!sat_sweep(half, q, tmin, tmax, edges, face, has_face, centroid, margin, maxd, &sw)) return false;
  if (sw.has_sep) {
    V3 nw = normalize_or_zero(box_trans.m * sw.best_sep_axis);
    if (length_sq(nw) < 0.5f) return false;
    const SatPair first = first_min_pair(half, q, edges);
    if (length_sq(first.pbox - first.ptri) > 1e-24f) {
      float plen = length(first.pbox - first.ptri);
      V3 pb = first.ptri;
      bool snapped = false;
      V3 s0, s1;
      if (plen - margin < 0.0f) {  // the flat-tie segment reads every pair: they are generated again, the same way
        SatPair pr[SAT_NPAIRS];
        enum_pairs(half, q, edges, pr);
        if (flat_tie_segment(half, pr, &s0, &s1)) {
          pb = closest_on_segment(centroid, s0, s1);
          snapped = true;
        }
      }
      float dist = snapped ? length(clamp_to_box(pb, half) - pb) - margin : plen - margin;
      if (dist >= maxd) return false;
      out->point_world = xform_point(box_trans, pb);
      out->distance = dist;
      return true;
    }
    out->point_world = xform_point(box_trans, tri_support(q, sw.best_sep_axis));
    out->distance = sw.best_gap - margin;
    return true;
  }
  if (sw.best_overlap == k::F32_MAX) return false;
  V3 n;
  if (length_sq(sw.best_pen_axis) > 0.5f) {
    n = sw.best_pen_axis;
  } else if (has_face) {
    n = dot(face, -centroid) < 0.0f ? -face : face;
  } else {
    return false;
  }
  float dist = fneg(sw.best_overlap) - margin;
  if (dist >= maxd) return false;
  V3 pb;
  if (!pen_witness(half, q, sw.pen_kind, n, &pb))
    if (!clipped_centroid(half, q, &pb)) pb = tri_support(q, n);
  V3 nw = normalize_or_zero(box_trans.m * n);
  if (length_sq(nw) < 0.5f) return false;
  out->point_world = xform_point(box_trans, pb);
  out->distance = dist;
  return true;
}
// The test as a function of its own, its inputs and result passed in registers: through pointers to the caller's
// private memory, the caller would store them for the call and the callee read them back behind those stores (and the
// result the other way), on every test of the pooled turns. The box transform and half extents fill the calling
// convention's register budget for aggregates, so the triangle's corners go as nine scalars.
struct SatResult {
  SatContact c;
  bool ok;
};
RS_NOINL SatResult box_triangle_sat_in(Aff box_trans, V3 half, float margin, float q0x, float q0y, float q0z, float q1x,
                                       float q1y, float q1z, float q2x, float q2y, float q2z, float maxd) {
  const V3 q[3] = {v3(q0x, q0y, q0z), v3(q1x, q1y, q1z), v3(q2x, q2y, q2z)};
  SatResult r;
  r.c = SatContact{};
  r.ok = box_triangle_sat_body(box_trans, half, margin, q, maxd, &r.c);
  return r;
}
RS_FN bool box_triangle_sat(const Aff& box_trans, V3 half, float margin, const V3 q[3], float maxd, SatContact* out) {
  const SatResult r = box_triangle_sat_in(box_trans, half, margin, q[0].x, q[0].y, q[0].z, q[1].x, q[1].y, q[1].z, q[2].x,
                                          q[2].y, q[2].z, maxd);
  if (r.ok) *out = r.c;
  return r.ok;
}

}  // namespace rs
