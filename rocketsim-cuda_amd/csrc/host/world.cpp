// The synthetic Soccar world and start states (see world.h). Everything here runs once on the host; its float outputs are
// computed in double and rounded once, or with the core's own float functions where the tick recomputes them.
#include "host/world.h"
#include "rsim.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <map>
#include <vector>

namespace rs {
namespace world {
namespace {

// ---- arena geometry (uu) ----------------------------------------------------------------------------------------------------
constexpr double AX = 4096.0, AY = 5120.0, AH = 2044.0;  // half width, half length, height
constexpr double FR = 256.0;                               // floor / ceiling fillet radius
constexpr double CR = 1024.0;                              // vertical corner radius
constexpr double GX = 893.0, GH = 642.775, GD = 880.0;     // goal half width, height, depth
constexpr double PI_D = 3.14159265358979323846;

struct D3 {
  double x, y, z;
};
D3 operator+(D3 a, D3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
D3 operator-(D3 a, D3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
D3 operator*(D3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
D3 dcross(D3 a, D3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
D3 dnorm(D3 a) {
  double l = sqrt(ddot(a, a));
  return l > 0 ? a * (1.0 / l) : a;
}
V3 to_bt(D3 p) { return V3{(float)(p.x / 50.0), (float)(p.y / 50.0), (float)(p.z / 50.0)}; }

struct MeshBuild {
  std::vector<Tri> tris;
};
// A quad a-b-c-d (in order around it) as two triangles whose normals face `inward`.
void add_quad(MeshBuild& m, D3 a, D3 b, D3 c, D3 d, D3 inward) {
  V3 pa = to_bt(a), pb = to_bt(b), pc = to_bt(c), pd = to_bt(d);
  auto tri = [&](V3 p0, V3 p1, V3 p2) {
    V3 n = cross(p1 - p0, p2 - p0);
    D3 nd = {n.x, n.y, n.z};
    if (ddot(nd, inward) < 0) {
      V3 t = p1;
      p1 = p2;
      p2 = t;
      n = cross(p1 - p0, p2 - p0);
    }
    double l = sqrt((double)n.x * n.x + (double)n.y * n.y + (double)n.z * n.z);
    if (l <= 0) return;  // degenerate
    Tri t;
    t.p0 = p0;
    t.p1 = p1;
    t.p2 = p2;
    t.normal = V3{(float)(n.x / l), (float)(n.y / l), (float)(n.z / l)};
    m.tris.push_back(t);
  };
  tri(pa, pb, pc);
  tri(pa, pc, pd);
}
// A surface patch p(u, v) on a (nu x nv) grid of quads; `inward(u, v)` the side the normals face.
template <class P, class N>
void add_patch(MeshBuild& m, int nu, int nv, P p, N inward) {
  for (int i = 0; i < nu; ++i)
    for (int j = 0; j < nv; ++j) {
      double u0 = (double)i / nu, u1 = (double)(i + 1) / nu, v0 = (double)j / nv, v1 = (double)(j + 1) / nv;
      add_quad(m, p(u0, v0), p(u1, v0), p(u1, v1), p(u0, v1), inward((u0 + u1) * 0.5, (v0 + v1) * 0.5));
    }
}
// The wall profile at parameter t in [0, 1] over [z0 part]: bottom fillet, vertical wall, top fillet. Returns the inset d
// from the wall line (toward the inside) and the height z, and the inward profile normal (nh horizontal, nz vertical).
struct Prof {
  double d, z, nh, nz;
};
Prof bottom_fillet(double t) {
  double ph = t * PI_D * 0.5;
  return {FR - FR * sin(ph), FR - FR * cos(ph), sin(ph), cos(ph)};
}
Prof top_fillet(double t) {
  double ph = (1.0 - t) * PI_D * 0.5;
  return {FR - FR * sin(ph), AH - FR + FR * cos(ph), sin(ph), -cos(ph)};
}
Prof wall(double z0, double z1, double t) { return {0.0, z0 + (z1 - z0) * t, 1.0, 0.0}; }

// A straight wall segment from a to b (horizontal points on the wall line), inward horizontal direction n.
template <class F>
void wall_strip(MeshBuild& m, D3 a, D3 b, D3 n, int nu, int nv, F prof) {
  add_patch(
      m, nu, nv,
      [&](double u, double v) {
        Prof p = prof(v);
        D3 q = a + (b - a) * u;
        return D3{q.x + n.x * p.d, q.y + n.y * p.d, p.z};
      },
      [&](double, double v) {
        Prof p = prof(v);
        return D3{n.x * p.nh, n.y * p.nh, p.nz};
      });
}
// A vertical corner arc: centre c, radius CR, angles th0..th1 (outward direction (cos, sin)).
template <class F>
void corner_strip(MeshBuild& m, D3 c, double th0, double th1, int nu, int nv, F prof) {
  add_patch(
      m, nu, nv,
      [&](double u, double v) {
        double th = th0 + (th1 - th0) * u;
        D3 out = {cos(th), sin(th), 0.0};
        Prof p = prof(v);
        double r = CR - p.d;
        return D3{c.x + out.x * r, c.y + out.y * r, p.z};
      },
      [&](double u, double v) {
        double th = th0 + (th1 - th0) * u;
        Prof p = prof(v);
        return D3{-cos(th) * p.nh, -sin(th) * p.nh, p.nz};
      });
}

std::vector<MeshBuild> arena_meshes() {
  std::vector<MeshBuild> ms(NUM_MESHES);
  auto full = [](int nb, int nw, int nt) {  // counts of a full profile
    return nb + nw + nt;
  };
  (void)full;
  const int NFIL = 6;  // fillet segments
  // 0-3: the side walls' floor and ceiling fillets (x = +-AX, straight part)
  for (int side = 0; side < 2; ++side) {
    double x = side == 0 ? AX : -AX;
    D3 n = {side == 0 ? -1.0 : 1.0, 0.0, 0.0};
    D3 a = {x, -(AY - CR), 0.0}, b = {x, AY - CR, 0.0};
    wall_strip(ms[side * 2 + 0], a, b, n, 40, NFIL, bottom_fillet);
    wall_strip(ms[side * 2 + 1], a, b, n, 40, NFIL, top_fillet);
  }
  // 4-7: the four vertical corners with their fillets
  const D3 cc[4] = {{AX - CR, AY - CR, 0}, {-(AX - CR), AY - CR, 0}, {-(AX - CR), -(AY - CR), 0}, {AX - CR, -(AY - CR), 0}};
  for (int q = 0; q < 4; ++q) {
    double th0 = q * PI_D * 0.5, th1 = th0 + PI_D * 0.5;
    corner_strip(ms[4 + q], cc[q], th0, th1, 14, NFIL, bottom_fillet);
    corner_strip(ms[4 + q], cc[q], th0, th1, 14, 7, [](double t) { return wall(FR, AH - FR, t); });
    corner_strip(ms[4 + q], cc[q], th0, th1, 14, NFIL, top_fillet);
  }
  // 8-11: the back walls beside the goals; 12-13 above the goals; 14-15 the goals
  for (int end = 0; end < 2; ++end) {
    double y = end == 0 ? AY : -AY, sy = end == 0 ? 1.0 : -1.0;
    D3 n = {0.0, -sy, 0.0};
    for (int half = 0; half < 2; ++half) {
      double sx = half == 0 ? -1.0 : 1.0;
      D3 a = {sx * (AX - CR), y, 0.0}, b = {sx * GX, y, 0.0};
      MeshBuild& m = ms[8 + end * 2 + half];
      wall_strip(m, a, b, n, 12, NFIL, bottom_fillet);
      wall_strip(m, a, b, n, 12, 7, [](double t) { return wall(FR, AH - FR, t); });
      wall_strip(m, a, b, n, 12, NFIL, top_fillet);
    }
    MeshBuild& above = ms[12 + end];
    D3 a = {-GX, y, 0.0}, b = {GX, y, 0.0};
    wall_strip(above, a, b, n, 12, 6, [](double t) { return wall(GH, AH - FR, t); });
    wall_strip(above, a, b, n, 12, NFIL, top_fillet);
    MeshBuild& g = ms[14 + end];
    double yb = y + sy * GD;
    add_patch(  // side walls x = -GX (normal +x) and x = +GX (normal -x)
        g, 8, 6, [&](double u, double v) { return D3{-GX, y + sy * GD * u, GH * v}; }, [](double, double) { return D3{1, 0, 0}; });
    add_patch(
        g, 8, 6, [&](double u, double v) { return D3{GX, y + sy * GD * u, GH * v}; }, [](double, double) { return D3{-1, 0, 0}; });
    add_patch(  // back
        g, 12, 6, [&](double u, double v) { return D3{-GX + 2 * GX * u, yb, GH * v}; }, [&](double, double) { return D3{0, -sy, 0}; });
    add_patch(  // ceiling
        g, 12, 8, [&](double u, double v) { return D3{-GX + 2 * GX * u, y + sy * GD * v, GH}; }, [](double, double) { return D3{0, 0, -1}; });
  }
  return ms;
}

// ---- BVH4 (one leaf per triangle) ----------------------------------------------------------------------------------------------
Aabb tri_box(const Tri& t) {
  Aabb b;
  b.min = vmin(vmin(t.p0, t.p1), t.p2);
  b.max = vmax(vmax(t.p0, t.p1), t.p2);
  const float e = 1e-4f;  // a hair of slack, as V3's quantised leaf boxes have
  b.min = b.min - splat(e);
  b.max = b.max + splat(e);
  return b;
}
Aabb box_union(const Aabb& a, const Aabb& b) { return Aabb{vmin(a.min, b.min), vmax(a.max, b.max)}; }

struct BvhOut {
  std::vector<WideNode> nodes;
  std::vector<Leaf> leaves;
};
// Builds the subtree over items (triangle indices); returns its child word and its box.
uint32_t bvh_build(BvhOut& o, const std::vector<Tri>& tris, std::vector<uint32_t> items, Aabb* box_out) {
  if (items.size() == 1) {
    Leaf lf;
    lf.bounds = tri_box(tris[items[0]]);
    lf.tri = items[0];
    o.leaves.push_back(lf);
    *box_out = lf.bounds;
    return BVH_LEAF | (uint32_t)(o.leaves.size() - 1);
  }
  uint32_t node_idx = (uint32_t)o.nodes.size();
  o.nodes.push_back(WideNode{});
  // split into up to four groups: sort along the longest centroid axis and cut it into slabs of whole subtrees (the largest
  // power of four below the count), which keeps the node count near a third of the triangle count
  auto centroid = [&](uint32_t i) {
    const Tri& t = tris[i];
    return V3{(t.p0.x + t.p1.x + t.p2.x) / 3.0f, (t.p0.y + t.p1.y + t.p2.y) / 3.0f, (t.p0.z + t.p1.z + t.p2.z) / 3.0f};
  };
  std::vector<std::vector<uint32_t>> groups;
  if (items.size() <= 4) {
    for (uint32_t i : items) groups.push_back({i});
  } else {
    V3 lo = centroid(items[0]), hi = lo;
    for (uint32_t i : items) {
      lo = vmin(lo, centroid(i));
      hi = vmax(hi, centroid(i));
    }
    V3 ext = hi - lo;
    int ax = ext.x >= ext.y && ext.x >= ext.z ? 0 : ext.y >= ext.z ? 1 : 2;
    std::stable_sort(items.begin(), items.end(), [&](uint32_t p, uint32_t q) { return comp(centroid(p), ax) < comp(centroid(q), ax); });
    size_t cap = 4;
    while (cap * 4 < items.size()) cap *= 4;
    for (size_t at = 0; at < items.size(); at += cap)
      groups.push_back(std::vector<uint32_t>(items.begin() + at, items.begin() + std::min(items.size(), at + cap)));
  }
  WideNode n = WideNode{};
  Aabb total;
  for (size_t c = 0; c < groups.size(); ++c) {
    Aabb cb;
    uint32_t child = bvh_build(o, tris, groups[c], &cb);
    n.children[c] = child;
    n.min_x[c] = cb.min.x;
    n.min_y[c] = cb.min.y;
    n.min_z[c] = cb.min.z;
    n.max_x[c] = cb.max.x;
    n.max_y[c] = cb.max.y;
    n.max_z[c] = cb.max.z;
    total = c == 0 ? cb : box_union(total, cb);
  }
  n.child_count = (uint32_t)groups.size();
  n.valid_mask = (1u << n.child_count) - 1u;
  o.nodes[node_idx] = n;
  *box_out = total;
  return node_idx;
}

// ---- internal-edge angles (Bullet btGenerateInternalEdgeInfo's method) --------------------------------------------------------
struct VKey {
  uint32_t x, y, z;
  bool operator<(const VKey& o) const { return x != o.x ? x < o.x : y != o.y ? y < o.y : z < o.z; }
  bool operator==(const VKey& o) const { return x == o.x && y == o.y && z == o.z; }
};
VKey vkey(V3 p) { return VKey{fbits(p.x), fbits(p.y), fbits(p.z)}; }
D3 d3(V3 p) { return D3{p.x, p.y, p.z}; }
V3 tri_vertex(const Tri& t, int i) { return i == 0 ? t.p0 : i == 1 ? t.p1 : t.p2; }

void edge_infos(const std::vector<Tri>& tris, TriInfo* out) {
  const float NO_EDGE = 6.28318548f;  // 2 pi: not an internal edge
  std::map<std::pair<VKey, VKey>, std::vector<std::pair<uint32_t, int>>> edges;
  for (uint32_t i = 0; i < tris.size(); ++i) {
    out[i] = TriInfo{0u, NO_EDGE, NO_EDGE, NO_EDGE};
    for (int e = 0; e < 3; ++e) {
      VKey a = vkey(tri_vertex(tris[i], e)), b = vkey(tri_vertex(tris[i], (e + 1) % 3));
      if (b < a) std::swap(a, b);
      edges[{a, b}].push_back({i, e});
    }
  }
  for (uint32_t i = 0; i < tris.size(); ++i) {
    const Tri& A = tris[i];
    for (int e = 0; e < 3; ++e) {
      VKey ka = vkey(tri_vertex(A, e)), kb = vkey(tri_vertex(A, (e + 1) % 3));
      if (kb < ka) std::swap(ka, kb);
      const auto& list = edges[{ka, kb}];
      int other = -1, oe = -1;
      for (auto& pr : list)
        if (pr.first != i) {
          other = (int)pr.first;
          oe = pr.second;
        }
      if (other < 0) continue;
      const Tri& B = tris[other];
      D3 v0 = d3(tri_vertex(A, e)), v1 = d3(tri_vertex(A, (e + 1) % 3)), va = d3(tri_vertex(A, (e + 2) % 3));
      D3 vb = d3(tri_vertex(B, (oe + 2) % 3));
      D3 nA = d3(A.normal), nB = d3(B.normal);
      D3 edge = v1 - v0;
      D3 ecA = dnorm(dcross(edge, nA));
      if (ddot(ecA, va - v0) < 0) ecA = ecA * -1.0;
      D3 ecB = dnorm(dcross(edge, nB));
      if (ddot(ecB, vb - v0) < 0) ecB = ecB * -1.0;
      D3 calc = dcross(ecA, ecB);
      double corrected = 0.0;
      bool convex = false;
      if (ddot(calc, calc) >= 0.0001) {
        calc = dnorm(calc);
        D3 cnA = dnorm(dcross(calc, ecA));
        double angle2 = atan2(ddot(ecB, cnA), ddot(ecB, ecA));
        double ang4 = PI_D - angle2;
        convex = ddot(nA, ecB) < 0.0;
        corrected = convex ? ang4 : -ang4;
      }
      // the stored angle rotates A's normal about (v_e - v_e+1) onto +-B's normal
      double stored = -corrected;
      D3 axis = dnorm(v0 - v1);
      double s = sin(stored * 0.5), c = cos(stored * 0.5);
      D3 qv = axis * s;  // rotate nA by quaternion (qv, c)
      D3 t = dcross(qv, nA) * 2.0;
      D3 rot = nA + t * c + dcross(qv, t);
      uint32_t swap_bit = e == 0 ? 8u : e == 1 ? 16u : 32u, convex_bit = e == 0 ? 1u : e == 1 ? 2u : 4u;
      if (ddot(rot, nB) < 0) out[i].flags |= swap_bit;
      if (convex) out[i].flags |= convex_bit;
      float fa = (float)stored;
      if (e == 0) out[i].edge_v0_v1_angle = fa;
      else if (e == 1) out[i].edge_v1_v2_angle = fa;
      else out[i].edge_v2_v0_angle = fa;
    }
  }
}

// ---- car presets (uu): hitbox size, hitbox offset, front / back wheel radius, suspension rest length, connection point ----
struct Preset {
  float hx, hy, hz, ox, oy, oz;
  float fr, frest, fcx, fcy, fcz;
  float br, brest, bcx, bcy, bcz;
  uint32_t three_wheels;
};
const Preset PRESETS[NUM_PRESETS] = {
    {120.507f, 86.6994f, 38.6591f, 13.8757f, 0.0f, 20.755f, 12.50f, 38.755f, 51.25f, 25.90f, 20.755f, 15.00f, 37.055f, -33.75f, 29.50f, 20.755f, 0},
    {130.427f, 85.7799f, 33.8f, 9.0f, 0.0f, 15.75f, 12.00f, 33.95f, 50.30f, 31.10f, 15.75f, 13.50f, 33.85f, -34.75f, 33.00f, 15.75f, 0},
    {131.32f, 87.1704f, 31.8944f, 9.008f, 0.0f, 12.0942f, 12.50f, 31.9242f, 49.97f, 27.80f, 12.0942f, 17.00f, 27.9242f, -35.43f, 20.28f, 12.0942f, 0},
    {133.992f, 83.021f, 32.8f, 12.5f, 0.0f, 11.75f, 13.50f, 29.70f, 51.50f, 26.67f, 11.75f, 15.00f, 29.666f, -35.75f, 35.00f, 11.75f, 0},
    {129.519f, 84.6879f, 36.6591f, 13.8757f, 0.0f, 20.755f, 12.50f, 38.755f, 51.25f, 25.90f, 20.755f, 15.00f, 37.055f, -34.00f, 29.50f, 20.755f, 0},
    {123.22f, 79.2103f, 44.1591f, 11.3757f, 0.0f, 21.505f, 15.00f, 39.505f, 51.25f, 25.90f, 21.505f, 15.00f, 39.105f, -33.75f, 29.50f, 21.505f, 0},
    {124.0f, 82.0f, 39.0f, 12.0f, 0.0f, 20.0f, 13.00f, 37.5f, 52.0f, 0.0f, 20.0f, 15.00f, 37.0f, -34.0f, 31.0f, 20.0f, 1},
};

const float PAD_POS[NUM_PADS][3] = {
    {0, -4240, 70},     {-1792, -4184, 70}, {1792, -4184, 70}, {-3072, -4096, 73}, {3072, -4096, 73}, {-940, -3308, 70},
    {940, -3308, 70},   {0, -2816, 70},     {-3584, -2484, 70}, {3584, -2484, 70}, {-1788, -2300, 70}, {1788, -2300, 70},
    {-2048, -1036, 70}, {0, -1024, 70},     {2048, -1036, 70}, {-3584, 0, 73},    {-1024, 0, 70},     {1024, 0, 70},
    {3584, 0, 73},      {-2048, 1036, 70},  {0, 1024, 70},     {2048, 1036, 70},  {-1788, 2300, 70},  {1788, 2300, 70},
    {-3584, 2484, 70},  {3584, 2484, 70},   {0, 2816, 70},     {-940, 3310, 70},  {940, 3308, 70},    {-3072, 4096, 73},
    {3072, 4096, 73},   {-1792, 4184, 70},  {1792, 4184, 70},  {0, 4240, 70}};

M3 rot_from_ypr(double yaw, double pitch, double roll) {  // forward x, right y, up z; yaw about z, pitch about -y, roll about x
  double cy = cos(yaw), sy = sin(yaw), cp = cos(pitch), sp = sin(pitch), cr = cos(roll), sr = sin(roll);
  D3 f = {cp * cy, cp * sy, sp};
  D3 r0 = {-sy, cy, 0.0}, u0 = dcross(f, r0);
  D3 r = r0 * cr + u0 * sr, u = u0 * cr - r0 * sr;
  return M3{V3{(float)f.x, (float)f.y, (float)f.z}, V3{(float)r.x, (float)r.y, (float)r.z}, V3{(float)u.x, (float)u.y, (float)u.z}};
}

// Bullet AxisSweep3 sorts with the overlap update switchable (construction sorts axes 0 and 1 without it).
void sap_min_down(ArenaState& s, const StaticData& w, int ax, uint32_t i, bool upd) {
  Edge* e = s.edges[ax];
  uint32_t mv = e[i].handle;
  while (i > 0 && e[i].pos < e[i - 1].pos) {
    Edge prev = e[i - 1];
    if (prev.pos & 1u) {
      if (upd && overlap_2d(s, mv, prev.handle, ax)) add_pair(s, w, mv, prev.handle);
      s.handle_edges[prev.handle].max[ax] += 1;
    } else {
      s.handle_edges[prev.handle].min[ax] += 1;
    }
    s.handle_edges[mv].min[ax] -= 1;
    e[i - 1] = e[i];
    e[i] = prev;
    --i;
  }
}
void sap_max_down(ArenaState& s, const StaticData& w, int ax, uint32_t i, bool upd) {
  Edge* e = s.edges[ax];
  uint32_t mv = e[i].handle;
  while (i > 0 && e[i].pos < e[i - 1].pos) {
    Edge prev = e[i - 1];
    if (!(prev.pos & 1u)) {
      if (upd && overlap_2d(s, mv, prev.handle, ax)) remove_pair(s, w, mv, prev.handle);
      s.handle_edges[prev.handle].min[ax] += 1;
    } else {
      s.handle_edges[prev.handle].max[ax] += 1;
    }
    s.handle_edges[mv].max[ax] -= 1;
    e[i - 1] = e[i];
    e[i] = prev;
    --i;
  }
}
void sap_add_handle(ArenaState& s, const StaticData& w, uint32_t h, const Aabb& box, uint32_t nreal) {
  uint16_t qmin[3], qmax[3];
  sap_quantize(w, box.min, 0, qmin);
  sap_quantize(w, box.max, 1, qmax);
  uint32_t limit = 2 * nreal;  // nreal counts this handle
  for (int ax = 0; ax < 3; ++ax) {
    s.handle_edges[0].max[ax] = (uint16_t)(s.handle_edges[0].max[ax] + 2);
    s.edges[ax][limit + 1] = s.edges[ax][limit - 1];
    s.edges[ax][limit - 1] = Edge{qmin[ax], (uint16_t)h};
    s.edges[ax][limit] = Edge{qmax[ax], (uint16_t)h};
    s.handle_edges[h].min[ax] = (uint16_t)(limit - 1);
    s.handle_edges[h].max[ax] = (uint16_t)limit;
  }
  for (int ax = 0; ax < 3; ++ax) {
    sap_min_down(s, w, ax, s.handle_edges[h].min[ax], ax == 2);
    sap_max_down(s, w, ax, s.handle_edges[h].max[ax], ax == 2);
  }
}

Aabb static_box(const StaticData& w, uint32_t body) {
  const BodyConst& bc = w.bodies[body];
  if (bc.shape_kind == k::SHAPE_PLANE) return w.planes[body - NUM_MESHES].aabb;
  const MeshInfo& m = w.meshes[body];
  return Aabb{m.aabb_ident.min + bc.pose.t, m.aabb_ident.max + bc.pose.t};
}

}  // namespace

// ---- the static world ---------------------------------------------------------------------------------------------------------
void build_world(StaticData* wp, int p0, int p1, const Mutators& mut) {
  StaticData& w = *wp;
  memset(&w, 0, sizeof(StaticData));
  w.game_mode = 0;
  w.no_ball_rot = 0;
  w.ball_body = 20;
  w.gravity_uu = V3{0.0f, 0.0f, mut.gravity_z};
  const float mf[NUM_MUT] = {180.0f, 30.0f, 6000.0f, 0.03f, 4375.0f / 3.0f, 875.0f / 3.0f, 2975.0f / 3.0f * mut.boost_accel_scale,
                             3175.0f / 3.0f * mut.boost_accel_scale, 100.0f / 3.0f, 3.0f, 0.25f, 100.0f, 100.0f / 3.0f, 12.0f,
                             100.0f, 10.0f, 4.0f, 1.0f, 1.0f, mut.ball_radius};
  for (int i = 0; i < NUM_MUT; ++i) w.mut_f[i] = mf[i];
  w.recharge_boost_per_second = 10.0f;
  w.recharge_boost_delay = 0.25f;
  w.goal_base_threshold_y = 5124.25f;
  w.boost_pads_enabled = 1;
  w.gravity_bt = w.gravity_uu * k::UU_TO_BT;
  w.dyn_body[0] = 20;
  w.dyn_body[1] = 21;
  w.dyn_body[2] = 22;
  // ray grid: cubic cells of 23 BT over the arena and goals
  w.grid_min = V3{-84.0f, -122.0f, -2.0f};
  w.grid_cell_size = 23.0f;
  w.grid_cell_size_sq = 23.0f * 23.0f;
  w.grid_num_cells[0] = 8;
  w.grid_num_cells[1] = 11;
  w.grid_num_cells[2] = 2;
  w.grid_max = w.grid_min + V3{8 * 23.0f, 11 * 23.0f, 2 * 23.0f};
  // sweep and prune: u16 positions over a box around the world
  w.sap_world_min = V3{-100.0f, -130.0f, -20.0f};
  V3 sap_max = V3{100.0f, 130.0f, 60.0f};
  w.sap_quantize = V3{65533.0f, 65533.0f, 65533.0f} / (sap_max - w.sap_world_min);
  w.sap_handle_mask = 0xFFFEu;
  w.sap_sentinel = 0xFFFFu;
  w.sap_max_handles = NUM_HANDLES;
  w.sap_world_half_extent = 100.0f;
  w.dyn_handle[0] = 21;
  w.dyn_handle[1] = 22;
  w.dyn_handle[2] = 23;

  // meshes, BVHs, triangle infos
  std::vector<MeshBuild> ms = arena_meshes();
  uint32_t tri_at = 0, node_at = 0, leaf_at = 0;
  for (int mi = 0; mi < NUM_MESHES; ++mi) {
    const std::vector<Tri>& tris = ms[mi].tris;
    BvhOut bo;
    std::vector<uint32_t> items(tris.size());
    for (uint32_t i = 0; i < tris.size(); ++i) items[i] = i;
    Aabb root;
    bvh_build(bo, tris, items, &root);
    if (tri_at + tris.size() > (size_t)MAX_TRIS || node_at + bo.nodes.size() > (size_t)MAX_NODES ||
        leaf_at + bo.leaves.size() > (size_t)MAX_LEAVES) {
      fprintf(stderr, "world: mesh %d does not fit the StaticData caps (%u+%zu tris, %u+%zu nodes, %u+%zu leaves)\n", mi, tri_at,
              tris.size(), node_at, bo.nodes.size(), leaf_at, bo.leaves.size());
      abort();
    }
    MeshInfo& m = w.meshes[mi];
    m.body = (uint32_t)mi;
    m.tri_first = tri_at;
    m.tri_count = (uint32_t)tris.size();
    m.node_first = node_at;
    m.node_count = (uint32_t)bo.nodes.size();
    m.leaf_first = leaf_at;
    m.leaf_count = (uint32_t)bo.leaves.size();
    m.aabb_ident = root;
    m.root_aabb = root;
    for (size_t i = 0; i < tris.size(); ++i) w.tris[tri_at + i] = tris[i];
    for (size_t i = 0; i < bo.nodes.size(); ++i) w.nodes[node_at + i] = bo.nodes[i];
    for (size_t i = 0; i < bo.leaves.size(); ++i) w.leaves[leaf_at + i] = bo.leaves[i];
    edge_infos(tris, w.infos + tri_at);
    tri_at += (uint32_t)tris.size();
    node_at += (uint32_t)bo.nodes.size();
    leaf_at += (uint32_t)bo.leaves.size();
  }
  // planes: floor, ceiling, +x wall, -x wall (bodies 16-19); the plane passes through the body's origin
  const float H = (float)(AH / 50.0), X = (float)(AX / 50.0), L = 200.0f, T = 100.0f;
  const V3 pn[4] = {{0, 0, 1}, {0, 0, -1}, {-1, 0, 0}, {1, 0, 0}};
  const V3 pt[4] = {{0, 0, 0}, {0, 0, H}, {X, 0, 0}, {-X, 0, 0}};
  const Aabb pid[4] = {{{-L, -L, -T}, {L, L, 0}}, {{-L, -L, 0}, {L, L, T}}, {{0, -L, -L}, {T, L, L}}, {{-T, -L, -L}, {0, L, L}}};
  for (int p = 0; p < NUM_PLANES; ++p) {
    Plane& pl = w.planes[p];
    pl.body = (uint32_t)(NUM_MESHES + p);
    pl.normal = pn[p];
    pl.is_single_axis = 1;
    pl.single_axis_idx = p < 2 ? 2u : 0u;
    pl.single_axis_backwards = (p == 1 || p == 2) ? 1u : 0u;
    pl.aabb_ident = pid[p];
    pl.aabb = Aabb{pid[p].min + pt[p], pid[p].max + pt[p]};
  }
  // bodies
  for (uint32_t b = 0; b < (uint32_t)NUM_BODIES; ++b) {
    BodyConst& bc = w.bodies[b];
    bc.world_array_idx = b;
    bc.broadphase_handle = b + 1;
    bc.contact_processing_threshold = 1e18f;
    bc.contact_breaking_threshold = k::CONTACT_BREAKING_THRESHOLD;
    bc.linear_sleeping_threshold = 0.8f;
    bc.angular_sleeping_threshold = 1.0f;
    bc.pose = aff_identity();
    if (b < (uint32_t)(NUM_MESHES + NUM_PLANES)) {  // static arena
      bc.user_idx = k::USER_NONE;
      bc.filter_group = k::FILTER_STATIC;
      bc.filter_mask = k::FILTER_ALL ^ k::FILTER_STATIC;
      bc.friction = 0.6f;
      bc.restitution = 0.3f;
      if (b < (uint32_t)NUM_MESHES) {
        bc.shape_kind = k::SHAPE_MESH;
        bc.mesh_tri_count = w.meshes[b].tri_count;
      } else {
        bc.shape_kind = k::SHAPE_PLANE;
        bc.pose.t = pt[b - NUM_MESHES];
      }
    } else if (b == 20) {  // ball
      float r = mut.ball_radius * k::UU_TO_BT;
      bc.user_idx = k::USER_BALL;
      bc.filter_group = k::FILTER_DEFAULT;
      bc.filter_mask = k::FILTER_ALL;
      bc.mass = 30.0f;
      bc.inv_mass = 1.0f / 30.0f;
      float inertia = 0.4f * bc.mass * r * r;
      bc.inv_inertia_local = splat(1.0f / inertia);
      bc.inv_mass_splat = splat(bc.inv_mass);
      bc.friction = 0.35f;
      bc.restitution = 0.6f;
      bc.linear_damping = 0.03f;
      bc.shape_kind = k::SHAPE_SPHERE;
      bc.sphere_implicit_dim = splat(r);
      bc.sphere_margin = r;
    } else {  // cars
      int ci = (int)b - 21;
      const Preset& pr = PRESETS[ci == 0 ? p0 : p1];
      bc.user_idx = k::USER_CAR;
      bc.user_pointer = (uint32_t)ci;
      bc.filter_group = k::FILTER_DEFAULT;
      bc.filter_mask = k::FILTER_ALL;
      bc.mass = 180.0f;
      bc.inv_mass = 1.0f / 180.0f;
      bc.inv_mass_splat = splat(bc.inv_mass);
      float lx = pr.hx * k::UU_TO_BT, ly = pr.hy * k::UU_TO_BT, lz = pr.hz * k::UU_TO_BT;
      bc.inv_inertia_local = V3{12.0f / (bc.mass * (ly * ly + lz * lz)), 12.0f / (bc.mass * (lx * lx + lz * lz)),
                                12.0f / (bc.mass * (lx * lx + ly * ly))};
      bc.friction = 0.3f;
      bc.restitution = 0.1f;
      bc.shape_kind = k::SHAPE_COMPOUND;
      bc.box_margin = 0.04f;
      bc.box_half_extents = V3{lx * 0.5f, ly * 0.5f, lz * 0.5f} - splat(0.04f);
      bc.child_trans = aff_identity();
      bc.child_trans.t = V3{pr.ox, pr.oy, pr.oz} * k::UU_TO_BT;
      V3 hm = add_s(bc.box_half_extents, bc.box_margin);
      bc.compound_local_aabb = Aabb{bc.child_trans.t - hm, bc.child_trans.t + hm};
      CarConst& cc = w.cars[ci];
      cc.car_idx = (uint32_t)ci;
      cc.team = (uint32_t)ci;
      cc.body_idx = b;
      cc.three_wheels = pr.three_wheels;
      cc.hitbox_size = V3{pr.hx, pr.hy, pr.hz};
      cc.hitbox_offset = V3{pr.ox, pr.oy, pr.oz};
      for (int wi = 0; wi < NUM_WHEELS; ++wi) {  // FL FR BL BR; the left wheels' y negated
        bool front = wi < 2, left = (wi & 1) == 0;
        WheelConst& wc = cc.wheels[wi];
        float cx = front ? pr.fcx : pr.bcx, cyy = front ? pr.fcy : pr.bcy, cz = front ? pr.fcz : pr.bcz;
        if (pr.three_wheels && front) cyy = 0.0f;
        wc.connection_point = V3{cx, left ? -cyy : cyy, cz} * k::UU_TO_BT;
        wc.radius = (front ? pr.fr : pr.br) * k::UU_TO_BT;
        wc.suspension_rest_length = ((front ? pr.frest : pr.brest) - k::MAX_SUSPENSION_TRAVEL) * k::UU_TO_BT;
        wc.suspension_force_scale = front ? 36.0f - 0.25f : 54.0f + 0.25f + 0.015f;
        wc.ray_length = wc.suspension_rest_length + wc.radius;
      }
    }
  }
  // proxies: handle 0 is the sweep-and-prune sentinel, handle b + 1 is body b
  for (uint32_t h = 1; h < (uint32_t)NUM_HANDLES; ++h) {
    ProxyConst& p = w.proxies[h];
    const BodyConst& bc = w.bodies[h - 1];
    p.client_body = h - 1;
    p.filter_group = bc.filter_group;
    p.filter_mask = bc.filter_mask;
    p.unique_id = h;
    if (h - 1 < (uint32_t)(NUM_MESHES + NUM_PLANES)) {
      p.aabb = static_box(w, h - 1);
      cell_indices(w, p.aabb.min, p.indices);
      p.cell_idx = cell_index(w, p.indices);
    }
  }
  // static cell lists: every static handle whose box comes within 2 BT of the cell
  uint32_t ncell = w.grid_num_cells[0] * w.grid_num_cells[1] * w.grid_num_cells[2], at = 0;
  for (uint32_t x = 0; x < w.grid_num_cells[0]; ++x)
    for (uint32_t y = 0; y < w.grid_num_cells[1]; ++y)
      for (uint32_t z = 0; z < w.grid_num_cells[2]; ++z) {
        uint32_t c3[3] = {x, y, z}, c = cell_index(w, c3);
        V3 lo = w.grid_min + V3{(float)x, (float)y, (float)z} * w.grid_cell_size - splat(2.0f);
        V3 hi = lo + splat(w.grid_cell_size + 4.0f);
        w.cells[c].first = at;
        for (uint32_t h = 1; h <= (uint32_t)(NUM_MESHES + NUM_PLANES); ++h) {
          if (!overlaps(w.proxies[h].aabb, Aabb{lo, hi})) continue;
          if (at >= (uint32_t)MAX_CELL_ENTRIES) {
            fprintf(stderr, "world: static cell lists overflow\n");
            abort();
          }
          w.cell_handles[at++] = h;
        }
        w.cells[c].count = at - w.cells[c].first;
      }
  (void)ncell;
  // pads
  for (int i = 0; i < NUM_PADS; ++i) {
    PadConst& p = w.pads[i];
    p.pos = V3{PAD_POS[i][0], PAD_POS[i][1], PAD_POS[i][2]};
    p.is_big = PAD_POS[i][2] > 71.0f ? 1u : 0u;
    p.box_radius = p.is_big ? 208.0f : 144.0f;
    p.cyl_radius = p.is_big ? 208.0f : 144.0f;
    p.max_cooldown = p.is_big ? 10.0f : 4.0f;
    p.boost_amount = p.is_big ? 100.0f : 12.0f;
    V3 e = V3{p.cyl_radius, p.cyl_radius, k::CYL_HEIGHT};
    p.aabb = Aabb{p.pos - e, p.pos + e};
  }
  w.max_pad_z = 73.0f + k::CYL_HEIGHT;
  for (int i = 0; i < TABLE_SIN_LEN; ++i) w.table_sin[i] = (float)sin((double)i * 6.283185307179586476925286766559 / 16384.0);
}

// ---- a start state ---------------------------------------------------------------------------------------------------------------
void build_arena(ArenaState* sp, const StaticData& w, const CarInit cars[NUM_CARS], const BallInit& ball, const bool* pad_active,
                 uint64_t tick_count, uint64_t rng_seed) {
  ArenaState& s = *sp;
  memset(&s, 0, sizeof(ArenaState));
  s.tick_count = u64w(tick_count);
  s.rng = u64w(rng_seed);
  const float dt = 1.0f / 120.0f;
  // dynamic bodies: the ball, then the cars
  for (int d = 0; d < NUM_DYN; ++d) {
    Body& b = s.bodies[d];
    const BodyConst& bc = w.bodies[w.dyn_body[d]];
    M3 rot;
    V3 pos, vel, av;
    if (d == 0) {
      rot = m3_identity();
      pos = ball.pos;
      vel = ball.vel;
      av = ball.ang_vel;
      b.collision_flags = k::CF_CUSTOM_MATERIAL | k::CF_CAN_SLEEP;
    } else {
      const CarInit& c = cars[d - 1];
      rot = rot_from_ypr(c.yaw, c.pitch, c.roll);
      pos = c.pos;
      vel = c.vel;
      av = c.ang_vel;
      b.collision_flags = k::CF_CUSTOM_MATERIAL | (c.demoed ? k::CF_NO_CONTACT_RESPONSE : 0u);
      b.activation = c.demoed ? k::ACT_DISABLED : k::ACT_ACTIVE;
    }
    set_world_trans(b, Aff{rot, pos * k::UU_TO_BT});
    b.interp_world = b.world;
    b.lin_vel = vel * k::UU_TO_BT;
    b.ang_vel = av;
    update_inertia(b, bc);
    if (bc.linear_damping != 0.0f) {
      b.cached_linear_damping = fpow(1.0f - bc.linear_damping, dt);
      b.cached_damping_time_step = dt;
      b.cached_linear_damping_value = bc.linear_damping;
    }
  }
  s.ball.pos = s.bodies[0].world.t * k::BT_TO_UU;
  s.ball.rot = s.bodies[0].world.m;
  s.ball.vel = s.bodies[0].lin_vel * k::BT_TO_UU;
  s.ball.ang_vel = s.bodies[0].ang_vel;
  for (int ci = 0; ci < NUM_CARS; ++ci) {
    Car& car = s.cars[ci];
    const Body& b = s.bodies[1 + ci];
    car.dodge_deadzone = 0.5f;
    car.pos = b.world.t * k::BT_TO_UU;
    car.rot = b.world.m;
    car.vel = b.lin_vel * k::BT_TO_UU;
    car.ang_vel = b.ang_vel;
    car.boost = cars[ci].boost;
    car.is_on_ground = cars[ci].on_ground ? 1u : 0u;
    car.is_demoed = cars[ci].demoed ? 1u : 0u;
    car.demo_respawn_timer = cars[ci].demoed ? cars[ci].demo_timer : 0.0f;
    car.air_time = cars[ci].on_ground ? 0.0f : 0.5f;
    for (int wi = 0; wi < NUM_WHEELS; ++wi) {
      car.wheels[wi].lat_friction = 1.0f;
      car.wheels[wi].long_friction = 1.0f;
    }
  }
  for (int i = 0; i < NUM_PADS; ++i)
    if (pad_active && !pad_active[i]) {  // used a moment ago: the cooldown still runs
      s.pad_gave_boost[i].some = 1;
      uint64_t ago = 1 + (rng_seed >> (i % 32)) % 300;
      s.pad_gave_boost[i].v = u64w(tick_count > ago ? tick_count - ago : 0);
    }
  // broadphase: insert every handle in body order (Bullet AxisSweep3::addHandle), then the dynamic grid cells
  for (int ax = 0; ax < 3; ++ax) {
    s.edges[ax][0] = Edge{0, 0};
    s.edges[ax][1] = Edge{(uint16_t)w.sap_sentinel, 0};
    s.handle_edges[0].min[ax] = 0;
    s.handle_edges[0].max[ax] = 1;
  }
  const V3 cbt = splat(k::CONTACT_BREAKING_THRESHOLD);
  for (uint32_t h = 1; h < (uint32_t)NUM_HANDLES; ++h) {
    uint32_t body = h - 1;
    Aabb box;
    int d = dyn_slot(w, body);
    if (d < 0) {
      box = w.proxies[h].aabb;
    } else {
      const BodyConst& bc = w.bodies[body];
      const Body& b = s.bodies[d];
      box = bc.shape_kind == k::SHAPE_SPHERE ? sphere_box(bc.sphere_margin, b.world.t) : compound_box(bc, b.world);
      box.min -= cbt;
      box.max += cbt;
      s.dyn_proxy[d].aabb = box;
    }
    sap_add_handle(s, w, h, box, h);
  }
  for (int d = 0; d < NUM_DYN; ++d) {
    uint32_t ni[3];
    cell_indices(w, s.dyn_proxy[d].aabb.min, ni);
    s.dyn_proxy[d].cell_idx = cell_index(w, ni);
    for (int a = 0; a < 3; ++a) s.dyn_proxy[d].indices[a] = ni[a];
    update_cells(s, w, true, w.dyn_handle[d], ni);
  }
  s.n_active = 0;
}

// ---- scenario samplers ------------------------------------------------------------------------------------------------------------
namespace {
const float KICKOFF[5][3] = {{-2048, -2560, 0.25f}, {2048, -2560, 0.75f}, {-256, -3840, 0.5f}, {256, -3840, 0.5f}, {0, -4608, 0.5f}};
CarInit ground_car(V3 pos, float yaw, float speed, float boost) {
  CarInit c = CarInit{};
  c.pos = V3{pos.x, pos.y, 17.0f};
  c.yaw = yaw;
  c.vel = V3{(float)(cos(yaw) * speed), (float)(sin(yaw) * speed), 0.0f};
  c.boost = boost;
  c.on_ground = true;
  return c;
}
}  // namespace

Controls sample_controls(Rng& rng) {
  Controls c;
  float ax[5];
  for (int i = 0; i < 5; ++i) ax[i] = (float)(int)(rng.next() % 3) - 1.0f;
  c.throttle = ax[0];
  c.steer = ax[1];
  c.pitch = ax[2];
  c.yaw = ax[3];
  c.roll = ax[4];
  c.jump = rng.unit() < 0.3 ? 1u : 0u;
  c.boost = rng.unit() < 0.3 ? 1u : 0u;
  c.handbrake = rng.unit() < 0.3 ? 1u : 0u;
  return c;
}

void sample_arena(ArenaState* s, const StaticData& w, int kind, Rng& rng) {
  CarInit cars[2];
  BallInit ball = BallInit{};
  bool pads[NUM_PADS];
  for (int i = 0; i < NUM_PADS; ++i) pads[i] = true;
  const float R = w.mut_f[MF_BALL_RADIUS];
  const double TAU = 6.283185307179586;
  switch (kind) {
    case SCEN_KICKOFF: {
      int spot = (int)(rng.next() % 5);
      float yaw = KICKOFF[spot][2] * (float)PI_D;
      cars[0] = ground_car(V3{KICKOFF[spot][0], KICKOFF[spot][1], 0}, yaw, 0.0f, 100.0f / 3.0f);
      cars[1] = ground_car(V3{-KICKOFF[spot][0], -KICKOFF[spot][1], 0}, yaw + (float)PI_D, 0.0f, 100.0f / 3.0f);
      ball.pos = V3{0.0f, 0.0f, R + 1.9f};
      break;
    }
    case SCEN_OPEN: {
      for (int ci = 0; ci < 2; ++ci) {
        V3 p = V3{rng.uniform(-3800, 3800), rng.uniform(-4800, 4800), 0};
        float yaw = (float)(rng.unit() * TAU);
        cars[ci] = ground_car(p, yaw, rng.uniform(0, 2300), rng.uniform(0, 100));
        if (rng.unit() < 0.2) {
          cars[ci].on_ground = false;
          cars[ci].pos.z = rng.uniform(100, 1500);
          cars[ci].pitch = rng.uniform(-1.0f, 1.0f);
          cars[ci].roll = rng.uniform(-1.0f, 1.0f);
          cars[ci].vel.z = rng.uniform(-500, 500);
          cars[ci].ang_vel = V3{rng.uniform(-3, 3), rng.uniform(-3, 3), rng.uniform(-3, 3)};
        }
      }
      ball.pos = V3{rng.uniform(-3800, 3800), rng.uniform(-4800, 4800), rng.uniform(93, 1800)};
      V3 v = V3{(float)(rng.gauss() * 1500), (float)(rng.gauss() * 1500), (float)(rng.gauss() * 1500)};
      float sp = length(v);
      if (sp > 6000.0f) v = v * (6000.0f / sp);
      ball.vel = v;
      for (int i = 0; i < NUM_PADS; ++i) pads[i] = rng.unit() < 0.5;
      break;
    }
    case SCEN_PINCH: {  // the ball within 200 uu of a wall or fillet, a car driving into it
      int where = (int)(rng.next() % 4);
      float gap = rng.uniform(0, 200);
      V3 bp, inward;
      if (where == 0 || where == 1) {  // side wall x = +-AX
        float sx = where == 0 ? 1.0f : -1.0f;
        bp = V3{sx * (float)(AX - R - gap), rng.uniform(-3800, 3800), rng.uniform(R, 400)};
        inward = V3{-sx, 0, 0};
      } else if (where == 2) {  // back wall beside the goal
        float sy = rng.unit() < 0.5 ? 1.0f : -1.0f, sx = rng.unit() < 0.5 ? 1.0f : -1.0f;
        bp = V3{sx * rng.uniform(1100, 2900), sy * (float)(AY - R - gap), rng.uniform(R, 400)};
        inward = V3{0, -sy, 0};
      } else {  // a corner
        int q = (int)(rng.next() % 4);
        double th = (q + rng.unit()) * PI_D * 0.5;
        double cx = (q == 0 || q == 3) ? AX - CR : -(AX - CR), cy = q < 2 ? AY - CR : -(AY - CR);
        double rr = CR - R - gap;
        bp = V3{(float)(cx + cos(th) * rr), (float)(cy + sin(th) * rr), rng.uniform(R, 400)};
        inward = V3{(float)-cos(th), (float)-sin(th), 0};
      }
      ball.pos = bp;
      ball.vel = V3{rng.uniform(-300, 300), rng.uniform(-300, 300), 0};
      float dist = rng.uniform(300, 900), ang = rng.uniform(-0.6f, 0.6f);
      float yaw_in = atan2f(inward.y, inward.x) + ang;
      V3 cp = V3{bp.x + cosf(yaw_in) * dist, bp.y + sinf(yaw_in) * dist, 0};
      cars[0] = ground_car(cp, yaw_in + (float)PI_D, rng.uniform(1400, 2300), rng.uniform(0, 100));
      V3 p1 = V3{rng.uniform(-3500, 3500), rng.uniform(-4500, 4500), 0};
      cars[1] = ground_car(p1, (float)(rng.unit() * TAU), rng.uniform(0, 1500), rng.uniform(0, 100));
      break;
    }
    case SCEN_HEADON: {  // two cars closing on each other
      V3 mid = V3{rng.uniform(-3000, 3000), rng.uniform(-4000, 4000), 0};
      float yaw = (float)(rng.unit() * TAU), gap = rng.uniform(300, 1200), closing = rng.uniform(1500, 2300);
      float f0 = rng.uniform(0.3f, 0.7f);
      V3 dir = V3{cosf(yaw), sinf(yaw), 0};
      cars[0] = ground_car(mid - dir * (gap * 0.5f + 120.0f), yaw, closing * f0, rng.uniform(0, 100));
      cars[1] = ground_car(mid + dir * (gap * 0.5f + 120.0f), yaw + (float)PI_D, closing * (1.0f - f0), rng.uniform(0, 100));
      ball.pos = V3{rng.uniform(-3800, 3800), rng.uniform(-4800, 4800), rng.uniform(93, 600)};
      break;
    }
    default: {  // SCEN_GOALFRAME: the ball near a post or the crossbar, moving at it
      float sy = rng.unit() < 0.5 ? 1.0f : -1.0f;
      float d = rng.uniform(0, 150);
      V3 target;
      if (rng.unit() < 0.5) {  // a post
        float sx = rng.unit() < 0.5 ? 1.0f : -1.0f;
        target = V3{sx * (float)GX, sy * (float)AY, rng.uniform(R, (float)GH - R)};
        ball.pos = V3{target.x - sx * (R + d) * 0.5f, target.y - sy * (R + d), target.z};
      } else {  // the crossbar
        target = V3{rng.uniform(-800, 800), sy * (float)AY, (float)GH};
        ball.pos = V3{target.x, target.y - sy * (R + d), target.z - (R + d) * 0.5f};
      }
      V3 to = target - ball.pos;
      float sp = rng.uniform(500, 3000);
      ball.vel = normalize_or_zero(to) * sp;
      for (int ci = 0; ci < 2; ++ci)
        cars[ci] = ground_car(V3{rng.uniform(-3000, 3000), sy * rng.uniform(2000, 4500), 0}, (float)(rng.unit() * TAU),
                              rng.uniform(0, 2000), rng.uniform(0, 100));
      break;
    }
  }
  build_arena(s, w, cars, ball, pads, 1000 + (rng.next() % 100000), rng.next());
}

}  // namespace world
}  // namespace rs

// ---- C ABI of the builder (declared in include/rsim.h) ------------------------------------------------------------------------------
namespace {
uint64_t arena_seed(uint64_t seed, uint32_t i) {  // an independent SplitMix64 stream per arena index
  rs::world::Rng r{seed ^ (0xd1b54a32d192ed03ull * ((uint64_t)i + 1))};
  r.next();
  return r.next();
}
int scenario_of(int scenario, uint64_t pick) {
  switch (scenario) {
    case RSIM_SCEN_CONTACT: {  // 40% pinches, 30% head-on, 30% goal frame
      uint32_t p = (uint32_t)(pick % 10);
      return p < 4 ? rs::world::SCEN_PINCH : p < 7 ? rs::world::SCEN_HEADON : rs::world::SCEN_GOALFRAME;
    }
    case RSIM_SCEN_MIXED: {  // 1:1:1 kickoff, open play, contact-heavy
      uint32_t p = (uint32_t)(pick % 3);
      return p == 0 ? rs::world::SCEN_KICKOFF : p == 1 ? rs::world::SCEN_OPEN : scenario_of(RSIM_SCEN_CONTACT, pick / 3);
    }
    case RSIM_SCEN_OPEN_CONTACT:  // 1:1 open play, contact-heavy
      return (pick & 1) == 0 ? rs::world::SCEN_OPEN : scenario_of(RSIM_SCEN_CONTACT, pick >> 1);
    default:
      return scenario;
  }
}
}  // namespace

extern "C" {

int rsim_build_world(int preset0, int preset1, float gravity_z, float boost_accel_scale, float ball_radius, void* world_out) {
  if (!world_out || preset0 < 0 || preset0 >= rs::world::NUM_PRESETS || preset1 < 0 || preset1 >= rs::world::NUM_PRESETS)
    return RSIM_ERR_INVALID;
  if (!(gravity_z < 0.0f) || !(boost_accel_scale > 0.0f) || !(ball_radius > 0.0f)) return RSIM_ERR_INVALID;
  rs::world::Mutators m;
  m.gravity_z = gravity_z;
  m.boost_accel_scale = boost_accel_scale;
  m.ball_radius = ball_radius;
  rs::world::build_world((rs::StaticData*)world_out, preset0, preset1, m);
  return RSIM_OK;
}

int rsim_sample_arenas(const void* world, int scenario, uint64_t seed, uint32_t first, uint32_t n, void* states_out) {
  if (!world || !states_out || scenario < 0 || scenario > RSIM_SCEN_OPEN_CONTACT) return RSIM_ERR_INVALID;
  const rs::StaticData& w = *(const rs::StaticData*)world;
  rs::ArenaState* out = (rs::ArenaState*)states_out;
  for (uint32_t i = 0; i < n; ++i) {
    rs::world::Rng r{arena_seed(seed, first + i)};
    int kind = scenario_of(scenario, r.next());
    rs::world::sample_arena(&out[i], w, kind, r);
  }
  return RSIM_OK;
}

int rsim_sample_arenas_in(const void* worlds, uint32_t count, const uint32_t* variant, int scenario, uint64_t seed,
                          uint32_t first, uint32_t n, void* states_out) {
  if (!worlds || !variant || !states_out || scenario < 0 || scenario > RSIM_SCEN_OPEN_CONTACT) return RSIM_ERR_INVALID;
  for (uint32_t i = 0; i < n; ++i)
    if (variant[i] >= count) return RSIM_ERR_INVALID;
  const rs::StaticData* w = (const rs::StaticData*)worlds;
  rs::ArenaState* out = (rs::ArenaState*)states_out;
  for (uint32_t i = 0; i < n; ++i) {  // the same draws as rsim_sample_arenas, each in its own world
    rs::world::Rng r{arena_seed(seed, first + i)};
    int kind = scenario_of(scenario, r.next());
    rs::world::sample_arena(&out[i], w[variant[i]], kind, r);
  }
  return RSIM_OK;
}

int rsim_sample_controls(uint64_t seed, uint32_t decision, uint32_t first, uint32_t n, void* controls_out) {
  if (!controls_out) return RSIM_ERR_INVALID;
  rs::Controls* out = (rs::Controls*)controls_out;
  for (uint32_t i = 0; i < n; ++i) {
    rs::world::Rng r{arena_seed(seed ^ (0x9e3779b97f4a7c15ull * ((uint64_t)decision + 1)), first + i)};
    for (int c = 0; c < rs::NUM_CARS; ++c) out[(size_t)i * rs::NUM_CARS + c] = rs::world::sample_controls(r);
  }
  return RSIM_OK;
}

}  // extern "C"
