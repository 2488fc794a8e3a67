// Phases P4 (islands: union-find over the dynamic pairs, sleep, the island sort of the active manifolds) and P5 (the
// sequential-impulse solver: contact and friction rows, special synthetic rows for the ball's static contacts, 10 split-impulse
// iterations, up to 10 velocity iterations with a zero-residual exit, finish). Reference: cuda/sim/simulation_island_manager.cuh,
// seq_impulse_constraint_solver.cuh, solver_constraint.cuh.
#pragma once
#include "rs_narrow.h"

namespace rs {

// ---- P4 ----------------------------------------------------------------------------------------------------------------------
RS_FN int island_root(int* parent, int i) {  // find with full path compression (:29-39)
  int r = i;
  while (parent[r] != r) r = parent[r];
  while (parent[i] != r && i != r) {
    int nx = parent[i];
    parent[i] = r;
    i = nx;
  }
  return r;
}
// Hoare partition sort of `idx` by keys[idx] (:43-66, the recursive form's visiting order on an explicit stack)
RS_FN void hoare_sort(uint32_t* idx, int n, const int* keys) {
  if (n <= 1) return;
  int slo[32], shi[32], sp = 0;
  slo[0] = 0;
  shi[0] = n - 1;
  sp = 1;
  while (sp > 0) {
    --sp;
    int lo = slo[sp], hi = shi[sp], i = lo, j = hi;
    int pivot = keys[idx[(lo + hi) / 2]];
    for (;;) {
      while (keys[idx[i]] < pivot) ++i;
      while (pivot < keys[idx[j]]) --j;
      if (i > j) break;
      uint32_t t = idx[i];
      idx[i] = idx[j];
      idx[j] = t;
      ++i;
      if (j == 0) break;
      --j;
      if (i > j) break;
    }
    if (i < hi) { slo[sp] = i; shi[sp] = hi; ++sp; }
    if (lo < j) { slo[sp] = lo; shi[sp] = j; ++sp; }
  }
}
RS_INLINE void phase_p4(ArenaState& s, const StaticData& w) {  // :100-150
  const int n = (int)s.n_active;
  int parent[NUM_DYN];
  for (int d = 0; d < NUM_DYN; ++d) parent[d] = d;
  for (uint32_t p = 0; p < s.n_pairs; ++p) {
    const ProxyConst &x0 = w.proxies[s.pairs[p].proxy0], &x1 = w.proxies[s.pairs[p].proxy1];
    if ((x0.filter_group & k::FILTER_STATIC) != 0 || (x1.filter_group & k::FILTER_STATIC) != 0) continue;
    int a = dyn_slot(w, x0.client_body), b = dyn_slot(w, x1.client_body);
    if (a < 0 || b < 0) continue;
    if ((s.bodies[a].collision_flags & k::CF_NO_CONTACT_RESPONSE) || (s.bodies[b].collision_flags & k::CF_NO_CONTACT_RESPONSE)) continue;
    int r0 = island_root(parent, a), r1 = island_root(parent, b);
    if (r0 != r1) parent[r0] = r1;
  }
  int tag[NUM_DYN];
  for (int d = 0; d < NUM_DYN; ++d) tag[d] = island_root(parent, d);
  bool awake[NUM_DYN] = {false, false, false};
  for (int d = 0; d < NUM_DYN; ++d)
    if (s.bodies[d].activation == k::ACT_ACTIVE) awake[tag[d]] = true;
  for (int d = 0; d < NUM_DYN; ++d)
    if (!awake[tag[d]] && s.bodies[d].activation == k::ACT_WANTS_DEACTIVATION) body_sleep(s.bodies[d]);
  if (n <= 1) return;
  uint32_t idx[MAX_ACTIVE];
  for (int i = 0; i < n; ++i) idx[i] = s.active[i];
  for (int i = 1; i < n; ++i) {  // ascending (distinct indices)
    uint32_t v = idx[i];
    int j = i - 1;
    while (j >= 0 && idx[j] > v) {
      idx[j + 1] = idx[j];
      --j;
    }
    idx[j + 1] = v;
  }
  int keys[MAX_MANIFOLDS];
  for (int i = 0; i < n; ++i) {
    const Manifold& m = s.manifolds[idx[i]];
    int d0 = dyn_slot(w, m.body0), d1 = dyn_slot(w, m.body1);
    int t0 = d0 < 0 ? -1 : tag[d0], t1 = d1 < 0 ? -1 : tag[d1];
    keys[idx[i]] = t0 >= 0 ? t0 : t1;
  }
  hoare_sort(idx, n, keys);
  for (int i = 0; i < n; ++i) s.active[i] = idx[i];
}

// ---- P5 rows (solver_constraint.cuh) -----------------------------------------------------------------------------------------
constexpr int MAX_ROWS = 32;
constexpr int MAX_SOLVER_BODIES = 8;
constexpr int SOLVER_ITERATIONS = 10;
constexpr float ERP2 = 0.8f;
constexpr float SPLIT_TURN_ERP = 0.1f;

struct SolverBody {
  V3 dlin, dang, inv_mass, push, turn, lin, ang, ext_force, ext_torque;
  int has_orig, orig;
};
struct Row {
  V3 r1xn, n1, r2xn, n2, ang_a, ang_b;
  float applied_push, applied, friction, jinv, rhs, lower, upper, rhs_pen;
  int friction_idx, a, b, split_only;
};
struct RowPoint { V3 normal; float friction, restitution, distance; };
struct Side { const Body* rb; const BodyConst* c; };  // the dynamic body behind a row side, or none

RS_FN SolverBody sbody_fixed() {
  V3 z = splat(0.0f);
  return SolverBody{z, z, z, z, z, z, z, z, z, 0, 0};
}
RS_FN SolverBody sbody_of(const Body& rb, const BodyConst& c, int d) {
  V3 z = splat(0.0f);
  return SolverBody{z, z, c.inv_mass_splat, z, z, rb.lin_vel, rb.ang_vel, rb.accum_lin_vel, rb.accum_ang_vel, 1, d};
}
RS_FN V3 sbody_vel(const SolverBody& b, V3 rel) { return b.lin + b.ext_force + cross(b.ang + b.ext_torque, rel); }
RS_FN float bullet_dot(V3 a, V3 b) {  // x + (y + z)
  V3 r = a * b;
  return r.x + (r.y + r.z);
}
RS_FN V3 plane_space(V3 n) {  // util.rs plane_space_1
  if (fabs_(n.z) > 0.70710678118654752440f) {
    float kk = 1.0f / __builtin_sqrtf(n.y * n.y + n.z * n.z);
    return V3{0.0f, fneg(n.z) * kk, n.y * kk};
  }
  float kk = 1.0f / __builtin_sqrtf(n.x * n.x + n.y * n.y);
  return V3{fneg(n.y) * kk, n.x * kk, 0.0f};
}
RS_FN V3 lat_friction_dir(const RowPoint& cp, const SolverBody& a, const SolverBody& b, V3 r1, V3 r2) {
  V3 v = sbody_vel(a, r1) - sbody_vel(b, r2);
  float rv = dot(cp.normal, v);
  V3 lat = v - cp.normal * rv;
  float l2 = length_sq(lat);
  if (l2 > 1.1920929e-7f) return lat * (1.0f / __builtin_sqrtf(l2));
  return plane_space(cp.normal);
}
RS_FN Row row_new(int a, int b, int fidx, float friction, float lower, float upper) {
  Row r;
  V3 z = splat(0.0f);
  r.r1xn = r.n1 = r.r2xn = r.n2 = r.ang_a = r.ang_b = z;
  r.applied_push = r.applied = r.jinv = r.rhs = r.rhs_pen = 0.0f;
  r.friction = friction;
  r.lower = lower;
  r.upper = upper;
  r.friction_idx = fidx;
  r.a = a;
  r.b = b;
  r.split_only = 0;
  return r;
}
RS_FN Row contact_row(int ia, int ib, const SolverBody& sa, const SolverBody& sb, Side s0, Side s1, V3 r1, V3 r2, const RowPoint& cp,
                      int fidx, float dt) {  // get_contact_constraint (warm start is the identity: :205-214)
  Row r = row_new(ia, ib, fidx, cp.friction, 0.0f, 1e10f);
  float inv_dt = 1.0f / dt;
  V3 n = cp.normal;
  float den0 = 0.0f, v1n = 0.0f, den1 = 0.0f, v2n = 0.0f;
  V3 vel0 = splat(0.0f), vel1 = splat(0.0f);
  if (s0.rb) {
    V3 ta = cross(r1, n);
    r.ang_a = mul_t(s0.rb->inv_inertia_world, ta);
    den0 = s0.c->inv_mass + dot(n, cross(r.ang_a, r1));
    vel0 = s0.rb->lin_vel + cross(s0.rb->ang_vel, r1);
    r.n1 = n;
    r.r1xn = ta;
    v1n = dot(r.n1, sa.lin + sa.ext_force) + dot(r.r1xn, sa.ang + sa.ext_torque);
  }
  if (s1.rb) {
    V3 ta = cross(r2, n);
    r.ang_b = mul_t(s1.rb->inv_inertia_world, -ta);
    den1 = s1.c->inv_mass + dot(n, cross(-r.ang_b, r2));
    vel1 = s1.rb->lin_vel + cross(s1.rb->ang_vel, r2);
    r.n2 = -n;
    r.r2xn = -ta;
    v2n = dot(r.n2, sb.lin + sb.ext_force) + dot(r.r2xn, sb.ang + sb.ext_torque);
  }
  r.jinv = 1.0f / (den0 + den1);
  float rel_vel = dot(n, vel0 - vel1);
  float restitution = fmax_rs(cp.restitution * fneg(rel_vel), 0.0f);
  float pen = cp.distance;
  float pos_err = pen > 0.0f ? 0.0f : (fneg(pen) * ERP2) * inv_dt;
  float vel_err = restitution - (v1n + v2n);
  float pen_imp = pos_err * r.jinv, vel_imp = vel_err * r.jinv;
  if (pen > 1e30f) {  // SPLIT_IMPULSE_PENETRATION_THRESHOLD
    r.rhs = pen_imp + vel_imp;
    r.rhs_pen = 0.0f;
  } else {
    r.rhs = vel_imp;
    r.rhs_pen = pen_imp;
  }
  return r;
}
RS_FN Row split_only_row(int ia, int ib, Side s0, Side s1, V3 r1, V3 r2, const RowPoint& cp, float dt) {  // :100-172
  Row r = row_new(ia, ib, -1, cp.friction, 0.0f, 1e10f);
  r.split_only = 1;
  V3 n = cp.normal;
  float den0 = 0.0f, den1 = 0.0f;
  if (s0.rb) {
    V3 ta = cross(r1, n);
    r.ang_a = mul_t(s0.rb->inv_inertia_world, ta);
    den0 = s0.c->inv_mass + dot(n, cross(r.ang_a, r1));
    r.n1 = n;
    r.r1xn = ta;
  }
  if (s1.rb) {
    V3 ta = cross(r2, n);
    r.ang_b = mul_t(s1.rb->inv_inertia_world, -ta);
    den1 = s1.c->inv_mass + dot(n, cross(-r.ang_b, r2));
    r.n2 = -n;
    r.r2xn = -ta;
  }
  r.jinv = 1.0f / (den0 + den1);
  float inv_dt = 1.0f / dt, pen = cp.distance;
  float pos_err = pen > 0.0f ? 0.0f : (fneg(pen) * ERP2) * inv_dt;
  float pen_imp = pos_err * r.jinv;
  r.rhs_pen = pen > 1e30f ? 0.0f : pen_imp;
  return r;
}
RS_FN Row friction_row(int ia, int ib, const SolverBody& sa, const SolverBody& sb, Side s0, Side s1, V3 r1, V3 r2, float friction,
                       V3 axis, int fidx) {  // get_friction_constraint (:39-67, :291-340)
  Row r = row_new(ia, ib, fidx, friction, fneg(friction), friction);
  float v1n = 0.0f, den1 = 0.0f, v2n = 0.0f, den2 = 0.0f;
  if (s0.rb) {
    r.n1 = axis;
    r.r1xn = cross(r1, r.n1);
    r.ang_a = mul_t(s0.rb->inv_inertia_world, r.r1xn);
    den1 = s0.c->inv_mass + dot(axis, cross(r.ang_a, r1));
    v1n = dot(r.n1, sa.lin + sa.ext_force) + dot(r.r1xn, sa.ang);
  }
  if (s1.rb) {
    r.n2 = -axis;
    r.r2xn = cross(r2, r.n2);
    r.ang_b = mul_t(s1.rb->inv_inertia_world, r.r2xn);
    den2 = s1.c->inv_mass + dot(axis, cross(-r.ang_b, r2));
    v2n = dot(r.n2, sb.lin + sb.ext_force) + dot(r.r2xn, sb.ang);
  }
  r.jinv = 1.0f / (den1 + den2);
  r.rhs = fneg(v1n + v2n) * r.jinv;
  return r;
}
// The row solves read and write the bodies' deltas through the bank (fields: 0 dlin, 1 dang, 2 push, 3 turn); the
// read-only body constants (inverse mass, whether it is a dynamic body) come from the SolverBody.
enum BankField { BF_DLIN = 0, BF_DANG = 1, BF_PUSH = 2, BF_TURN = 3, BF_INV_MASS = 4 };
// The row solves read the bodies' inverse masses and kinds from the bank (a body that is not dynamic takes no delta).
RS_FN float resolve_split(Row& r, DeltaBank& bank) {  // :518-565
  if (r.rhs_pen == 0.0f) return 0.0f;
  const bool ha = bank.dynamic(r.a), hb = bank.dynamic(r.b);
  float di = r.rhs_pen;
  V3 pa = ha ? bank.get(r.a, BF_PUSH) : splat(0.0f), ta = ha ? bank.get(r.a, BF_TURN) : splat(0.0f);
  V3 pb = hb ? bank.get(r.b, BF_PUSH) : splat(0.0f), tb = hb ? bank.get(r.b, BF_TURN) : splat(0.0f);
  float d1 = ha ? bullet_dot(r.n1, pa) + bullet_dot(r.r1xn, ta) : 0.0f;
  float d2 = hb ? bullet_dot(r.n2, pb) + bullet_dot(r.r2xn, tb) : 0.0f;
  di = di - d1 * r.jinv;
  di = di - d2 * r.jinv;
  float sum = r.applied_push + di;
  if (sum < r.lower) {
    di = r.lower - r.applied_push;
    r.applied_push = r.lower;
  } else {
    r.applied_push = sum;
  }
  if (ha) {
    bank.set(r.a, BF_PUSH, pa + r.n1 * bank.get(r.a, BF_INV_MASS) * di);
    bank.set(r.a, BF_TURN, ta + r.ang_a * di);
  }
  if (hb) {  // re-read: a and b may be the same body
    bank.set(r.b, BF_PUSH, bank.get(r.b, BF_PUSH) + r.n2 * bank.get(r.b, BF_INV_MASS) * di);
    bank.set(r.b, BF_TURN, bank.get(r.b, BF_TURN) + r.ang_b * di);
  }
  return di / r.jinv;
}
RS_FN float resolve_row(Row& r, DeltaBank& bank, bool two_sided) {  // :427-516
  const bool ha = bank.dynamic(r.a), hb = bank.dynamic(r.b);
  float di = r.rhs;
  V3 la = ha ? bank.get(r.a, BF_DLIN) : splat(0.0f), aa = ha ? bank.get(r.a, BF_DANG) : splat(0.0f);
  V3 lb = hb ? bank.get(r.b, BF_DLIN) : splat(0.0f), ab = hb ? bank.get(r.b, BF_DANG) : splat(0.0f);
  float d1 = ha ? bullet_dot(r.n1, la) + bullet_dot(r.r1xn, aa) : 0.0f;
  float d2 = hb ? bullet_dot(r.n2, lb) + bullet_dot(r.r2xn, ab) : 0.0f;
  di = di - d1 * r.jinv;
  di = di - d2 * r.jinv;
  float sum = r.applied + di;
  if (sum < r.lower) {
    di = r.lower - r.applied;
    r.applied = r.lower;
  } else if (two_sided && sum > r.upper) {
    di = r.upper - r.applied;
    r.applied = r.upper;
  } else {
    r.applied = sum;
  }
  if (ha) {
    bank.set(r.a, BF_DLIN, la + r.n1 * bank.get(r.a, BF_INV_MASS) * di);
    bank.set(r.a, BF_DANG, aa + r.ang_a * di);
  }
  if (hb) {  // re-read: a and b may be the same body
    bank.set(r.b, BF_DLIN, bank.get(r.b, BF_DLIN) + r.n2 * bank.get(r.b, BF_INV_MASS) * di);
    bank.set(r.b, BF_DANG, bank.get(r.b, BF_DANG) + r.ang_b * di);
  }
  return di / r.jinv;
}

// ---- P5 (seq_impulse_constraint_solver.cuh) -----------------------------------------------------------------------------------
struct SpecialAcc {
  int d;
  V3 normal;
  float lever, friction, restitution;
  uint32_t count;
};
struct Solver {
  SolverBody bodies[MAX_SOLVER_BODIES];
  Row contact[MAX_ROWS];
  Row friction[MAX_ROWS];
  SpecialAcc special[4];
  int n_bodies, n_contact, n_friction, n_special;
  int fixed;
  int companion[NUM_DYN];
};
RS_FN int solver_body_of(Solver& sv, const ArenaState& s, const StaticData& w, int d, uint32_t body) {  // :55-70
  if (d < 0) {
    if (sv.fixed >= 0) return sv.fixed;
    sv.fixed = sv.n_bodies;
    sv.bodies[sv.n_bodies++] = sbody_fixed();
    return sv.fixed;
  }
  if (sv.companion[d] >= 0) return sv.companion[d];
  int id = sv.n_bodies;
  sv.companion[d] = id;
  sv.bodies[sv.n_bodies++] = sbody_of(s.bodies[d], w.bodies[body], d);
  return id;
}
RS_FN void add_special(Solver& sv, int d, V3 normal, float lever, float friction, float restitution) {  // :84-99
  for (int i = 0; i < sv.n_special; ++i) {
    SpecialAcc& a = sv.special[i];
    if (a.d == d) {
      a.normal += normal;
      a.lever = a.lever + lever;
      a.friction = a.friction + friction;
      a.restitution = a.restitution + restitution;
      a.count += 1;
      return;
    }
  }
  sv.special[sv.n_special++] = SpecialAcc{d, normal, lever, friction, restitution, 1u};
}
RS_FN void push_special(Solver& sv, ArenaState& s, const StaticData& w, const SpecialAcc& acc, float dt) {  // :102-129
  float ns = (float)acc.count;
  V3 mn = normalize_or_zero(acc.normal / ns);
  float lever = acc.lever / ns, fr = acc.friction / ns, rest = acc.restitution / ns;
  int d = acc.d;
  int ia = sv.companion[d], ib;
  if (sv.fixed >= 0) {
    ib = sv.fixed;
  } else {
    ib = sv.n_bodies;
    sv.fixed = ib;
    sv.bodies[sv.n_bodies++] = sbody_fixed();
  }
  const SolverBody &sa = sv.bodies[ia], &sb = sv.bodies[ib];
  V3 r1 = mn * fneg(lever), r2 = splat(0.0f);
  RowPoint cp = RowPoint{mn, fr, rest, 0.0f};
  Side s0 = Side{sa.has_orig ? &s.bodies[d] : nullptr, &w.bodies[w.dyn_body[d]]}, s1 = Side{nullptr, nullptr};
  int fidx = sv.n_contact;
  if (sv.n_contact >= MAX_ROWS || sv.n_friction >= MAX_ROWS) {
    s.overflow[OVF_ACTIVE]++;
    return;
  }
  sv.contact[sv.n_contact++] = contact_row(ia, ib, sa, sb, s0, s1, r1, r2, cp, fidx, dt);
  V3 lat = lat_friction_dir(cp, sa, sb, r1, r2);
  sv.friction[sv.n_friction++] = friction_row(ia, ib, sa, sb, s0, s1, r1, r2, cp.friction, lat, fidx);
}
RS_FN void solver_setup(Solver& sv, ArenaState& s, const StaticData& w, float dt) {  // :132-193
  sv.fixed = -1;
  for (int d = 0; d < NUM_DYN; ++d) {  // setup_solver_bodies
    sv.companion[d] = -1;
    if (!body_active(s.bodies[d])) continue;
    sv.companion[d] = sv.n_bodies;
    sv.bodies[sv.n_bodies++] = sbody_of(s.bodies[d], w.bodies[w.dyn_body[d]], d);
  }
  sv.n_special = 0;
  for (uint32_t ai = 0; ai < s.n_active; ++ai) {
    const Manifold& m = s.manifolds[s.active[ai]];
    int d0 = dyn_slot(w, m.body0), d1 = dyn_slot(w, m.body1);
    int ia = solver_body_of(sv, s, w, d0, m.body0), ib = solver_body_of(sv, s, w, d1, m.body1);
    // both candidates loaded, one picked: a conditional address would mix the state's and the world's memory into a
    // generic pointer (flat loads)
    const V3 s0 = s.bodies[d0 >= 0 ? d0 : 0].world.t, s1 = s.bodies[d1 >= 0 ? d1 : 0].world.t;
    const V3 p0 = w.bodies[m.body0].pose.t, p1 = w.bodies[m.body1].pose.t;
    V3 t0 = d0 >= 0 ? s0 : p0;
    V3 t1 = d1 >= 0 ? s1 : p1;
    for (uint32_t pi = 0; pi < m.n_points; ++pi) {
      const ManifoldPoint& mp = m.pts[pi];
      V3 r1 = mp.world_on_a - t0, r2 = mp.world_on_b - t1;
      RowPoint cp = RowPoint{mp.normal_on_b, mp.friction, mp.restitution, mp.distance};
      Side s0 = Side{sv.bodies[ia].has_orig ? &s.bodies[d0] : nullptr, d0 >= 0 ? &w.bodies[m.body0] : nullptr};
      Side s1 = Side{sv.bodies[ib].has_orig ? &s.bodies[d1] : nullptr, d1 >= 0 ? &w.bodies[m.body1] : nullptr};
      if (mp.is_special) {
        if (sv.n_contact >= MAX_ROWS) s.overflow[OVF_ACTIVE]++;
        else sv.contact[sv.n_contact++] = split_only_row(ia, ib, s0, s1, r1, r2, cp, dt);
        int obj = -1;
        V3 lever = splat(0.0f);
        if (d0 >= 0) {
          obj = d0;
          lever = r1;
        } else if (d1 >= 0) {
          obj = d1;
          lever = r2;
        }
        if (obj >= 0) add_special(sv, obj, mp.normal_on_b, length(lever), mp.friction, mp.restitution);
        continue;
      }
      int fidx = sv.n_contact;
      V3 lat = lat_friction_dir(cp, sv.bodies[ia], sv.bodies[ib], r1, r2);
      if (sv.n_contact >= MAX_ROWS || sv.n_friction >= MAX_ROWS) {
        s.overflow[OVF_ACTIVE]++;
        continue;
      }
      sv.contact[sv.n_contact++] = contact_row(ia, ib, sv.bodies[ia], sv.bodies[ib], s0, s1, r1, r2, cp, fidx, dt);
      sv.friction[sv.n_friction++] = friction_row(ia, ib, sv.bodies[ia], sv.bodies[ib], s0, s1, r1, r2, mp.friction, lat, fidx);
    }
  }
  s.n_active = 0;
  for (int i = 0; i < MAX_ACTIVE; ++i) s.active[i] = 0;
  for (int i = 0; i < sv.n_special; ++i) {
    SpecialAcc acc = sv.special[i];
    push_special(sv, s, w, acc, dt);
  }
}
// The iterations (:196-251). Each loop reads its next row while it solves the current one (a row's solve never writes
// another row), keeps the rows still running in a bit mask, and takes the contact impulses a friction row needs from the
// bank; what a solve changes in a row is written back to it.
RS_FN void solver_iterate(Solver& sv) {  // :196-251
  static_assert(MAX_ROWS <= BANK_ROWS && MAX_SOLVER_BODIES <= BANK_BODIES, "the bank holds every row and body");
  static_assert(MAX_ROWS <= 32, "the running rows are a 32-bit mask");
  DeltaBank bank;
  for (int i = 0; i < sv.n_bodies && i < BANK_BODIES; ++i) {
    bank.set(i, BF_DLIN, sv.bodies[i].dlin);
    bank.set(i, BF_DANG, sv.bodies[i].dang);
    bank.set(i, BF_PUSH, sv.bodies[i].push);
    bank.set(i, BF_TURN, sv.bodies[i].turn);
    bank.set(i, BF_INV_MASS, sv.bodies[i].inv_mass);
    bank.set_dynamic(i, sv.bodies[i].has_orig != 0);
  }
  const int nc = sv.n_contact, nf = sv.n_friction;
  for (int i = 0; i < nc; ++i) bank.set_applied(i, sv.contact[i].applied);
  uint32_t run = nc >= 32 ? ~0u : (1u << nc) - 1u;
  for (int it = 0; it < SOLVER_ITERATIONS && run != 0; ++it) {
    Row next = nc > 0 ? sv.contact[0] : Row{};
    for (int i = 0; i < nc; ++i) {
      Row r = next;
      if (i + 1 < nc) next = sv.contact[i + 1];
      if (!((run >> i) & 1u)) continue;
      float res = resolve_split(r, bank);
      sv.contact[i].applied_push = r.applied_push;
      if (res * res == 0.0f) run &= ~(1u << i);
    }
  }
  for (int it = 0; it < SOLVER_ITERATIONS; ++it) {
    float lsr = 0.0f;
    Row next = nc > 0 ? sv.contact[0] : Row{};
    for (int i = 0; i < nc; ++i) {
      Row r = next;
      if (i + 1 < nc) next = sv.contact[i + 1];
      if (r.split_only) continue;
      float res = resolve_row(r, bank, false);
      sv.contact[i].applied = r.applied;
      bank.set_applied(i, r.applied);
      lsr = fmax_rs(res * res, lsr);
    }
    next = nf > 0 ? sv.friction[0] : Row{};
    for (int i = 0; i < nf; ++i) {
      Row r = next;
      if (i + 1 < nf) next = sv.friction[i + 1];
      float total = bank.applied(r.friction_idx);
      if (total <= 0.0f) continue;
      float lim = r.friction * total;
      r.lower = fneg(lim);
      r.upper = lim;
      float res = resolve_row(r, bank, true);
      sv.friction[i].lower = r.lower;
      sv.friction[i].upper = r.upper;
      sv.friction[i].applied = r.applied;
      lsr = fmax_rs(res * res, lsr);
    }
    if (lsr == 0.0f) break;
  }
  for (int i = 0; i < sv.n_bodies && i < BANK_BODIES; ++i) {
    sv.bodies[i].dlin = bank.get(i, BF_DLIN);
    sv.bodies[i].dang = bank.get(i, BF_DANG);
    sv.bodies[i].push = bank.get(i, BF_PUSH);
    sv.bodies[i].turn = bank.get(i, BF_TURN);
  }
}
RS_FN void solver_finish(Solver& sv, ArenaState& s, float dt) {  // :271-301
  for (int i = 0; i < sv.n_bodies; ++i) {
    SolverBody& b = sv.bodies[i];
    if (!b.has_orig) continue;
    Body& body = s.bodies[b.orig];
    b.lin += b.dlin;
    b.ang += b.dang;
    if (length_sq(b.push) != 0.0f || length_sq(b.turn) != 0.0f) {
      Aff t = body.world;
      if (body.collision_flags & k::CF_NO_ANGULAR_MOTION) {
        t.t += b.push * dt;
      } else {
        Quat q = body.world_quat;
        integrate_trans(&t, &q, b.push, b.turn * SPLIT_TURN_ERP, dt);
      }
      set_world_trans(body, t);
    }
    V3 lv = b.lin + b.ext_force, av = b.ang + b.ext_torque;
    bool was_sleeping = body.activation == k::ACT_SLEEPING;
    body.lin_vel = lv;
    body.ang_vel = av;
    if (was_sleeping && (lv.x != 0.0f || lv.y != 0.0f || lv.z != 0.0f || av.x != 0.0f || av.y != 0.0f || av.z != 0.0f)) {
      set_activation(body, k::ACT_WANTS_DEACTIVATION);
      body.deactivation_time = 0.0f;
    }
  }
}
RS_INLINE void phase_p5(ArenaState& s, const StaticData& w) {
  const float dt = 1.0f / 120.0f;
  Solver sv;
  sv.n_bodies = sv.n_contact = sv.n_friction = sv.n_special = 0;
  sv.fixed = -1;
  solver_setup(sv, s, w, dt);
  solver_iterate(sv);
  solver_finish(sv, s, dt);
}

}  // namespace rs
