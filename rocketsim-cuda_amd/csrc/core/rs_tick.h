// The tick and the decision: phases P0 (limits, UE3 quantisation), P2 (gravity, damping, predicted transforms), P6
// (integration, inertia, activation), P7 (contact records: ball hits, bumps, demos, folded world events), P8 (car post-tick,
// boost pads, ball finish, tick count), the ten phases in V3's order, and the 8-tick decision with the goal early exit.
// Reference: cuda/sim/arena_base.cuh, ball_base.cuh, car_base.cuh, boost_pad_grid.cuh, discrete_dynamics_world.cuh,
// phase.cuh, arena_step.cuh.
#pragma once
#include "rs_car.h"
#include "rs_solver.h"

namespace rs {

// ---- P0 (arena_base.cuh:21-38) ---------------------------------------------------------------------------------------------------
RS_INLINE void phase_p0(ArenaState& s, const StaticData& w) {
  s.events = 0;
  for (int d = 0; d < NUM_DYN; ++d) clear_accum(s.bodies[d]);
  for (int ci = 0; ci < NUM_CARS; ++ci) {
    Body& rb = s.bodies[1 + ci];
    limit_vels(rb, k::CAR_MAX_SPEED * k::UU_TO_BT, k::CAR_MAX_ANG_SPEED);
    quantize_body(&rb.world.t, &rb.lin_vel, &rb.ang_vel);
    s.cars[ci].vel = rb.lin_vel * k::BT_TO_UU;
    s.cars[ci].ang_vel = rb.ang_vel;
  }
  Body& ball = s.bodies[0];
  limit_vels(ball, w.mut_f[MF_BALL_MAX_SPEED] * k::UU_TO_BT, k::BALL_MAX_ANG_SPEED);
  quantize_body(&ball.world.t, &ball.lin_vel, &ball.ang_vel);
  s.ball.vel = ball.lin_vel * k::BT_TO_UU;
  s.ball.ang_vel = ball.ang_vel;
}

// ---- P2 / P6 (discrete_dynamics_world.cuh) ---------------------------------------------------------------------------------------
RS_INLINE void phase_p2(ArenaState& s, const StaticData& w) {
  const float dt = 1.0f / 120.0f;
  V3 g = w.gravity_bt * dt;
  for (int d = 0; d < NUM_DYN; ++d)
    if (body_active(s.bodies[d])) add_lin_impulse(s.bodies[d], g, true);
  for (int d = 0; d < NUM_DYN; ++d) {
    Body& b = s.bodies[d];
    if (b.activation == k::ACT_SLEEPING) {
      b.interp_world = b.world;
      continue;
    }
    apply_damping(b, w.bodies[w.dyn_body[d]], dt);
    b.interp_world = predict_trans(b, dt);
  }
}
RS_INLINE void phase_p6(ArenaState& s, const StaticData& w) {
  const float dt = 1.0f / 120.0f;
  for (int d = 0; d < NUM_DYN; ++d) {
    Body& b = s.bodies[d];
    if (!body_active(b)) continue;
    Aff t = predict_trans(b, dt);
    set_com_trans(b, w.bodies[w.dyn_body[d]], t);
  }
  for (int d = 0; d < NUM_DYN; ++d) update_activation(s.bodies[d], w.bodies[w.dyn_body[d]], dt);
}

// ---- P7 (arena_base.cuh:43-205, ball_base.cuh:22-48) ------------------------------------------------------------------------------
RS_INLINE bool within_rect_cone(V3 fwd, V3 right, V3 up, V3 delta, bool reverse, float yaw_deg, float pitch_deg) {
  V3 f = reverse ? -fwd : fwd, d;
  if (!try_normalize(delta, &d)) return false;
  float fd = dot(d, f);
  if (fd <= 0.0f) return false;
  const float rad = k::PI_F / 180.0f;
  return fabs_(dot(d, right)) <= fd * ftan(yaw_deg * rad) && fabs_(dot(d, up)) <= fd * ftan(pitch_deg * rad);
}
RS_FN void ball_hit(ArenaState& s, const StaticData& w, int ci) {  // ball/base.rs on_hit
  Car& car = s.cars[ci];
  Body& rb = s.bodies[0];
  V3 fwd = car.rot.c0;
  V3 rel_pos = s.ball.pos - car.pos, rel_vel = s.ball.vel - car.vel;
  float rel_speed = fmin_rs(length(rel_vel), k::HIT_MAX_DELTA_VEL_UU);
  bool can = !car.last_extra_hit_tick.some || (u64(car.last_extra_hit_tick.v) + 1 < u64(s.tick_count));
  if (rel_speed > 0.0f && can) {
    V3 dir = normalize_or_zero(rel_pos * V3{1.0f, 1.0f, k::HIT_Z_SCALE_NORMAL});
    V3 adj = (fwd * dot(dir, fwd)) * (1.0f - k::HIT_FORWARD_SCALE);
    dir = normalize_or_zero(dir - adj);
    V3 imp = ((dir * rel_speed) * curve_ball_car_extra_impulse(rel_speed)) * w.mut_f[MF_BALL_HIT_EXTRA_FORCE_SCALE];
    add_lin_impulse(rb, imp * k::UU_TO_BT, false);
    car.last_extra_hit_tick.some = 1;
    car.last_extra_hit_tick.v = s.tick_count;
  }
}
RS_FN void car_car_hit(ArenaState& s, const StaticData& w, int c1, int c2, const RecordPoint& mp) {  // base.rs on_car_car_collision
  int att = c1, vic = c2;
  if (s.cars[vic].is_demoed || s.cars[att].is_demoed) return;
  for (int swapped = 0; swapped < 2; ++swapped) {
    if (swapped) {
      int t = att;
      att = vic;
      vic = t;
    }
    Car& a = s.cars[att];
    Car& v = s.cars[vic];
    if (a.bump_cooldown_timer > 0.0f && (!a.bump_cooldown_victim.some || a.bump_cooldown_victim.v == (uint32_t)vic)) continue;
    V3 dpos = v.pos - a.pos;
    if (dot(a.vel, dpos) < 0.0f) continue;
    V3 vel_dir = normalize_or_zero(a.vel), to_vic = normalize_or_zero(dpos);
    if (dot(a.vel, to_vic) <= dot(v.vel, vel_dir)) continue;
    V3 af = a.rot.c0, ar = a.rot.c1, au = a.rot.c2;
    bool center_in = within_rect_cone(af, ar, au, dpos, false, k::BUMP_YAW_LIMIT_DEG, k::BUMP_PITCH_LIMIT_DEG);
    V3 on_att = swapped ? mp.world_on_b : mp.world_on_a, on_vic = swapped ? mp.world_on_a : mp.world_on_b;
    V3 gate = (v.pos - on_vic * k::BT_TO_UU) - (a.pos - on_att * k::BT_TO_UU);
    bool resid_in = within_rect_cone(af, ar, au, gate, false, k::BUMP_YAW_LIMIT_DEG, k::BUMP_PITCH_LIMIT_DEG);
    if (!center_in && !resid_in) continue;
    bool demo;
    if (w.demo_mode == k::DEMO_ON_CONTACT) {
      demo = true;
    } else if (w.demo_mode == k::DEMO_DISABLED) {
      demo = false;
    } else {
      demo = false;
      if (a.is_supersonic) {
        float fs = dot(a.vel, af);
        demo = fabs_(fs) >= k::SUPERSONIC_MAINTAIN_MIN_SPEED &&
               within_rect_cone(af, ar, au, dpos, fs < 0.0f, k::DEMO_YAW_LIMIT_DEG, k::DEMO_PITCH_LIMIT_DEG);
      }
    }
    if (demo && !w.enable_team_demos) demo = w.cars[att].team != w.cars[vic].team;
    if (demo) {
      v.is_demoed = 1;
      v.demo_respawn_timer = w.mut_f[MF_RESPAWN_DELAY];
    } else {
      float sp = fmin_rs(length(a.vel), k::BUMP_ATTACKER_SPEED_MAX);
      if (v.is_on_ground != 0) {
        float base = curve_bump_vel_ground(sp);
        float upward = curve_bump_upward_vel(sp) * w.mut_f[MF_BUMP_FORCE_SCALE];
        V3 imp = (vel_dir * base) + (v.rot.c2 * upward);
        v.vel_impulse_cache += imp * k::UU_TO_BT;
      } else {
        v.vel_impulse_cache += (vel_dir * curve_bump_vel_air(sp)) * k::UU_TO_BT;
      }
    }
    a.bump_cooldown_timer = w.mut_f[MF_BUMP_COOLDOWN_TIME];
    a.bump_cooldown_victim.some = 1;
    a.bump_cooldown_victim.v = (uint32_t)vic;
    s.events |= 1u << ((demo ? EV_DEMO : EV_BUMP) + vic);
  }
}
RS_INLINE void phase_p7(ArenaState& s, const StaticData& w) {
  s.events |= s.fold_events;
  for (int ci = 0; ci < NUM_CARS; ++ci)
    if (s.fold_events & (1u << (EV_CAR_HIT_WORLD + ci))) {
      s.cars[ci].world_contact_normal.some = 1;
      s.cars[ci].world_contact_normal.v = s.fold_car_normal[ci];
    }
  s.fold_events = 0;
  s.fold_car_normal[0] = V3{};
  s.fold_car_normal[1] = V3{};
  uint32_t n = s.n_records;
  for (uint32_t i = 0; i < n; ++i) {
    ContactRecord r = s.records[i];
    const BodyConst &ba = w.bodies[r.body_a], &bb = w.bodies[r.body_b];
    if (bb.user_idx == k::USER_BALL) {
      ball_hit(s, w, (int)ba.user_pointer);
      s.events |= 1u << (EV_CAR_HIT_BALL + ba.user_pointer);
    } else {
      car_car_hit(s, w, (int)ba.user_pointer, (int)bb.user_pointer, r.pt);
    }
  }
  for (uint32_t i = 0; i < s.n_records; ++i) s.records[i] = ContactRecord{};
  s.n_records = 0;
}

// ---- P8 (arena_base.cuh:210-220, car_base.cuh:28-95, boost_pad_grid.cuh, ball_base.cuh:11-19) ----------------------------------------
RS_FN void car_post_tick(Car& car, const Body& rb) {
  const float START2 = k::SUPERSONIC_START_SPEED * k::SUPERSONIC_START_SPEED;
  const float MAINT2 = k::SUPERSONIC_MAINTAIN_MIN_SPEED * k::SUPERSONIC_MAINTAIN_MIN_SPEED;
  if (car.is_demoed) return;
  car.rot = rb.world.m;
  float sp2 = length_sq(rb.lin_vel * k::BT_TO_UU);
  if (car.is_supersonic) {
    if (sp2 >= START2) {
      car.supersonic_grace_timer = 0.0f;
    } else if (sp2 < MAINT2) {
      car.is_supersonic = 0;
      car.supersonic_grace_timer = 0.0f;
    } else {
      car.supersonic_grace_timer = car.supersonic_grace_timer + k::TICK_TIME;
      if (car.supersonic_grace_timer >= k::SUPERSONIC_MAINTAIN_MAX_TIME) {
        car.is_supersonic = 0;
        car.supersonic_grace_timer = 0.0f;
      }
    }
  } else if (sp2 >= START2) {
    car.is_supersonic = 1;
    car.supersonic_grace_timer = 0.0f;
  } else {
    car.supersonic_grace_timer = 0.0f;
  }
  if (car.has_jumped && !car.is_jumping && car.is_on_ground && car.air_time > 0.0f) {
    int nc = 0;
    bool extending = true;
    for (int i = 0; i < NUM_WHEELS; ++i) {
      if (!car.wheels[i].ray.some) continue;
      ++nc;
      if (!(car.wheels[i].ray.v.suspension_relative_vel > 1.0f)) {
        extending = false;
        break;
      }
    }
    if (!(nc != 0 && extending)) car.has_jumped = 0;
  } else if (car.has_jumped && !car.is_jumping && car.is_on_ground && car.jump_ticks > k::JUMP_SETTLED_REARM_TICKS) {
    car.has_jumped = 0;
  }
  car.bump_cooldown_timer = fmax_rs(car.bump_cooldown_timer - k::TICK_TIME, 0.0f);
  car.prev_controls = car.controls;
}
RS_FN void car_finish_tick(Car& car, Body& rb) {
  if (car.is_demoed) return;
  V3 v = car.vel_impulse_cache;
  if (!(v.x == 0.0f && v.y == 0.0f && v.z == 0.0f)) {
    rb.lin_vel += car.vel_impulse_cache;
    car.vel_impulse_cache = splat(0.0f);
  }
  car.pos = rb.world.t * k::BT_TO_UU;
  car.vel = rb.lin_vel * k::BT_TO_UU;
  car.ang_vel = rb.ang_vel;
}
RS_FN bool pad_cyl(V3 pad, float radius, V3 p) {
  float dx = pad.x - p.x, dy = pad.y - p.y;
  return (dx * dx + dy * dy) < fpowi(radius, 2) && fabs_(p.z - pad.z) <= k::CYL_HEIGHT;
}
RS_FN bool pad_box(V3 pad, float box_radius, const Car& car, const CarConst& cc) {  // 15-axis SAT of the hitbox against the pad box
  const M3 r = car.rot;
  const V3 center = car.pos + r * cc.hitbox_offset;
  const V3 ca[3] = {r.c0, r.c1, r.c2};
  const V3 ch = cc.hitbox_size * 0.5f, ph = V3{box_radius, box_radius, k::PAD_BOX_HEIGHT};
  const V3 delta = center - pad;
  for (int kk = 0; kk < 15; ++kk) {
    V3 ax = kk < 3 ? unit_axis(kk) : kk < 6 ? ca[kk - 3] : cross(ca[(kk - 6) / 3], unit_axis((kk - 6) % 3));
    if (length_sq(ax) <= 1.1920929e-7f) continue;
    float pr = dot(ph, V3{fabs_(dot(unit_axis(0), ax)), fabs_(dot(unit_axis(1), ax)), fabs_(dot(unit_axis(2), ax))});
    float cr = dot(ch, V3{fabs_(dot(ca[0], ax)), fabs_(dot(ca[1], ax)), fabs_(dot(ca[2], ax))});
    if (fabs_(dot(delta, ax)) > pr + cr) return false;
  }
  return true;
}
RS_FN bool can_take_boost(const Car& car, float mx) { return !car.is_demoed && !(car.boost >= mx); }
RS_FN int pad_winner(const ArenaState& s, int pad, float mx) {
  const PadTouching& t = s.pad_touching[pad];
  for (uint32_t q = 0; q < t.n; ++q)
    if (can_take_boost(s.cars[t.car_idx[q]], mx)) return (int)t.car_idx[q];
  return -1;
}
// The steps and their order are the reference's; the pad lists are read together (one bit per pad with a touching car),
// a list is written back only when it changed, the discovery screen is one mask over the pads' boxes, and the winner is
// computed for the (at most NUM_CARS) pads a car touches rather than for all 34: the same reads decide the same writes.
RS_FN void boost_pads_tick(ArenaState& s, const StaticData& w) {  // boost_pad_grid.cuh:56-133
  static_assert(NUM_PADS <= 64, "one bit per pad");
  const float mx = w.mut_f[MF_CAR_MAX_BOOST_AMOUNT];
  int touching[NUM_CARS];
  for (int ci = 0; ci < NUM_CARS; ++ci) touching[ci] = -1;
  uint64_t listed_pads = 0;
  RS_UNROLL
  for (int i = 0; i < NUM_PADS; ++i) listed_pads |= (uint64_t)(s.pad_touching[i].n != 0u) << i;
  for (uint64_t m = listed_pads; m != 0; m &= m - 1) {  // retention, pad order (a pad with no car keeps nothing)
    const int i = __builtin_ctzll(m);
    const PadConst& pad = w.pads[i];
    PadTouching& t = s.pad_touching[i];
    uint32_t kept = 0;
    for (uint32_t q = 0; q < t.n; ++q) {
      uint32_t idx = t.car_idx[q];
      const Car& car = s.cars[idx];
      if (!car.is_demoed && touching[idx] < 0 &&
          (pad_cyl(pad.pos, pad.cyl_radius, car.pos) || pad_box(pad.pos, pad.box_radius, car, w.cars[idx]))) {
        touching[idx] = i;
        t.car_idx[kept++] = idx;
      }
    }
    if (kept != t.n) {
      for (uint32_t q = kept; q < t.n; ++q) t.car_idx[q] = 0;
      t.n = kept;
    }
  }
  // discovery through the cylinder. Both cars' box screens first, in one pass over the pads (each pad's box read once):
  // a screen reads only the car's position and the pads' boxes, which the discovery does not change.
  bool look[NUM_CARS];
  V3 cpos[NUM_CARS];
  for (int ci = 0; ci < NUM_CARS; ++ci) {
    const Car& car = s.cars[ci];
    cpos[ci] = car.pos;
    look[ci] = !(car.is_demoed || car.pos.z > w.max_pad_z || touching[ci] >= 0);
  }
  uint64_t in_boxes[NUM_CARS] = {};
  if (look[0] || look[1]) {
    static_assert(NUM_CARS == 2, "the screens test two cars");
    RS_UNROLL
    for (int i = 0; i < NUM_PADS; ++i) {
      const Aabb b = w.pads[i].aabb;
      RS_UNROLL
      for (int ci = 0; ci < NUM_CARS; ++ci)
        in_boxes[ci] |= (uint64_t)(all3(cmple(b.min, cpos[ci])) && all3(cmpge(b.max, cpos[ci]))) << i;
    }
  }
  for (int ci = 0; ci < NUM_CARS; ++ci) {
    if (!look[ci]) continue;
    const Car& car = s.cars[ci];
    for (uint64_t m = in_boxes[ci]; m != 0; m &= m - 1) {
      const int i = __builtin_ctzll(m);
      const PadConst& pad = w.pads[i];
      if (!pad_cyl(pad.pos, pad.cyl_radius, car.pos)) continue;
      PadTouching& t = s.pad_touching[i];
      bool listed = false;
      for (uint32_t q = 0; q < t.n; ++q) listed |= t.car_idx[q] == (uint32_t)ci;
      if (!listed) {
        if (t.n < (uint32_t)NUM_CARS) t.car_idx[t.n++] = (uint32_t)ci;
        else ++s.overflow[OVF_PAD_TOUCHING];
      }
      touching[ci] = i;
      break;
    }
  }
  int win[NUM_CARS];  // the winner of the pad each car touches, as of now
  for (int ci = 0; ci < NUM_CARS; ++ci) win[ci] = touching[ci] >= 0 ? pad_winner(s, touching[ci], mx) : -1;
  for (int ci = 0; ci < NUM_CARS; ++ci) {  // pickups, car order; the pad's winner refreshed after a pickup
    Car& car = s.cars[ci];
    if (car.is_demoed || car.boost >= mx) continue;
    int i = touching[ci];
    if (i < 0 || win[ci] != ci) continue;
    const PadConst& pad = w.pads[i];
    int64_t tick = (int64_t)u64(s.tick_count);
    if (s.pad_gave_boost[i].some) {
      int64_t last = (int64_t)u64(s.pad_gave_boost[i].v);
      if (i64_to_f(tick - last) * k::TICK_TIME < pad.max_cooldown) continue;
    }
    car.boost = fmin_rs(car.boost + pad.boost_amount, mx);
    s.pad_gave_boost[i].some = 1;
    s.pad_gave_boost[i].v = u64w((uint64_t)tick);
    s.events |= 1u << (EV_PICKUP + ci);
    const int nw = pad_winner(s, i, mx);
    for (int cj = 0; cj < NUM_CARS; ++cj)
      if (touching[cj] == i) win[cj] = nw;
  }
}
RS_INLINE void phase_p8(ArenaState& s, const StaticData& w) {
  for (int ci = 0; ci < NUM_CARS; ++ci) {
    car_post_tick(s.cars[ci], s.bodies[1 + ci]);
    car_finish_tick(s.cars[ci], s.bodies[1 + ci]);
  }
  if (w.boost_pads_enabled) boost_pads_tick(s, w);
  const Body& rb = s.bodies[0];  // ball finish_physics_tick
  s.ball.vel = rb.lin_vel * k::BT_TO_UU;
  s.ball.ang_vel = rb.ang_vel;
  s.ball.pos = rb.world.t * k::BT_TO_UU;
  s.ball.rot = rb.world.m;
  s.ball.tick_count_since_kickoff = u64w(u64(s.ball.tick_count_since_kickoff) + 1);
  s.tick_count = u64w(u64(s.tick_count) + 1);
}

// ---- the tick and the decision (phase.cuh, arena_step.cuh) -------------------------------------------------------------------
constexpr int NUM_PHASES = 10;
enum PhaseStatus : int { PHASE_OK = 0, PHASE_UNIMPLEMENTED = -1, PHASE_BAD_INDEX = -2 };

// Phase k of one tick (0 P0, 1 P1, 2 P2, 3 P3a, 4 P3b, 5 P4, 6 P5, 7 P6, 8 P7, 9 P8): state at boundary k to boundary k + 1.
// UW: every lane steps in the same world (its pointer is then made wave-uniform); k is the same in every lane.
template <bool UW>
RS_NOINL int run_phase_in(int kk, RS_PRIVATE_AS ArenaState* sp, const RS_CONST_AS StaticData* wp) {
  if (UW) wp = rs_uniform_ptr(wp);
  RS_UNIFORM_INDEX(kk);
  ArenaState& s = *(ArenaState*)sp;
  const StaticData& w = *(const StaticData*)wp;
  switch (kk) {
    case 0: phase_p0(s, w); return PHASE_OK;
    case 1: phase_p1<UW>(s, w); return PHASE_OK;
    case 2: phase_p2(s, w); return PHASE_OK;
    case 3: phase_p3a(s, w); return PHASE_OK;
    case 4: return phase_p3b<UW>(s, w) == 0 ? PHASE_OK : PHASE_UNIMPLEMENTED;
    case 5: phase_p4(s, w); return PHASE_OK;
    case 6: phase_p5(s, w); return PHASE_OK;
    case 7: phase_p6(s, w); return PHASE_OK;
    case 8: phase_p7(s, w); return PHASE_OK;
    case 9: phase_p8(s, w); return PHASE_OK;
    default: return PHASE_BAD_INDEX;
  }
}
// Every caller holds its arena in a lane-private copy and the StaticData in global memory. UW (uniform world): the
// caller's lanes all step in the same world.
template <bool UW = false>
RS_FN int run_phase(int kk, ArenaState& s, const StaticData& w) {
  return run_phase_in<UW>(kk, (RS_PRIVATE_AS ArenaState*)&s, (const RS_CONST_AS StaticData*)&w);
}
// One tick (base.rs step_tick): the ten phases in order.
template <bool UW = false>
RS_FN int step_tick(ArenaState& s, const StaticData& w) {
  int st = PHASE_OK;
  for (int kk = 0; kk < NUM_PHASES; ++kk)
    if (run_phase<UW>(kk, s, w) != PHASE_OK) st = PHASE_UNIMPLEMENTED;
  return st;
}
RS_FN bool ball_scored(const ArenaState& s, const StaticData& w) {  // base.rs is_ball_scored, Soccar
  V3 p = s.bodies[0].world.t * k::BT_TO_UU;
  return fabs_(p.y) > w.goal_base_threshold_y + w.mut_f[MF_BALL_RADIUS];
}
RS_FN bool any_overflow(const ArenaState& s) {
  uint32_t o = 0;
  for (int i = 0; i < OVF_COUNT; ++i) o |= s.overflow[i];
  return o != 0;
}

// One decision (arena_step.cuh:83-116, sim.cu:78-93): up to n ticks, the cars' controls set before each. The decision ends
// after the tick on which the ball is first scored, so *s is the goal state then (the reference steps the remaining ticks
// and discards them; stopping is the same result). after_tick(t, s) sees the state after each tick up to the goal.
// Returns the result word: the scored tick (0 for none) | SIM8_UNIMPLEMENTED | SIM8_OVERFLOW.
// UW: as for step_tick.
template <bool UW = false, class AfterTick>
RS_FN uint32_t run_decision(ArenaState& s, const StaticData& w, const Controls* ctl, uint32_t n, AfterTick after_tick) {
  uint32_t unimpl = 0, scored = 0;
  for (uint32_t t = 1; t <= n; ++t) {
    for (int ci = 0; ci < NUM_CARS; ++ci) s.cars[ci].controls = ctl[ci];
    if (step_tick<UW>(s, w) != PHASE_OK) unimpl = SIM8_UNIMPLEMENTED;
    after_tick(t, s);
    if (ball_scored(s, w)) {
      scored = t;
      break;
    }
  }
  return scored | unimpl | (any_overflow(s) ? SIM8_OVERFLOW : 0u);
}
struct NoTrace {
  RS_FN void operator()(uint32_t, const ArenaState&) const {}
};

}  // namespace rs
