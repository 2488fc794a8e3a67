// Phase P1, the car pre-tick: per car the flip z-damp prepass, then game logic (respawn, wheels and steering, jump,
// auto-flip, double jump / flip, air control, auto-roll, boost) and the btVehicleRL update (four wheel rays, suspension,
// friction). Reference: cuda/sim/arena_base_p1.cuh, car_pre_tick.cuh, vehicle_rl.cuh, wheel_info.cuh, contact_constraint.cuh.
#pragma once
#include "rs_ray.h"

namespace rs {

// ---- the wheel (wheel_info.cuh) ---------------------------------------------------------------------------------------------
RS_FN void wheel_ray_ends(const WheelConst& wc, const Aff& chassis, V3* src, V3* dst) {  // :14-18 (the ray's ends)
  *src = xform_point(chassis, wc.connection_point);
  *dst = *src - chassis.m.c2 * wc.ray_length;
}
RS_FN void wheel_prepare(Wheel& wh, const WheelConst& wc, const Aff& chassis, V3* src, V3* dst) {  // :14-18
  wheel_ray_ends(wc, chassis, src, dst);
  wh.hard_point = *src;
}
RS_FN void wheel_reset(Wheel& wh) {
  wh.extra_pushback = 0.0f;
  wh.ray = Opt<RaycastInfo>{};
}
// contact_constraint.cuh:11-32 resolve_single_collision
RS_FN float resolve_single_collision(const BodyView& b1, const BodyView& b2, V3 pos, V3 n, float dt, float distance) {
  V3 rel1 = pos - b1.world.t, rel2 = pos - b2.world.t;
  V3 v = vel_at(b1, rel1) - vel_at(b2, rel2);
  float rel_vel = dot(n, v);
  float pos_err = (k::WHEEL_PUSHBACK_ERP * fneg(distance)) / dt;
  float vel_err = fneg(rel_vel);
  float jinv = 1.0f / (impulse_denom(b1, pos, n) + impulse_denom(b2, pos, n));
  return fmax_rs(pos_err * jinv + vel_err * jinv, 0.0f);
}
// contact_constraint.cuh:35-46 resolve_single_bilateral_fake_ground
RS_FN float resolve_bilateral_ground(const BodyView& b1, V3 inv_inertia_local, V3 pos, V3 n) {
  V3 rel1 = pos - b1.world.t;
  V3 v1 = vel_at(b1, rel1);
  V3 aj = mul_t(b1.world.m, cross(rel1, n));
  float jinv = 1.0f / (b1.inv_mass + dot(inv_inertia_local * aj, aj));
  return (-0.2f * dot(n, v1)) * jinv;
}
// wheel_info.cuh:21-87 apply_ray_cast
RS_FN void wheel_apply_ray(Wheel& wh, const WheelConst& wc, const BodyView& chassis, const Aff& ct, V3 front_axle,
                           const WheelHit& hit, const BodyView& ground, float dt, bool front) {
  V3 cp = hit.point, cn = hit.normal;
  bool world = view_static(ground);
  wh.axle_dir = front ? front_axle : ct.m.c1;
  V3 up = ct.m.c2;
  float trace_len = dot(wh.hard_point - cp, up);
  float max_susp = wc.suspension_rest_length + k::MAX_SUSPENSION_TRAVEL * k::UU_TO_BT;
  wh.vel_at_contact_point = vel_at(chassis, cp - ct.t);
  float proj_vel = dot(cn, wh.vel_at_contact_point);
  float denom = dot(cn, up);
  float srv, clipped;
  if (denom > 0.1f) {
    float inv = 1.0f / denom;
    srv = proj_vel * inv;
    clipped = inv;
  } else {
    srv = 0.0f;
    clipped = 10.0f;
  }
  float susp_len = fmin_rs(trace_len - wc.radius, max_susp);
  wh.extra_pushback = 0.0f;
  if (world || ground.user_idx == k::USER_CAR) {
    float thresh = (wc.suspension_rest_length + wc.radius) - k::SUSPENSION_SUBTRACTION;
    if (trace_len < thresh) {
      float r = resolve_single_collision(chassis, ground, hit.point, hit.normal, dt, trace_len - thresh);
      wh.extra_pushback = r / (float)NUM_WHEELS;
    }
  }
  RaycastInfo ri;
  ri.contact_normal = cn;
  ri.contact_point = cp;
  ri.ground_body = hit.body;
  ri.suspension_length = susp_len;
  ri.impulse = splat(0.0f);
  ri.ground_stick = Opt<V3>{};
  if (!world) {
    ri.ground_stick.some = 1;
    ri.ground_stick.v = -cn;
  }
  ri.in_contact_with_world = world ? 1u : 0u;
  ri.clipped_inv_contact_dot_suspension = clipped;
  ri.suspension_relative_vel = srv;
  wh.ray.some = 1;
  wh.ray.v = ri;
}
// wheel_info.cuh:96-127 refresh_friction_curves
RS_FN void wheel_friction_curves(Wheel& wh, const BodyView& chassis, V3 chassis_t, V3 cn, float handbrake, float throttle,
                                 bool three_wheels, bool dynamic_hit) {
  V3 lat = wh.axle_dir;
  V3 lng = cross(lat, cn);
  V3 cv = (cross(chassis.ang_vel, wh.hard_point - chassis_t) + chassis.lin_vel) * k::BT_TO_UU;
  float base = fabs_(dot(cv, lat));
  float in = base > 5.0f ? base / (fabs_(dot(cv, lng)) + base) : 0.0f;
  float lat_f = three_wheels ? curve_lat_friction_3w(in) : curve_lat_friction(in);
  float long_f = 1.0f;
  if (handbrake != 0.0f) {
    lat_f = lat_f * (1.0f + (curve_handbrake_lat(in) - 1.0f) * handbrake);
    long_f = long_f * (1.0f + (curve_handbrake_long(in) - 1.0f) * handbrake);
  }
  if (throttle == 0.0f || dynamic_hit) {
    float ns = curve_non_sticky_friction(cn.z);
    lat_f = lat_f * ns;
    long_f = long_f * ns;
  }
  wh.lat_friction = lat_f;
  wh.long_friction = long_f;
}
// What the impulse passes of the vehicle update read of a wheel (its ray result, axle, drive and friction), read for all
// four wheels at once, after the last write to any of it: no wheel's loads then wait behind another wheel's branch.
struct WheelContact {
  bool some;
  V3 cn, cp, axle, vcp;
  float susp_len, clipped, srv, pushback, engine, brake, lng, lat;
};
RS_FN WheelContact wheel_contact(const Wheel& wh) {
  const RaycastInfo& ri = wh.ray.v;
  return WheelContact{wh.ray.some != 0, ri.contact_normal, ri.contact_point, wh.axle_dir, wh.vel_at_contact_point,
                      ri.suspension_length, ri.clipped_inv_contact_dot_suspension, ri.suspension_relative_vel,
                      wh.extra_pushback, wh.engine_force, wh.brake, wh.long_friction, wh.lat_friction};
}
// wheel_info.cuh:130-171 calc_friction_impulses / update_friction_impulse: the wheel's friction impulse (a wheel out of
// contact keeps its stored one)
RS_FN V3 wheel_friction_impulse(const WheelContact& wc, const BodyView& chassis, V3 inv_inertia_local, float dt, float friction_scale) {
  V3 cn = wc.cn, cp = wc.cp;
  V3 axle = normalize_or_zero(wc.axle);
  V3 fwd = normalize_or_zero(cross(cn, axle));
  float side = resolve_bilateral_ground(chassis, inv_inertia_local, cp, axle);
  float rolling;
  if (wc.engine == 0.0f) {
    if (wc.brake == 0.0f) {
      rolling = 0.0f;
    } else {
      float rel_vel = dot(wc.vcp, fwd);
      if (dt > 1.0f / 80.0f) {
        float th = 0.8f - 1.0f / (dt * 150.0f);
        if (fabs_(rel_vel) < th) rel_vel = 0.0f;
      }
      rolling = fclamp_rs(fneg(rel_vel) * 113.73963f, fneg(wc.brake), wc.brake);
    }
  } else {
    rolling = fneg(wc.engine) / friction_scale;
  }
  V3 total = fwd * rolling * wc.lng + axle * side * wc.lat;
  return total * friction_scale;
}
// wheel_info.cuh:174-191 update_suspension
RS_FN void wheel_suspension(const WheelContact& wcn, const WheelConst& wc, Body& cb, const BodyConst& cbc, V3 chassis_t, float dt) {
  if (!wcn.some) return;
  float force = ((wc.suspension_rest_length - wcn.susp_len) * k::SUSPENSION_STIFFNESS) * wcn.clipped;
  float damp = wcn.srv < 0.0f ? k::WHEELS_DAMPING_COMPRESSION : k::WHEELS_DAMPING_RELAXATION;
  float f = (force - damp * wcn.srv) * wc.suspension_force_scale;
  if (f <= 0.0f) return;
  float scale = f * dt + wcn.pushback;
  add_impulse(cb, cbc, IMP_LINEAR_REL_POS, wcn.cn * scale, wcn.cp - chassis_t, true, false);
}
// wheel_info.cuh:194-201 apply_friction_impulses
RS_FN void wheel_apply_friction(const WheelContact& wcn, V3 impulse, Body& cb, const BodyConst& cbc, const Aff& ct, float dt) {
  if (!wcn.some) return;
  V3 off = wcn.cp - ct.t;
  float up_dot = dot(ct.m.c2, off);
  add_impulse(cb, cbc, IMP_LINEAR_REL_POS, impulse * dt, off - ct.m.c2 * up_dot, true, false);
}

// ---- the vehicle (vehicle_rl.cuh) --------------------------------------------------------------------------------------------
RS_FN V3 steered_axle(const Aff& ct, float steer) {  // :11-17: Quat::from_axis_angle(up, steer) * right
  float s = fsin(steer * 0.5f), c = fcos(steer * 0.5f);
  V3 v = ct.m.c2 * s;
  return qrotate(Quat{v.x, v.y, v.z, c}, ct.m.c1);
}
RS_FN V3 upwards_from_wheels(const Car& car, V3 chassis_up) {  // :20-29
  V3 sum = splat(0.0f);
  for (int i = 0; i < NUM_WHEELS; ++i)
    if (car.wheels[i].ray.some && car.wheels[i].ray.v.in_contact_with_world) sum += car.wheels[i].ray.v.contact_normal;
  V3 out;
  return try_normalize(sum, &out) ? out : chassis_up;
}
RS_FN void wheel_dynamic_effects(ArenaState& s, const StaticData& w, const Wheel& wh) {  // :32-52
  if (!wh.ray.some) return;
  const RaycastInfo& ri = wh.ray.v;
  if (ri.ground_stick.some) {
    Body& g = s.bodies[dyn_slot(w, ri.ground_body)];
    add_impulse(g, w.bodies[ri.ground_body], IMP_LINEAR_REL_POS, ri.ground_stick.v, ri.contact_point - g.world.t, true, false);
  }
  if (wh.extra_pushback <= 0.0f) return;
  if (w.bodies[ri.ground_body].user_idx != k::USER_CAR) return;
  Body& victim = s.bodies[dyn_slot(w, ri.ground_body)];
  float full = wh.extra_pushback * (float)NUM_WHEELS;
  add_impulse(victim, w.bodies[ri.ground_body], IMP_LINEAR_REL_POS, -ri.contact_normal * full, ri.contact_point - victim.world.t,
              true, false);
}
RS_FN void vehicle_update(ArenaState& s, const StaticData& w, int ci, float dt, float handbrake, float throttle, bool three_wheels,
                          const WheelRayPrep& prep) {
  Car& car = s.cars[ci];
  const CarConst& cc = w.cars[ci];
  const uint32_t chassis_body = cc.body_idx;
  const int cd = dyn_slot(w, chassis_body);
  const BodyConst& cbc = w.bodies[chassis_body];
  const Aff ct = s.bodies[cd].world;
  const V3 chassis_t = ct.t;
  float friction_scale = (cbc.inv_mass == 0.0f ? 0.0f : 1.0f / cbc.inv_mass) / 3.0f;
  V3 src[NUM_WHEELS], dst[NUM_WHEELS];
  for (int i = 0; i < NUM_WHEELS; ++i) wheel_prepare(car.wheels[i], cc.wheels[i], ct, &src[i], &dst[i]);
  WheelHit hits[NUM_WHEELS];
  cast_wheel_rays(s, w, src, dst, chassis_body, prep, ci, hits);
  bool cache_some = false;  // the front wheels share one steered axle while their angle matches
  float cache_angle = 0.0f;
  V3 cache_axle = splat(0.0f);
  int in_contact = 0;
  // The chassis and the wheels' grounds read before the loop: no wheel's effects reach the chassis (its rays ignore
  // it), and a ground changes in the loop only when it is a dynamic body an earlier wheel pushed, so a dynamic ground is
  // read again at its wheel.
  const BodyView chassis = body_view(s, w, chassis_body);
  BodyView grounds[NUM_WHEELS];
  RS_UNROLL for (int i = 0; i < NUM_WHEELS; ++i) grounds[i] = body_view(s, w, hits[i].some ? hits[i].body : chassis_body);
  RS_UNROLL for (int i = 0; i < NUM_WHEELS; ++i) {
    Wheel& wh = car.wheels[i];
    if (!hits[i].some) {
      wheel_reset(wh);
      continue;
    }
    BodyView ground = grounds[i];
    if (!view_static(ground)) ground = body_view(s, w, hits[i].body);
    ++in_contact;
    bool front = i < 2;
    V3 axle;
    if (front) {
      if (cache_some && cache_angle == wh.steer_angle) {
        axle = cache_axle;
      } else {
        axle = steered_axle(ct, wh.steer_angle);
        cache_some = true;
        cache_angle = wh.steer_angle;
        cache_axle = axle;
      }
    } else {
      axle = ct.m.c1;
    }
    wheel_apply_ray(wh, cc.wheels[i], chassis, ct, axle, hits[i], ground, dt, front);
    wheel_friction_curves(wh, chassis, chassis_t, hits[i].normal, handbrake, throttle, three_wheels, !view_static(ground));
    wheel_dynamic_effects(s, w, wh);
  }
  if (in_contact < 3)
    for (int i = 0; i < NUM_WHEELS; ++i) car.wheels[i].engine_force = car.wheels[i].engine_force / 4.0f;
  WheelContact wcn[NUM_WHEELS];
  RS_UNROLL for (int i = 0; i < NUM_WHEELS; ++i) wcn[i] = wheel_contact(car.wheels[i]);
  RS_UNROLL for (int i = 0; i < NUM_WHEELS; ++i) wheel_suspension(wcn[i], cc.wheels[i], s.bodies[cd], cbc, chassis_t, dt);
  V3 imp[NUM_WHEELS];
  {
    BodyView chassis = body_view(s, w, chassis_body);
    RS_UNROLL for (int i = 0; i < NUM_WHEELS; ++i) {
      imp[i] = car.wheels[i].ray.v.impulse;
      if (wcn[i].some) {
        imp[i] = wheel_friction_impulse(wcn[i], chassis, cbc.inv_inertia_local, dt, friction_scale);
        car.wheels[i].ray.v.impulse = imp[i];
      }
    }
  }
  RS_UNROLL for (int i = 0; i < NUM_WHEELS; ++i) wheel_apply_friction(wcn[i], imp[i], s.bodies[cd], cbc, ct, dt);
}

// ---- auto-flip roll sign (car_pre_tick.cuh:23-72) ------------------------------------------------------------------------------
RS_FN int32_t axis_angle_to_int(float a0, float a1) {  // car/base.rs axis_angle_to_int
  double ang = datan2((double)a0, (double)a1);
  float u = ((float)ang * 32768.0f) / k::PI_F;
  return f2i_rs(sse_round_even(u * 2.0f + 0.5f)) >> 1;
}
RS_FN V3 rebuild_row1(const StaticData& w, int32_t pitch, int32_t yaw) {  // car/base.rs rebuild_row1
  float tp = w.table_sin[(uint32_t)((pitch >> 2) & 0x3FFF)], ty = w.table_sin[(uint32_t)((yaw >> 2) & 0x3FFF)];
  float tr = w.table_sin[0];
  float tpq = w.table_sin[(uint32_t)(((pitch + 0x4000) >> 2) & 0x3FFF)], tyq = w.table_sin[(uint32_t)(((yaw + 0x4000) >> 2) & 0x3FFF)];
  float trq = w.table_sin[0x1000];
  return V3{(tp * tr) * tyq - trq * ty, tyq * trq + (tp * tr) * ty, fneg(tr) * tpq};
}
RS_FN float autoflip_sign(const StaticData& w, const M3& rot) {
  int32_t yaw = axis_angle_to_int(rot.c0.y, rot.c0.x);
  float len = (float)dhypot((double)rot.c0.x, (double)rot.c0.y);
  int32_t pitch = axis_angle_to_int(rot.c0.z, len);
  V3 r1 = rebuild_row1(w, pitch, yaw);
  int16_t roll = (int16_t)axis_angle_to_int(dot(rot.c2, r1), dot(rot.c1, r1));
  return roll > 0 ? 1.0f : -1.0f;
}

// ---- fastrand WyRand (fastrand 2.5.0; car_pre_tick.cuh:76-101) ------------------------------------------------------------------
RS_FN uint64_t mulhi64(uint64_t a, uint64_t b) {
  uint64_t hi, lo;
  crm::mul64(a, b, &hi, &lo);
  return hi;
}
RS_FN uint64_t wyrand_next(U64w* st) {
  uint64_t s = u64(*st) + 0x2d358dccaa6c78a5ull;
  *st = u64w(s);
  uint64_t o = s ^ 0x8bb84b93962eacc9ull;
  return (s * o) ^ mulhi64(s, o);
}
RS_FN uint64_t wyrand_below(U64w* st, uint64_t n) {  // Lemire's unbiased multiply-shift
  uint64_t r = wyrand_next(st);
  uint64_t hi = mulhi64(r, n), lo = r * n;
  if (lo < n) {
    uint64_t t = (0ull - n) % n;
    while (lo < t) {
      r = wyrand_next(st);
      hi = mulhi64(r, n);
      lo = r * n;
    }
  }
  return hi;
}

// ---- respawn (car_pre_tick.cuh:106-193) ----------------------------------------------------------------------------------------
RS_FN M3 m3_from_yaw(float yaw) {  // glam Mat3A::from_euler(ZYX, yaw, 0, 0)
  float si = fsin(yaw), ci = fcos(yaw), sj = fsin(0.0f), cj = fcos(0.0f), sh = fsin(0.0f), ch = fcos(0.0f);
  float cc = ci * ch, cs = ci * sh, sc = si * ch, ss = si * sh;
  return M3{{cj * ci, cj * si, fneg(sj)}, {sj * cs - sc, sj * ss + cc, cj * sh}, {sj * cc + ss, sj * sc - cs, cj * ch}};
}
RS_FN void car_respawn(ArenaState& s, const StaticData& w, int ci, float boost_amount) {
  Car& car = s.cars[ci];
  const CarConst& cc = w.cars[ci];
  Body& rb = s.bodies[dyn_slot(w, cc.body_idx)];
  const BodyConst& bc = w.bodies[cc.body_idx];
  const float loc_x[4] = {-2304.0f, -2688.0f, 2304.0f, 2688.0f};
  uint32_t spot = (uint32_t)wyrand_below(&s.rng, 4);
  float y_dir = cc.team == 0 ? -1.0f : 1.0f;
  V3 pos = V3{loc_x[spot], -4608.0f * fneg(y_dir), k::RESPAWN_Z};
  M3 rot = m3_from_yaw(1.57079632679489661923f + (cc.team == 0 ? 0.0f : k::PI_F));
  float dodge = car.dodge_deadzone;
  V3 vic = car.vel_impulse_cache;
  Wheel wheels[NUM_WHEELS];
  for (int i = 0; i < NUM_WHEELS; ++i) wheels[i] = car.wheels[i];
  car = Car{};
  car.dodge_deadzone = dodge;
  car.vel_impulse_cache = vic;
  for (int i = 0; i < NUM_WHEELS; ++i) car.wheels[i] = wheels[i];
  car.pos = pos;
  car.rot = rot;
  car.boost = fclamp_rs(boost_amount, 0.0f, k::BOOST_MAX);
  car.is_on_ground = 1;
  // sync_state_to_body (car_base.cuh:17-25)
  set_world_trans(rb, Aff{car.rot, car.pos * k::UU_TO_BT});
  rb.interp_world = rb.world;
  rb.lin_vel = car.vel * k::UU_TO_BT;
  rb.ang_vel = car.ang_vel;
  clear_accum(rb);
  update_inertia(rb, bc);
  car.vel_impulse_cache = splat(0.0f);
  for (int i = 0; i < NUM_WHEELS; ++i) {  // reset_transient_contacts
    Wheel& wh = car.wheels[i];
    wh = Wheel{};
    wh.lat_friction = 1.0f;
    wh.long_friction = 1.0f;
  }
  car.sticky_gate_prev = 0;
}

// ---- the per-tick updates (car_pre_tick.cuh:197-635) --------------------------------------------------------------------------
RS_FN uint32_t wheels_in_contact(const Car& car) {
  uint32_t n = 0;
  for (int i = 0; i < NUM_WHEELS; ++i) n += car.wheels_with_contact[i].some ? 1u : 0u;
  return n;
}
RS_FN void car_update_wheels(ArenaState& s, const StaticData& w, int ci, Body& rb, float fwd_speed, bool* up_some, V3* up_dir) {
  Car& car = s.cars[ci];
  const CarConst& cc = w.cars[ci];
  const BodyConst& bc = w.bodies[cc.body_idx];
  float hb_delta = (car.controls.handbrake ? k::POWERSLIDE_RISE_RATE : -k::POWERSLIDE_FALL_RATE) * k::TICK_TIME;
  car.handbrake_val = fclamp_rs(car.handbrake_val + hb_delta, 0.0f, 1.0f);
  float real_brake = 0.0f;
  float raw_throttle = car.controls.throttle;
  V3 q = quant_axes(V3{raw_throttle, car.controls.steer, 0.0f});
  float throttle = q.x, steer = q.y;
  bool all_contact = wheels_in_contact(car) == (uint32_t)NUM_WHEELS;
  float real_throttle = (car.controls.boost && car.boost > 0.0f && all_contact) ? 1.0f : throttle;
  float abs_speed = fabs_(fwd_speed);
  float engine_throttle = real_throttle;
  if (!car.controls.handbrake) {
    if (fabs_(real_throttle) > k::THROTTLE_DEADZONE) {
      if (abs_speed > k::STOPPING_FORWARD_VEL && fsignum_rs(real_throttle) != fsignum_rs(fwd_speed)) {
        real_brake = 1.0f;
        if (abs_speed > k::BRAKING_NO_THROTTLE_SPEED_THRESH) engine_throttle = 0.0f;
      }
    } else {
      engine_throttle = 0.0f;
      real_brake = abs_speed < k::STOPPING_FORWARD_VEL ? 1.0f : k::COASTING_BRAKE_FACTOR;
    }
  }
  float drive_scale = curve_drive_torque(abs_speed);
  float engine_force = (engine_throttle * (k::THROTTLE_TORQUE_AMOUNT * k::UU_TO_BT)) * drive_scale;
  float brake_force = real_brake * (k::BRAKE_TORQUE_AMOUNT * k::UU_TO_BT);
  for (int i = 0; i < NUM_WHEELS; ++i) {
    car.wheels[i].engine_force = engine_force;
    car.wheels[i].brake = brake_force;
  }
  float steer_angle = cc.three_wheels ? curve_steer_angle_3w(abs_speed) : curve_steer_angle(abs_speed);
  if (car.handbrake_val != 0.0f)
    steer_angle = steer_angle + (curve_powerslide_steer(abs_speed) - steer_angle) * car.handbrake_val;
  steer_angle = steer_angle * steer;
  car.wheels[0].steer_angle = steer_angle;
  car.wheels[1].steer_angle = steer_angle;
  if (car.sticky_gate_prev) {
    V3 up;
    if (*up_some) {
      up = *up_dir;
    } else {
      up = upwards_from_wheels(car, rb.world.m.c2);
      *up_dir = up;
      *up_some = true;
    }
    bool full_stick = raw_throttle != 0.0f || abs_speed > k::STOPPING_FORWARD_VEL;
    float sticky = (cc.three_wheels ? 0.0f : 1.0f) * 0.5f;
    if (full_stick) sticky = sticky + (1.0f - fabs_(up.z));
    add_impulse(rb, bc, IMP_LINEAR, up * sticky * -fabs_(w.gravity_uu.z) * k::TICK_TIME * k::UU_TO_BT, splat(0.0f), false, true);
  }
}
RS_FN void car_update_air_torque(ArenaState& s, const StaticData& w, int ci, Body& rb, uint32_t nwic, bool prev_flipping,
                                 float prev_flip_time) {
  Car& car = s.cars[ci];
  const BodyConst& bc = w.bodies[w.cars[ci].body_idx];
  bool allow_dodge = nwic < 3, allow_air = nwic == 0;
  if (car.is_flipping && allow_dodge && !is_zero(car.flip_rel_torque)) {
    V3 rel = car.flip_rel_torque;
    float pitch_scale = 1.0f;
    if (rel.y != 0.0f && car.controls.pitch != 0.0f && fsignum_rs(rel.y) == fsignum_rs(car.controls.pitch) &&
        prev_flip_time >= k::FLIP_PITCH_CANCEL_GATE_MIN_TIME)
      pitch_scale = 1.0f - fmin_rs(fabs_(quant_axis(fneg(car.controls.pitch))), 1.0f);
    rel.y = rel.y * pitch_scale;
    V3 torque = rel * V3{k::FLIP_TORQUE_X, k::FLIP_TORQUE_Y, k::FLIP_TORQUE_Z} * k::TICK_TIME;
    add_impulse(rb, bc, IMP_ANGULAR, rb.world.m * torque, splat(0.0f), false, true);
  }
  if (allow_air) {
    V3 in = quant_axes(V3{car.controls.pitch, car.controls.yaw, car.controls.roll});
    float pitch_torque_scale = 1.0f;
    if (!is_zero(in) && (prev_flipping || car.is_flipping || prev_flip_time < k::FLIP_PITCHLOCK_EXTRA_TIME)) pitch_torque_scale = 0.0f;
    M3 dirs = M3{-car.rot.c1, car.rot.c2, -car.rot.c0};
    V3 scaled = in * V3{pitch_torque_scale, 1.0f, 1.0f};
    V3 torque = !is_zero(in) ? dirs * (scaled * V3{k::AIR_TORQUE_X, k::AIR_TORQUE_Y, k::AIR_TORQUE_Z}) : splat(0.0f);
    V3 retain = splat(1.0f) - vabs(scaled);
    retain.z = 1.0f;
    V3 damping = dirs * (mul_t(dirs, car.ang_vel) * V3{k::AIR_DAMPING_X, k::AIR_DAMPING_Y, k::AIR_DAMPING_Z} * retain);
    add_impulse(rb, bc, IMP_ANGULAR, (torque - damping) * k::AIR_TORQUE_APPLY_SCALE * k::TICK_TIME, splat(0.0f), false, true);
  }
  if (allow_air) {
    float ts = car.is_boosting ? 1.0f : quant_axis(car.controls.throttle);
    if (ts != 0.0f)
      add_impulse(rb, bc, IMP_LINEAR, car.rot.c0 * ts * k::THROTTLE_AIR_ACCEL * k::UU_TO_BT * k::TICK_TIME, splat(0.0f), false, true);
  }
}
RS_FN void car_update_jump(ArenaState& s, const StaticData& w, int ci, Body& rb, bool jump_pressed) {
  Car& car = s.cars[ci];
  const BodyConst& bc = w.bodies[w.cars[ci].body_idx];
  V3 up = car.rot.c2;
  if (!car.has_jumped && car.is_on_ground && jump_pressed) {
    car.is_jumping = 1;
    car.has_jumped = 1;
    car.jump_ticks = 0;
    car.jump_hold_broken = 0;
  }
  car.jump_ticks += 1;
  if (car.is_jumping) {
    if (!car.controls.jump) car.jump_hold_broken = 1;
    car.is_jumping = (car.jump_ticks <= k::JUMP_MIN_TICKS ||
                      (!car.jump_hold_broken && car.controls.jump && car.jump_ticks <= k::JUMP_MAX_TICKS)) ? 1u : 0u;
    if (!car.is_jumping) car.jump_ticks = 1;
  }
  if (car.is_jumping) {
    if (car.jump_ticks == 1) {
      add_impulse(rb, bc, IMP_LINEAR, up * w.mut_f[MF_JUMP_IMMEDIATE_FORCE] * k::UU_TO_BT, splat(0.0f), false, false);
      limit_vels(rb, k::CAR_MAX_SPEED * k::UU_TO_BT, k::CAR_MAX_ANG_SPEED);
    }
    add_impulse(rb, bc, IMP_LINEAR, up * w.mut_f[MF_JUMP_ACCEL] * (k::UU_TO_BT * k::TICK_TIME), splat(0.0f), false, true);
  }
}
RS_FN void car_update_auto_flip(ArenaState& s, const StaticData& w, int ci, Body& rb, bool jump_pressed) {
  Car& car = s.cars[ci];
  const BodyConst& bc = w.bodies[w.cars[ci].body_idx];
  if (jump_pressed && car.world_contact_normal.some && car.world_contact_normal.v.z > k::AUTOFLIP_NORM_Z_THRESH) {
    float y_z = car.rot.c1.z, z_z = car.rot.c2.z;
    float cy = __builtin_sqrtf(z_z * z_z + y_z * y_z);
    float roll = cy > 16.0f * k::F32_EPS ? fatan2(y_z, z_z) : 0.0f;
    float abs_roll = fabs_(roll);
    if (abs_roll > k::AUTOFLIP_ROLL_THRESH) {
      car.auto_flip_timer = k::AUTOFLIP_TIME * (abs_roll / k::PI_F);
      car.auto_flip_torque_scale = autoflip_sign(w, car.rot);
      car.is_auto_flipping = 1;
      add_impulse(rb, bc, IMP_LINEAR, -car.rot.c2 * (k::AUTOFLIP_IMPULSE * k::UU_TO_BT), splat(0.0f), false, false);
    }
  }
  if (car.is_auto_flipping) {
    if (car.auto_flip_timer <= 0.0f) {
      car.is_auto_flipping = 0;
      car.auto_flip_timer = 0.0f;
    } else {
      rb.ang_vel += car.rot.c0 * k::AUTOFLIP_ANG_VEL_DELTA * car.auto_flip_torque_scale;
      limit_vels(rb, k::CAR_MAX_SPEED * k::UU_TO_BT, k::CAR_MAX_ANG_SPEED);
      car.auto_flip_timer = car.auto_flip_timer - k::TICK_TIME;
    }
  }
}
RS_FN bool car_update_double_jump_or_flip(ArenaState& s, const StaticData& w, int ci, Body& rb, bool jump_pressed, float fwd_speed,
                                          bool was_jumping) {
  Car& car = s.cars[ci];
  const BodyConst& bc = w.bodies[w.cars[ci].body_idx];
  if (car.is_on_ground) {
    car.has_double_jumped = 0;
    car.has_flipped = 0;
    car.air_time = 0.0f;
    car.air_time_since_jump = 0.0f;
    car.flip_time = car.flip_time + k::TICK_TIME;
    return false;
  }
  car.air_time = car.air_time + k::TICK_TIME;
  if (car.has_jumped && !was_jumping && !car.is_jumping) car.air_time_since_jump = car.air_time_since_jump + k::TICK_TIME;
  else car.air_time_since_jump = 0.0f;
  if (jump_pressed && car.air_time_since_jump < k::JUMP_DOUBLEJUMP_MAX_DELAY) {
    V3 dodge = quant_axes(V3{car.controls.pitch, car.controls.yaw, car.controls.roll} * V3{-1.0f, 1.0f, 1.0f});
    float mag = (fabs_(dodge.x) + fabs_(dodge.y)) + fabs_(dodge.z);
    bool flip_input = mag >= car.dodge_deadzone;
    bool unlimited = flip_input ? (w.unlimited_flips != 0) : (w.unlimited_double_jumps != 0);
    if ((!car.is_auto_flipping && !car.has_double_jumped && !car.has_flipped) || unlimited) {
      car.has_jumped = 1;
      if (flip_input) {
        car.flip_time = 0.0f;
        car.has_flipped = 1;
        car.is_flipping = 1;
        V3 dir = dodge;
        dir.y = dir.y + dir.z;
        if (fabs_(dir.x) < 0.1f && fabs_(dir.y) < 0.1f) {
          dir = splat(0.0f);
        } else {
          dir.z = 0.0f;
          dir = normalize_or_zero(dir);
        }
        V3 dz = vselect(cmplt(vabs(dir), splat(0.1f)), splat(0.0f), dir);
        if (length_sq(dz) > k::F32_EPS * k::F32_EPS) {
          car.flip_rel_torque = V3{fneg(dir.y), dir.x, 0.0f};
          bool backwards = fabs_(fwd_speed) < 100.0f ? (fbits(dz.x) >> 31) != 0 : fsignum_rs(dz.x) != fsignum_rs(fwd_speed);
          float max_scale_x = backwards ? k::FLIP_BACKWARD_IMPULSE_MAX_SPEED_SCALE : k::FLIP_FORWARD_IMPULSE_MAX_SPEED_SCALE;
          float ratio = fabs_(fwd_speed) / k::CAR_MAX_SPEED;
          V3 iv = dz * k::FLIP_INITIAL_VEL_SCALE;
          iv.x = iv.x * ((max_scale_x - 1.0f) * ratio + 1.0f);
          iv.y = iv.y * ((k::FLIP_SIDE_IMPULSE_MAX_SPEED_SCALE - 1.0f) * ratio + 1.0f);
          if (backwards) iv.x = iv.x * k::FLIP_BACKWARD_IMPULSE_SCALE_X;
          V3 fd = car.rot.c0;
          fd.z = 0.0f;
          V3 f2 = normalize_or_zero(fd);
          V3 r2 = V3{fneg(f2.y), f2.x, 0.0f};
          V3 dv = iv.x * f2 + iv.y * r2;
          add_impulse(rb, bc, IMP_LINEAR, dv * k::UU_TO_BT, splat(0.0f), false, false);
          limit_vels(rb, k::CAR_MAX_SPEED * k::UU_TO_BT, k::CAR_MAX_ANG_SPEED);
        }
      } else {
        add_impulse(rb, bc, IMP_LINEAR, car.rot.c2 * w.mut_f[MF_JUMP_IMMEDIATE_FORCE] * k::UU_TO_BT, splat(0.0f), false, false);
        limit_vels(rb, k::CAR_MAX_SPEED * k::UU_TO_BT, k::CAR_MAX_ANG_SPEED);
        car.has_double_jumped = 1;
      }
    }
  }
  if (car.is_flipping) {
    float ft = car.flip_time;
    bool still = car.has_flipped && ft < k::FLIP_TORQUE_TIME;
    car.is_flipping = still ? 1u : 0u;
    car.flip_time = still ? ft + k::TICK_TIME : k::TICK_TIME;
    return !still;
  }
  car.flip_time = car.flip_time + k::TICK_TIME;
  return false;
}
RS_FN void car_update_auto_roll(ArenaState& s, const StaticData& w, int ci, Body& rb, uint32_t nwic, bool* up_some, V3* up_dir) {
  Car& car = s.cars[ci];
  const BodyConst& bc = w.bodies[w.cars[ci].body_idx];
  if (!car.sticky_gate_prev && !car.world_contact_normal.some) return;
  V3 gup;
  if (car.sticky_gate_prev && nwic > 0) {
    if (*up_some) {
      gup = *up_dir;
    } else {
      gup = upwards_from_wheels(car, rb.world.m.c2);
      *up_dir = gup;
      *up_some = true;
    }
  } else {
    gup = car.world_contact_normal.v;
  }
  V3 gdown = -gup;
  V3 fwd = car.rot.c0, right = car.rot.c1;
  V3 cr = cross(gup, fwd);
  V3 cf = cross(gdown, cr);
  float rtf = 1.0f - fclamp_rs(dot(right, cr), 0.0f, 1.0f);
  float ftf = 1.0f - fclamp_rs(dot(fwd, cf), 0.0f, 1.0f);
  V3 tdr = fwd * fneg(fsignum_rs(dot(right, gup)));
  V3 tdf = right * fsignum_rs(dot(fwd, gup));
  V3 tr = tdr * rtf, tf = tdf * ftf;
  add_impulse(rb, bc, IMP_LINEAR, gdown * (k::AUTOROLL_FORCE * k::UU_TO_BT * k::TICK_TIME), splat(0.0f), false, true);
  add_impulse(rb, bc, IMP_ANGULAR, (tf + tr) * (k::AUTOROLL_TORQUE * k::TICK_TIME), splat(0.0f), false, true);
}
RS_FN void car_update_boost(ArenaState& s, const StaticData& w, int ci, Body& rb) {
  Car& car = s.cars[ci];
  const BodyConst& bc = w.bodies[w.cars[ci].body_idx];
  const float ag = w.mut_f[MF_BOOST_ACCEL_GROUND], aa = w.mut_f[MF_BOOST_ACCEL_AIR];
  if (car.is_boosting) {
    car.boost = fmax_rs(car.boost - w.mut_f[MF_BOOST_USED_PER_SECOND] * k::TICK_TIME, 0.0f);
    bool depleted = car.boost == 0.0f;
    bool expired = !car.controls.boost && car.boosting_time >= k::BOOST_MIN_TIME;
    if (depleted || expired) {
      car.is_boosting = 0;
      car.boosting_time = 0.0f;
      car.time_since_boosted = car.time_since_boosted + k::TICK_TIME;
      if (w.recharge_boost_enabled && car.time_since_boosted >= w.recharge_boost_delay)
        car.boost = car.boost + w.recharge_boost_per_second * k::TICK_TIME;
    } else {
      car.is_boosting = 1;
      car.boosting_time = car.boosting_time + k::TICK_TIME;
      car.time_since_boosted = 0.0f;
      float acc = car.is_on_ground ? ag : aa;
      add_impulse(rb, bc, IMP_LINEAR, acc * car.rot.c0 * (k::UU_TO_BT * k::TICK_TIME), splat(0.0f), false, true);
    }
  } else if (car.controls.boost && car.boost > 0.0f) {
    car.is_boosting = 1;
    car.boosting_time = car.boosting_time + k::TICK_TIME;
    car.time_since_boosted = 0.0f;
    float acc = car.is_on_ground ? ag : aa;
    add_impulse(rb, bc, IMP_LINEAR, acc * car.rot.c0 * (k::UU_TO_BT * k::TICK_TIME), splat(0.0f), false, true);
  } else {
    car.is_boosting = 0;
    car.boosting_time = 0.0f;
    car.time_since_boosted = car.time_since_boosted + k::TICK_TIME;
    if (w.recharge_boost_enabled && car.time_since_boosted >= w.recharge_boost_delay)
      car.boost = car.boost + w.recharge_boost_per_second * k::TICK_TIME;
  }
  car.boost = fclamp_rs(car.boost, 0.0f, w.mut_f[MF_CAR_MAX_BOOST_AMOUNT]);
}
RS_FN void car_update_contact_state(ArenaState& s, const StaticData& w, int ci) {
  Car& car = s.cars[ci];
  uint32_t n = 0;
  for (int i = 0; i < NUM_WHEELS; ++i) {
    const Wheel& wh = car.wheels[i];
    Opt<HitInfo> hit = Opt<HitInfo>{};
    if (wh.ray.some) {
      const RaycastInfo& ri = wh.ray.v;
      float trace_len = length(wh.hard_point - ri.contact_point);
      hit.some = 1;
      hit.v.hit_point = ri.contact_point * k::BT_TO_UU;
      hit.v.hit_normal = ri.contact_normal;
      hit.v.hit_fraction = trace_len / w.cars[ci].wheels[i].ray_length;
      hit.v.user_info = w.bodies[ri.ground_body].user_idx;
    }
    car.wheels_with_contact[i] = hit;
    n += hit.some ? 1u : 0u;
  }
  car.is_on_ground = n >= 3 ? 1u : 0u;
}
RS_FN void car_flip_zdamp(ArenaState& s, const StaticData& w, int ci) {  // apply_flip_zdamp_prepass
  Car& car = s.cars[ci];
  if (car.is_demoed || car.is_on_ground || !car.is_flipping) return;
  if (!(car.flip_time >= k::FLIP_Z_DAMP_START && car.flip_time <= k::FLIP_TORQUE_TIME)) return;
  Body& rb = s.bodies[dyn_slot(w, w.cars[ci].body_idx)];
  if (rb.lin_vel.z < 0.0f || car.flip_time < k::FLIP_Z_DAMP_END) rb.lin_vel.z = rb.lin_vel.z * (1.0f - k::FLIP_Z_DAMP_120);
}
RS_INLINE void car_pre_tick(ArenaState& s, const StaticData& w, int ci, const WheelRayPrep& prep) {  // car/base.rs pre_tick_update (car_pre_tick.cuh:638-703)
  Car& car = s.cars[ci];
  const CarConst& cc = w.cars[ci];
  Body& rb = s.bodies[dyn_slot(w, cc.body_idx)];
  if (car.is_demoed) {
    car.demo_respawn_timer = fmax_rs(car.demo_respawn_timer - k::TICK_TIME, 0.0f);
    if (car.demo_respawn_timer == 0.0f) car_respawn(s, w, ci, w.mut_f[MF_CAR_SPAWN_BOOST_AMOUNT]);
    if (rb.activation != k::ACT_DISABLED) rb.activation = k::ACT_DISABLED;
    rb.collision_flags |= k::CF_NO_CONTACT_RESPONSE;
    return;
  }
  rb.activation = k::ACT_ACTIVE;  // force_activate
  rb.deactivation_time = 0.0f;
  rb.collision_flags &= ~k::CF_NO_CONTACT_RESPONSE;
  car.controls.throttle = fclamp_rs(car.controls.throttle, -1.0f, 1.0f);
  car.controls.steer = fclamp_rs(car.controls.steer, -1.0f, 1.0f);
  car.controls.pitch = fclamp_rs(car.controls.pitch, -1.0f, 1.0f);
  car.controls.yaw = fclamp_rs(car.controls.yaw, -1.0f, 1.0f);
  car.controls.roll = fclamp_rs(car.controls.roll, -1.0f, 1.0f);
  float fwd_speed = dot(rb.lin_vel, rb.world.m.c0) * k::BT_TO_UU;
  bool jump_pressed = car.controls.jump && !car.prev_controls.jump;
  uint32_t nwic = wheels_in_contact(car);
  bool up_some = false;
  V3 up_dir = splat(0.0f);
  car_update_wheels(s, w, ci, rb, fwd_speed, &up_some, &up_dir);
  if (car.is_on_ground) {
    if (car.is_flipping) car.flip_time = 0.0f;
    car.is_flipping = 0;
  }
  bool was_jumping = car.is_jumping != 0;
  car_update_jump(s, w, ci, rb, jump_pressed);
  car_update_auto_flip(s, w, ci, rb, jump_pressed);
  bool prev_flipping = car.is_flipping != 0;
  float prev_flip_time = car.flip_time;
  bool flip_ended = car_update_double_jump_or_flip(s, w, ci, rb, jump_pressed, fwd_speed, was_jumping);
  if (!car.is_on_ground) car_update_air_torque(s, w, ci, rb, nwic, prev_flipping, prev_flip_time);
  float quant_throttle = quant_axis(car.controls.throttle);
  if (quant_throttle != 0.0f && !car.is_flipping && !flip_ended && ((0 < nwic && nwic < 4) || car.world_contact_normal.some))
    car_update_auto_roll(s, w, ci, rb, nwic, &up_some, &up_dir);
  car.world_contact_normal = Opt<V3>{};
  car_update_boost(s, w, ci, rb);
  float real_throttle = (car.controls.boost && car.boost > 0.0f) ? 1.0f : quant_throttle;
  vehicle_update(s, w, ci, k::TICK_TIME, car.handbrake_val, real_throttle, cc.three_wheels != 0, prep);
  car_update_contact_state(s, w, ci);
  uint32_t any_world = 0;
  for (int i = 0; i < NUM_WHEELS; ++i)
    if (car.wheels[i].ray.some && car.wheels[i].ray.v.in_contact_with_world) any_world = 1;
  car.sticky_gate_prev = any_world;
}

// P1 (arena_base_p1.cuh:11-20): the z-damp prepass for every car first, then each car's pre-tick; CarLanded events.
// A car's pre-tick as a function of its own, called once per car: the boundary lets the register allocator treat the
// car's update apart from the rest of the phase (measured faster than inlining it). Its state and world pointers carry
// their address spaces, so its accesses compile to scratch_ and global_ instructions.
// UW: every lane steps in the same world (its pointer is then made wave-uniform).
template <bool UW>
RS_NOINL void car_pre_tick_in(RS_PRIVATE_AS ArenaState* sp, const RS_CONST_AS StaticData* wp, int ci,
                              const RS_PRIVATE_AS WheelRayPrep* pp) {
  if (UW) wp = rs_uniform_ptr(wp);
  RS_UNIFORM_INDEX(ci);
  car_pre_tick(*(ArenaState*)sp, *(const StaticData*)wp, ci, *(const WheelRayPrep*)pp);
}
// Both cars' wheel-ray screens and mesh walks, before either car's pre-tick (exact: see WheelRayPrep). A demoed car casts
// no rays this tick (its pre-tick respawns it at most); its share is empty.
RS_FN void wheel_rays_prepare(const ArenaState& s, const StaticData& w, WheelRayPrep& prep) {
  prep.tb.n_jobs = prep.tb.n_leaves = prep.tb.next = 0;
  prep.tb.leaves = 1;
  for (int ci = 0; ci < NUM_CARS; ++ci) {
    prep.ncand[ci] = 0;
    prep.first_job[ci] = prep.tb.n_jobs;
    if (!s.cars[ci].is_demoed) {
      const CarConst& cc = w.cars[ci];
      const Aff ct = s.bodies[dyn_slot(w, cc.body_idx)].world;
      V3 src[NUM_WHEELS], dst[NUM_WHEELS];
      for (int i = 0; i < NUM_WHEELS; ++i) wheel_ray_ends(cc.wheels[i], ct, &src[i], &dst[i]);
      ClosestHits hits;
      wheel_hits_init(hits, src, dst, cc.body_idx);
      prep.ncand[ci] = ray_screen(s, w, hits, prep.cand[ci]);
      ray_jobs(s, w, hits, prep.cand[ci], prep.ncand[ci], prep.tb);
    }
    prep.end_job[ci] = prep.tb.n_jobs;
  }
  trav_run(w, prep.tb);
}
template <bool UW = false>
RS_INLINE void phase_p1(ArenaState& s, const StaticData& w) {
  for (int ci = 0; ci < NUM_CARS; ++ci) car_flip_zdamp(s, w, ci);
  WheelRayPrep prep;
  wheel_rays_prepare(s, w, prep);
  for (int ci = 0; ci < NUM_CARS; ++ci) {
    uint32_t was = s.cars[ci].is_on_ground;
    car_pre_tick_in<UW>((RS_PRIVATE_AS ArenaState*)&s, (const RS_CONST_AS StaticData*)&w, ci, (const RS_PRIVATE_AS WheelRayPrep*)&prep);
    if (!was && s.cars[ci].is_on_ground) s.events |= 1u << (EV_LANDED + ci);
  }
}

}  // namespace rs
