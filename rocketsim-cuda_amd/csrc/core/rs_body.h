// Rigid-body steps of the tick: the quaternion integrator, damping, inertia, activation, velocity limits and impulses
// (reference cuda/sim/rigid_body.cuh, rigid_body_p1.cuh, transform_util.cuh), the UE3 quantisation of P0 and the control
// quantisation (quantize.cuh), and the piecewise-linear curves (linear_piece_curve.cuh).
#pragma once
#include "rs_consts.h"

namespace rs {

// ---- transform_util.cuh:18-41 integrate_trans -------------------------------------------------------------------------------
RS_FN void integrate_trans(Aff* tr, Quat* rot, V3 lin_vel, V3 ang_vel, float dt) {
  tr->t = tr->t + lin_vel * dt;
  const float thresh = ffrom(0x3F490FDBu);  // FRAC_PI_4
  float angle = length(ang_vel);
  if (angle * dt > thresh) angle = thresh / dt;
  float half_angle = (angle * dt) * 0.5f;
  float k;
  if (angle < 0.001f) {
    k = 0.5f * dt - (((dt * dt) * dt) * 0.020833334f * angle) * angle;  // left to right
  } else {
    k = fsin(half_angle) / angle;
  }
  V3 axis = ang_vel * k;
  Quat dorn = Quat{axis.x, axis.y, axis.z, fcos(half_angle)};
  *rot = qnormalize(dorn * *rot);
  tr->m = m3_from_quat(*rot);
}

// ---- rigid_body.cuh ---------------------------------------------------------------------------------------------------------
RS_FN bool body_active(const Body& b) { return !(b.activation == k::ACT_SLEEPING || b.activation == k::ACT_DISABLED); }
RS_FN void set_activation(Body& b, uint32_t st) {
  if (b.activation != k::ACT_DISABLED) b.activation = st;
}
RS_FN void set_world_trans(Body& b, const Aff& t) {  // :29-33
  if ((b.collision_flags & k::CF_NO_ANGULAR_MOTION) == 0) b.world_quat = quat_from_m3(t.m);
  b.world = t;
}
RS_FN M3 inertia_world(const M3& m, V3 inv_inertia_local) {  // :36-42 m * diag(inv) * m^T
  M3 s = transpose(m);
  s.c0 *= inv_inertia_local;
  s.c1 *= inv_inertia_local;
  s.c2 *= inv_inertia_local;
  return m * s;
}
RS_FN void update_inertia(Body& b, const BodyConst& c) { b.inv_inertia_world = inertia_world(b.world.m, c.inv_inertia_local); }
RS_FN void apply_damping(Body& b, const BodyConst& c, float dt) {  // :61-70: the cached factor is reused while valid
  if (c.linear_damping != 0.0f) {
    if (b.cached_damping_time_step != dt || b.cached_linear_damping_value != c.linear_damping) {
      b.cached_linear_damping = fpow(1.0f - c.linear_damping, dt);
      b.cached_damping_time_step = dt;
      b.cached_linear_damping_value = c.linear_damping;
    }
    b.lin_vel *= b.cached_linear_damping;
  }
}
// inlined everywhere: a call would save and restore the caller's registers through private memory
RS_INLINE Aff predict_trans(const Body& b, float dt) {  // :73-82
  Aff t = b.world;
  Quat q = b.world_quat;
  if ((b.collision_flags & k::CF_NO_ANGULAR_MOTION) != 0) t.t = t.t + b.lin_vel * dt;
  else integrate_trans(&t, &q, b.lin_vel, b.ang_vel, dt);
  return t;
}
RS_FN void set_com_trans(Body& b, const BodyConst& c, const Aff& t) {  // :85-89
  b.interp_world = t;
  set_world_trans(b, t);
  update_inertia(b, c);
}
RS_FN void update_activation(Body& b, const BodyConst& c, float dt) {  // :92-108
  if ((b.collision_flags & k::CF_CAN_SLEEP) == 0) {
    set_activation(b, k::ACT_ACTIVE);
    return;
  }
  if (b.activation == k::ACT_SLEEPING) return;
  float lin2 = c.linear_sleeping_threshold * c.linear_sleeping_threshold;
  float ang2 = c.angular_sleeping_threshold * c.angular_sleeping_threshold;
  if (length_sq(b.lin_vel) < lin2 && length_sq(b.ang_vel) < ang2) {
    b.deactivation_time = b.deactivation_time + dt;
    if (b.deactivation_time > 2.0f && b.activation == k::ACT_ACTIVE) set_activation(b, k::ACT_WANTS_DEACTIVATION);
  } else {
    b.deactivation_time = 0.0f;
    set_activation(b, k::ACT_ACTIVE);
  }
}
RS_FN void body_sleep(Body& b) {
  set_activation(b, k::ACT_SLEEPING);
  b.lin_vel = splat(0.0f);
  b.ang_vel = splat(0.0f);
}
RS_FN void limit_vels(Body& b, float max_lin, float max_ang) {  // :125-132
  if (length_sq(b.lin_vel) > fpowi(max_lin, 2)) b.lin_vel = normalize_or_zero(b.lin_vel) * max_lin;
  if (length_sq(b.ang_vel) > fpowi(max_ang, 2)) b.ang_vel = normalize_or_zero(b.ang_vel) * max_ang;
}
RS_FN void clear_accum(Body& b) {
  b.accum_lin_vel = splat(0.0f);
  b.accum_ang_vel = splat(0.0f);
}
// add_impulse(Linear(v), massed = false): v * 1.0 into lin_vel (accum: the accumulators), ZERO into the angular one (:45-50,
// :136-142: the zero is still added, which turns a -0 lane into +0).
RS_FN void add_lin_impulse(Body& b, V3 v, bool accum) {
  V3 lin = v * 1.0f, ang = splat(0.0f);
  if (accum) {
    b.accum_lin_vel += lin;
    b.accum_ang_vel += ang;
  } else {
    b.lin_vel += lin;
    b.ang_vel += ang;
  }
}

// add_impulse (rigid_body_p1.cuh:79-103) with a run-time kind.
enum ImpulseKind { IMP_LINEAR, IMP_LINEAR_REL_POS, IMP_ANGULAR };
RS_FN void add_impulse(Body& b, const BodyConst& c, ImpulseKind kind, V3 v, V3 rel_pos, bool massed, bool accum) {
  V3 lin = splat(0.0f), ang = splat(0.0f);
  float sc = massed ? c.inv_mass : 1.0f;
  if (kind == IMP_LINEAR) {
    lin = v * sc;
  } else if (kind == IMP_LINEAR_REL_POS) {
    lin = v * sc;
    ang = b.inv_inertia_world * cross(rel_pos, v);
    if (!massed) ang *= c.mass;
  } else {
    ang = v * sc;
  }
  if (accum) {
    b.accum_lin_vel += lin;
    b.accum_ang_vel += ang;
  } else {
    b.lin_vel += lin;
    b.ang_vel += ang;
  }
}

// ---- quantize.cuh ---------------------------------------------------------------------------------------------------------
// sse2 m128_floor per lane: lanes below 2^23 become trunc(v) - (trunc(v) > v) (floor(-0) = +0); others pass (:16-23).
RS_FN float sse_floor(float v) {
  if ((int32_t)(fbits(v) & 0x7FFFFFFFu) < (int32_t)0x4B000000) {
    float t = (float)f2i_rs(v);
    return t + ((t > v) ? -1.0f : 0.0f);
  }
  return v;
}
RS_FN V3 sign3(V3 v) { return vselect(cmpeq(v, splat(0.0f)), splat(0.0f), vsignum(v)); }  // :27-30
RS_FN V3 quant_pos(V3 v, float scale) {  // :33-35
  V3 a = v * scale + splat(0.5f);
  return V3{sse_floor(a.x), sse_floor(a.y), sse_floor(a.z)} / scale;
}
RS_FN V3 quant_vel(V3 v, float scale) {  // :37-47
  float inv = 1.0f / scale;
  V3 sc = v * scale;
  V3 r = V3{(float)f2i_rs(sc.x) * inv, (float)f2i_rs(sc.y) * inv, (float)f2i_rs(sc.z) * inv};
  float off = k::QUANT_OFFSET_CORRECT_FRAC * inv;
  return r + sign3(r) * off;
}
RS_FN void quantize_body(V3* pos, V3* lin_vel, V3* ang_vel) {  // :50-57
  V3 np = quant_pos(*pos * k::BT_TO_UU, k::QUANT_POS_SCALE) * k::UU_TO_BT;
  V3 nv = quant_vel(*lin_vel * k::BT_TO_UU, k::QUANT_VEL_SCALE) * k::UU_TO_BT;
  V3 na = quant_vel(*ang_vel, k::QUANT_ANG_VEL_SCALE);
  *pos = np;
  *lin_vel = nv;
  *ang_vel = na;
}
// sse2 m128_round_ties_even and m128_round (:61-78)
RS_FN float sse_round_even(float v) {
  uint32_t sign = fbits(v) & 0x80000000u;
  float magic = ffrom(fbits(8388608.0f) | sign);
  float r1 = (v + magic) - magic;
  bool small = fabs_(v) <= 8388608.0f;
  return ffrom((small ? fbits(r1) : 0u) ^ (small ? 0u : fbits(v)));
}
RS_FN float sse_round(float v) {
  float r = sse_round_even(v);
  uint32_t sign = fbits(v) & 0x80000000u;
  bool tie_down = (v - r) == ffrom(fbits(0.5f) | sign);
  return r + (tie_down ? ffrom(fbits(1.0f) | sign) : 0.0f);
}
// quantize_axis_inputs: controls to the 8-bit wire format and back, three lanes (:81-92)
RS_FN V3 quant_axes(V3 c) {
  V3 cl = vclamp(c, splat(-1.0f), splat(1.0f));
  V3 sc = vselect(cmplt(cl, splat(0.0f)), splat(128.0f), splat(127.0f));
  V3 biased = cl * sc + splat(128.0f);
  V3 w = biased + biased + splat(0.5f);
  V3 r = V3{sse_round(w.x), sse_round(w.y), sse_round(w.z)};
  V3 byte = V3{(float)((f2i_rs(r.x) >> 1) & 0xFF), (float)((f2i_rs(r.y) >> 1) & 0xFF), (float)((f2i_rs(r.z) >> 1) & 0xFF)};
  V3 s = byte - splat(128.0f);
  return vselect(cmplt(s, splat(0.0f)), s * splat(1.0f / 128.0f), s / splat(127.0f));
}
// quantize_axis_input, one lane (:95-104)
RS_FN float quant_axis(float x) {
  float cl = fclamp_rs(x, -1.0f, 1.0f);
  float y = cl < 0.0f ? fmax_rs(cl * 128.0f, -128.0f) : fmin_rs(cl * 127.0f, 127.0f);
  float w = ((y + 128.0f) + (y + 128.0f)) + 0.5f;
  uint32_t byte = (uint32_t)((f2i_rs(sse_round_even(w)) >> 1) & 0xFF);
  float s = (float)byte - 128.0f;
  if (byte < 0x80u) return s * (1.0f / 128.0f);
  return s / 127.0f;
}

// ---- linear_piece_curve.cuh -------------------------------------------------------------------------------------------------
RS_FN float curve(const float (*p)[2], int n, float x) {  // :11-26
  if (n == 1) return p[0][1];
  if (x <= p[0][0]) return p[0][1];
  int i = 1;
  while (i < n && !(p[i][0] > x)) ++i;
  if (i == n) return p[n - 1][1];
  float frac = (x - p[i - 1][0]) / (p[i][0] - p[i - 1][0]);
  return (p[i][1] - p[i - 1][1]) * frac + p[i - 1][1];
}
#define RS_CURVE(name, N, ...)                        \
  RS_FN float name(float x) {                         \
    const float p[N][2] = __VA_ARGS__;                \
    return curve(p, N, x);                            \
  }
RS_CURVE(curve_ball_car_extra_impulse, 5, {{0.f, 0.65f}, {500.f, 0.65f}, {1400.f, 0.60f}, {2300.f, 0.55f}, {4600.f, 0.30f}})
RS_CURVE(curve_bump_vel_ground, 3, {{0.f, 150.0f / 180.0f}, {1400.f, 200000.0f / 180.0f}, {2200.f, 275000.0f / 180.0f}})
RS_CURVE(curve_bump_vel_air, 3, {{0.f, 150.0f / 180.0f}, {1400.f, 250000.0f / 180.0f}, {2200.f, 350000.0f / 180.0f}})
RS_CURVE(curve_bump_upward_vel, 3, {{0.f, 50.0f / 180.0f}, {1400.f, 50000.0f / 180.0f}, {2200.f, 75000.0f / 180.0f}})
RS_CURVE(curve_steer_angle, 6,
         {{0.f, 0.53356f}, {500.f, 0.31930f}, {1000.f, 0.18203f}, {1500.f, 0.10570f}, {1750.f, 0.08507f}, {3000.f, 0.03454f}})
RS_CURVE(curve_steer_angle_3w, 2, {{0.f, 0.342473f}, {2300.f, 0.034837f}})
RS_CURVE(curve_powerslide_steer, 2, {{0.f, 0.39235f}, {2500.f, 0.12610f}})
RS_CURVE(curve_drive_torque, 3, {{0.f, 1.0f}, {1400.f, 0.1f}, {1410.f, 0.0f}})
RS_CURVE(curve_non_sticky_friction, 3, {{0.f, 0.1f}, {0.7075f, 0.5f}, {1.0f, 1.0f}})
RS_CURVE(curve_lat_friction, 2, {{0.f, 1.0f}, {1.0f, 0.2f}})
RS_CURVE(curve_lat_friction_3w, 2, {{0.f, 0.30f}, {1.0f, 0.25f}})
RS_CURVE(curve_handbrake_lat, 1, {{0.f, 0.1f}})
RS_CURVE(curve_handbrake_long, 2, {{0.f, 0.5f}, {1.0f, 0.9f}})
#undef RS_CURVE

}  // namespace rs
