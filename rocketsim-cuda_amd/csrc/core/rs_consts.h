// Constants of the tick (RocketSim V3 sim/consts.rs as the reference restates them in cuda/sim/consts.cuh:9-100 and
// bullet_consts.cuh). Constant expressions are evaluated in binary32 by the compiler (IEEE round to nearest), exactly as
// the reference's are, so each constant is the same float.
#pragma once
#include "rs_state.h"
#include "rs_crmath.h"

namespace rs {
namespace k {

constexpr float BT_TO_UU = 50.0f;
constexpr float UU_TO_BT = 1.0f / 50.0f;
constexpr float TICK_RATE = 120.0f;
constexpr float TICK_TIME = 1.0f / TICK_RATE;

constexpr float CAR_MAX_SPEED = 2300.0f;
constexpr float CAR_MAX_ANG_SPEED = 5.5f;
constexpr float SUPERSONIC_START_SPEED = 2200.0f;
constexpr float SUPERSONIC_MAINTAIN_MIN_SPEED = 2200.0f - 100.0f;
constexpr float SUPERSONIC_MAINTAIN_MAX_TIME = 1.0f;
constexpr float DEMO_YAW_LIMIT_DEG = 45.572994f;
constexpr float DEMO_PITCH_LIMIT_DEG = 36.869896f;
constexpr uint32_t JUMP_SETTLED_REARM_TICKS = 15;
constexpr float BUMP_YAW_LIMIT_DEG = 70.0f;
constexpr float BUMP_PITCH_LIMIT_DEG = 36.869896f;
constexpr float BALL_MAX_ANG_SPEED = 6.0f;
constexpr float HIT_FORWARD_SCALE = 0.65f;
constexpr float HIT_MAX_DELTA_VEL_UU = 4600.0f;
constexpr float HIT_Z_SCALE_NORMAL = 0.35f;
constexpr float BUMP_ATTACKER_SPEED_MAX = 2200.0f;
constexpr float CYL_HEIGHT = 95.0f;
constexpr float PAD_BOX_HEIGHT = 64.0f;

constexpr float QUANT_POS_SCALE = 100.0f;
constexpr float QUANT_VEL_SCALE = 100.0f;
constexpr float QUANT_ANG_VEL_SCALE = 10000.0f;
constexpr float QUANT_OFFSET_CORRECT_FRAC = 0.1f;

constexpr float PI_F = 3.14159265358979323846264338327950288f;
constexpr float TAU_F = 6.28318530717958647692f;
constexpr float F32_EPS = 1.1920928955078125e-7f;
constexpr float F32_MAX = 3.40282347e+38f;

constexpr float BOOST_MAX = 100.0f;
constexpr float BOOST_MIN_TIME = 0.1f;
constexpr float THROTTLE_TORQUE_AMOUNT = 180.0f * 400.0f;
constexpr float BRAKE_TORQUE_AMOUNT = 180.0f * (14.25f + (1.0f / 3.0f));
constexpr float STOPPING_FORWARD_VEL = 25.0f;
constexpr float COASTING_BRAKE_FACTOR = 0.15f;
constexpr float BRAKING_NO_THROTTLE_SPEED_THRESH = 0.01f;
constexpr float THROTTLE_DEADZONE = 0.01f;
constexpr float THROTTLE_AIR_ACCEL = 200.0f / 3.0f;
constexpr float POWERSLIDE_RISE_RATE = 5.0f;
constexpr float POWERSLIDE_FALL_RATE = 2.0f;
constexpr uint32_t JUMP_MIN_TICKS = 3;
constexpr uint32_t JUMP_MAX_TICKS = 24;
constexpr float JUMP_DOUBLEJUMP_MAX_DELAY = 1.25f;
constexpr float FLIP_Z_DAMP_120 = 0.35f;
constexpr float FLIP_Z_DAMP_START = 0.15f;
constexpr float FLIP_Z_DAMP_END = 0.21f;
constexpr float FLIP_TORQUE_TIME = 0.65f;
constexpr float FLIP_PITCH_CANCEL_GATE_MIN_TIME = 0.041f;
constexpr float FLIP_PITCHLOCK_EXTRA_TIME = 0.3f;
constexpr float FLIP_INITIAL_VEL_SCALE = 500.0f;
constexpr float FLIP_TORQUE_X = 260.0f, FLIP_TORQUE_Y = 224.0f, FLIP_TORQUE_Z = 0.0f;
constexpr float FLIP_FORWARD_IMPULSE_MAX_SPEED_SCALE = 1.0f;
constexpr float FLIP_SIDE_IMPULSE_MAX_SPEED_SCALE = 1.9f;
constexpr float FLIP_BACKWARD_IMPULSE_MAX_SPEED_SCALE = 2.5f;
constexpr float FLIP_BACKWARD_IMPULSE_SCALE_X = 16.0f / 15.0f;
constexpr float AIR_TORQUE_X = 130.0f, AIR_TORQUE_Y = 95.0f, AIR_TORQUE_Z = 400.0f;
constexpr float AIR_DAMPING_X = 30.0f, AIR_DAMPING_Y = 20.0f, AIR_DAMPING_Z = 50.0f;
constexpr float AIR_TORQUE_APPLY_SCALE = 2.0f * PI_F / 65536.0f * 1000.0f;
constexpr float AUTOFLIP_IMPULSE = 200.0f;
constexpr float AUTOFLIP_ANG_VEL_DELTA = 1000.0f;
constexpr float AUTOFLIP_TIME = 0.4f;
constexpr float AUTOFLIP_NORM_Z_THRESH = 0.707106781186547524400844362104849039f;
constexpr float AUTOFLIP_ROLL_THRESH = 2.8f;
constexpr float AUTOROLL_FORCE = 100.0f;
constexpr float AUTOROLL_TORQUE = 80.0f;
constexpr float RESPAWN_Z = 36.0f;
constexpr float SUSPENSION_STIFFNESS = 500.0f;
constexpr float WHEELS_DAMPING_COMPRESSION = 25.0f;
constexpr float WHEELS_DAMPING_RELAXATION = 40.0f;
constexpr float MAX_SUSPENSION_TRAVEL = 12.0f;
constexpr float SUSPENSION_SUBTRACTION = 0.05f;
constexpr float WHEEL_PUSHBACK_ERP = 0.1f;

// Bullet
constexpr uint32_t CF_STATIC = 1u, CF_CUSTOM_MATERIAL = 1u << 1, CF_CAN_SLEEP = 1u << 2, CF_NO_CONTACT_RESPONSE = 1u << 3,
                   CF_NO_ANGULAR_MOTION = 1u << 5;
constexpr float CONTACT_BREAKING_THRESHOLD = 0.02f;
constexpr float SPHERE_RADIUS_MARGIN = 0.08f;
constexpr uint32_t ACT_ACTIVE = 0, ACT_SLEEPING = 1, ACT_DISABLED = 2, ACT_WANTS_DEACTIVATION = 3;
constexpr uint32_t USER_NONE = 0, USER_CAR = 1, USER_BALL = 2;
constexpr uint32_t SHAPE_MESH = 0, SHAPE_PLANE = 1, SHAPE_SPHERE = 2, SHAPE_COMPOUND = 3;
constexpr uint32_t DEMO_NORMAL = 0, DEMO_ON_CONTACT = 1, DEMO_DISABLED = 2;
constexpr uint32_t FILTER_DEFAULT = 1u, FILTER_STATIC = 1u << 1, FILTER_DROPSHOT_FLOOR = 16u, FILTER_ALL = 0xFFu;

}  // namespace k

// The transcendental calls of the tick (libm.cuh:40-69), correctly rounded, so that every build agrees bit for bit. Two
// host-only builds of the oracle choose otherwise, to attribute differences from V3 to the libm (libm.cuh:7-12):
// RS_LIBM_PLATFORM makes all seven the platform C library as V3's Rust reaches it (sinf, cosf, tanf, atan2f, powf, atan2,
// hypot); RS_LIBM_MIXED chooses per call from the thread-local mask libm_platform_mask (bit 0 sin, 1 cos, 2 tan, 3 atan2,
// 4 pow, 5 atan2_64, 6 hypot64; a set bit is the platform library). The product defines neither.
#if defined(RS_LIBM_MIXED) && defined(RS_LIBM_PLATFORM)
#error "define at most one of RS_LIBM_MIXED and RS_LIBM_PLATFORM"
#endif
#if (defined(RS_LIBM_MIXED) || defined(RS_LIBM_PLATFORM)) && defined(__HIP_DEVICE_COMPILE__)
#error "the platform-libm builds are host-only"
#endif
#if defined(RS_LIBM_MIXED)
inline thread_local uint32_t libm_platform_mask = 0;
#define RS_LIBM_PICK(bit, platform, cr) ((libm_platform_mask & (bit)) ? (platform) : (cr))
#elif defined(RS_LIBM_PLATFORM)
#define RS_LIBM_PICK(bit, platform, cr) (platform)
#else
#define RS_LIBM_PICK(bit, platform, cr) (cr)
#endif
RS_FN float fsin(float x) { return RS_LIBM_PICK(1u, ::sinf(x), crm::sinf_cr(x)); }
RS_FN float fcos(float x) { return RS_LIBM_PICK(2u, ::cosf(x), crm::cosf_cr(x)); }
RS_FN float ftan(float x) { return RS_LIBM_PICK(4u, ::tanf(x), crm::tanf_cr(x)); }
RS_FN float fatan2(float y, float x) { return RS_LIBM_PICK(8u, ::atan2f(y, x), crm::atan2f_cr(y, x)); }
RS_FN float fpow(float x, float y) { return RS_LIBM_PICK(16u, ::powf(x, y), crm::powf_cr(x, y)); }
// The f64 pair of the auto-flip path (car_pre_tick.cuh:23-72): a double-double atan2 and the correctly rounded hypot.
RS_FN double datan2(double y, double x) { return RS_LIBM_PICK(32u, ::atan2(y, x), crm::atan2_64(y, x)); }
RS_FN double dhypot(double x, double y) { return RS_LIBM_PICK(64u, ::hypot(x, y), crm::hypot64(x, y)); }
// f32::powi is LLVM's square-and-multiply (libm.cuh:54-65).
RS_FN float fpowi(float a, int b) {
  unsigned p = b < 0 ? 0u - (unsigned)b : (unsigned)b;
  float r = 1.0f;
  for (;;) {
    if (p & 1u) r = r * a;
    p >>= 1;
    if (p == 0) break;
    a = a * a;
  }
  return b < 0 ? 1.0f / r : r;
}

}  // namespace rs
