// Vector, matrix, quaternion and affine arithmetic with the rounding sequence of glam 0.34.1's x86-64 SSE2 paths, which
// RocketSim V3 uses and the reference restates op by op in cuda/common/glam.cuh. Each function below is a fixed sequence of
// IEEE single operations; the sequence is what has to match (the order of a sum, which operand a min returns on a tie or a
// NaN), so every comment names the sequence. Cited reference lines are cuda/common/glam.cuh unless stated.
#pragma once
#include "rs_defs.h"

namespace rs {

// ---- scalars ------------------------------------------------------------------------------------------------------------
RS_FN float fneg(float a) { return ffrom(fbits(a) ^ 0x80000000u); }      // sign-bit flip (glam.cuh:41)
RS_FN float fabs_(float a) { return ffrom(fbits(a) & 0x7FFFFFFFu); }      // glam.cuh:42
RS_FN bool fnan(float a) { return (fbits(a) & 0x7FFFFFFFu) > 0x7F800000u; }
RS_FN bool ffinite(float a) { return (fbits(a) & 0x7FFFFFFFu) < 0x7F800000u; }
RS_FN uint32_t mask_of(bool c) { return c ? 0xFFFFFFFFu : 0u; }
// _mm_min_ps / _mm_max_ps lane semantics: the second operand unless the comparison holds (glam.cuh:45-46).
RS_FN float mm_min(float a, float b) { return a < b ? a : b; }
RS_FN float mm_max(float a, float b) { return a > b ? a : b; }
// Rust f32::min / max (NaN ignored), clamp (two compares in order), signum (glam.cuh:49-56).
RS_FN float fmin_rs(float a, float b) { return fnan(a) ? b : fnan(b) ? a : (a < b ? a : b); }
RS_FN float fmax_rs(float a, float b) { return fnan(a) ? b : fnan(b) ? a : (a > b ? a : b); }
RS_FN float fclamp_rs(float x, float lo, float hi) {
  if (x < lo) x = lo;
  if (x > hi) x = hi;
  return x;
}
RS_FN float fsignum_rs(float a) { return fnan(a) ? a : ffrom((fbits(a) & 0x80000000u) | 0x3F800000u); }
// Rust `as` casts from f32: saturating, NaN to 0 (glam.cuh:59-76). The range tests come first, so the C conversion below
// never sees an out-of-range value.
RS_FN int32_t f2i_rs(float a) {
  if (fnan(a)) return 0;
  if (a >= 2147483648.0f) return 0x7FFFFFFF;
  if (a <= -2147483648.0f) return (int32_t)0x80000000u;
  return (int32_t)a;
}
RS_FN uint32_t f2u_rs(float a) {
  if (fnan(a) || a <= 0.0f) return 0u;
  if (a >= 4294967296.0f) return 0xFFFFFFFFu;
  return (uint32_t)a;
}
RS_FN float fselect(uint32_t m, float t, float f) { return ffrom((fbits(f) & ~m) | (fbits(t) & m)); }
// i64 -> f32, round to nearest even (hd.cuh:144-148): small values through i32, the rest through the magnitude.
RS_FN float i64_to_f(int64_t v) {
  if (v >= -2147483648LL && v <= 2147483647LL) return (float)(int32_t)v;
  float m = (float)(v < 0 ? 0ull - (uint64_t)v : (uint64_t)v);
  return v < 0 ? fneg(m) : m;
}

// ---- types (field layout is part of the state format: 4-byte words, no padding) ------------------------------------------
struct V3 { float x, y, z; };
struct V4 { float x, y, z, w; };
struct Quat { float x, y, z, w; };
struct M3 { V3 c0, c1, c2; };      // columns (glam Mat3A x_axis, y_axis, z_axis)
struct Aff { M3 m; V3 t; };        // glam Affine3A: matrix3, translation
struct B3 { uint32_t x, y, z; };   // lane masks, all-ones or zero

RS_FN V3 v3(float x, float y, float z) { return V3{x, y, z}; }
RS_FN V3 splat(float s) { return V3{s, s, s}; }
RS_FN float comp(const V3& a, int i) { return i == 0 ? a.x : i == 1 ? a.y : a.z; }
RS_FN void set_comp(V3& a, int i, float s) {
  if (i == 0) a.x = s; else if (i == 1) a.y = s; else a.z = s;
}

// ---- V3 lane-wise ops (vec3a.rs :1548-:2225 per glam.cuh:93-110) ----------------------------------------------------------
RS_FN V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
RS_FN V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
RS_FN V3 operator*(V3 a, V3 b) { return V3{a.x * b.x, a.y * b.y, a.z * b.z}; }
RS_FN V3 operator/(V3 a, V3 b) { return V3{a.x / b.x, a.y / b.y, a.z / b.z}; }
RS_FN V3 operator*(V3 a, float s) { return V3{a.x * s, a.y * s, a.z * s}; }
RS_FN V3 operator*(float s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
RS_FN V3 operator/(V3 a, float s) { return V3{a.x / s, a.y / s, a.z / s}; }
RS_FN V3 operator/(float s, V3 a) { return V3{s / a.x, s / a.y, s / a.z}; }
RS_FN V3 operator-(V3 a) { return V3{fneg(a.x), fneg(a.y), fneg(a.z)}; }
RS_FN V3& operator+=(V3& a, V3 b) { return a = a + b; }
RS_FN V3& operator-=(V3& a, V3 b) { return a = a - b; }
RS_FN V3& operator*=(V3& a, V3 b) { return a = a * b; }
RS_FN V3& operator*=(V3& a, float s) { return a = a * s; }
RS_FN V3& operator/=(V3& a, float s) { return a = a / s; }
RS_FN V3 add_s(V3 a, float s) { return a + splat(s); }
RS_FN bool is_zero(V3 a) { return a.x == 0.0f && a.y == 0.0f && a.z == 0.0f; }

// dot: (x x' + y y') + z z' (sse2.rs dot3_in_x; glam.cuh:113)
RS_FN float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
// cross: lane i = a_j b_k - b_j a_k (vec3a.rs:277; glam.cuh:115-118)
RS_FN V3 cross(V3 a, V3 b) { return V3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
RS_FN float length_sq(V3 a) { return dot(a, a); }
RS_FN float length(V3 a) { return __builtin_sqrtf(dot(a, a)); }
RS_FN float length_recip(V3 a) { return 1.0f / __builtin_sqrtf(dot(a, a)); }
// normalize divides by the length; normalize_or / try_normalize multiply by the reciprocal length (glam.cuh:122-135).
RS_FN V3 normalize(V3 a) { return a / splat(__builtin_sqrtf(dot(a, a))); }
RS_FN V3 normalize_or(V3 a, V3 fallback) {
  float r = length_recip(a);
  return (ffinite(r) && r > 0.0f) ? a * r : fallback;
}
RS_FN V3 normalize_or_zero(V3 a) { return normalize_or(a, splat(0.0f)); }
RS_FN bool try_normalize(V3 a, V3* out) {
  float r = length_recip(a);
  if (ffinite(r) && r > 0.0f) {
    *out = a * r;
    return true;
  }
  return false;
}
RS_FN V3 lerp(V3 a, V3 b, float s) { return a * (1.0f - s) + b * s; }  // vec3a.rs:1012 (glam.cuh:137)
RS_FN V3 vabs(V3 a) { return V3{fabs_(a.x), fabs_(a.y), fabs_(a.z)}; }
RS_FN V3 vmin(V3 a, V3 b) { return V3{mm_min(a.x, b.x), mm_min(a.y, b.y), mm_min(a.z, b.z)}; }
RS_FN V3 vmax(V3 a, V3 b) { return V3{mm_max(a.x, b.x), mm_max(a.y, b.y), mm_max(a.z, b.z)}; }
RS_FN V3 vclamp(V3 a, V3 lo, V3 hi) { return vmin(vmax(a, lo), hi); }  // max first (glam.cuh:141)
// the SSE reduction pairs (x, z) and (y, z) first (glam.cuh:143-144)
RS_FN float max_elem(V3 a) { return mm_max(mm_max(a.x, a.z), mm_max(a.y, a.z)); }
RS_FN B3 cmpeq(V3 a, V3 b) { return B3{mask_of(a.x == b.x), mask_of(a.y == b.y), mask_of(a.z == b.z)}; }
RS_FN B3 cmpge(V3 a, V3 b) { return B3{mask_of(a.x >= b.x), mask_of(a.y >= b.y), mask_of(a.z >= b.z)}; }
RS_FN B3 cmple(V3 a, V3 b) { return B3{mask_of(a.x <= b.x), mask_of(a.y <= b.y), mask_of(a.z <= b.z)}; }
RS_FN B3 cmplt(V3 a, V3 b) { return B3{mask_of(a.x < b.x), mask_of(a.y < b.y), mask_of(a.z < b.z)}; }
RS_FN bool all3(B3 m) { return (m.x & m.y & m.z) != 0u; }
RS_FN V3 vselect(B3 m, V3 t, V3 f) { return V3{fselect(m.x, t.x, f.x), fselect(m.y, t.y, f.y), fselect(m.z, t.z, f.z)}; }
RS_FN B3 finite_mask(V3 a) { return B3{mask_of(ffinite(a.x)), mask_of(ffinite(a.y)), mask_of(ffinite(a.z))}; }
// glam Vec3A::signum: (sign | 1.0) per lane, NaN lanes pass through (vec3a.rs:510)
RS_FN float lane_signum(float a) { return fnan(a) ? a : ffrom((fbits(a) & 0x80000000u) | 0x3F800000u); }
RS_FN V3 vsignum(V3 a) { return V3{lane_signum(a.x), lane_signum(a.y), lane_signum(a.z)}; }

// ---- V4 (the quaternion's vector ops and a few 4-lane reductions) --------------------------------------------------------
RS_FN V4 operator*(V4 a, V4 b) { return V4{a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
// dot4: (x x' + z z') + (y y' + w w') (sse2.rs dot4_into_m128; glam.cuh:179-181)
RS_FN float dot4(V4 a, V4 b) { return (a.x * b.x + a.z * b.z) + (a.y * b.y + a.w * b.w); }

// ---- M3 (column major; every product-sum is (x c0 + y c1) + z c2) ------------------------------------------------------------
RS_FN M3 m3_identity() { return M3{{1.0f, 0.0f, 0.0f}, {0.0f, 1.0f, 0.0f}, {0.0f, 0.0f, 1.0f}}; }
RS_FN V3 operator*(const M3& m, V3 v) { return (m.c0 * v.x + m.c1 * v.y) + m.c2 * v.z; }
RS_FN V3 mul_t(const M3& m, V3 v) { return V3{dot(m.c0, v), dot(m.c1, v), dot(m.c2, v)}; }  // transpose(m) * v
RS_FN M3 operator*(const M3& a, const M3& b) { return M3{a * b.c0, a * b.c1, a * b.c2}; }
RS_FN M3 transpose(const M3& m) {
  return M3{{m.c0.x, m.c1.x, m.c2.x}, {m.c0.y, m.c1.y, m.c2.y}, {m.c0.z, m.c1.z, m.c2.z}};
}
RS_FN M3 mabs(const M3& m) { return M3{vabs(m.c0), vabs(m.c1), vabs(m.c2)}; }
// mat3a.rs from_quat (glam.cuh:231-239)
RS_FN M3 m3_from_quat(Quat q) {
  float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
  float xx = q.x * x2, xy = q.x * y2, xz = q.x * z2;
  float yy = q.y * y2, yz = q.y * z2, zz = q.z * z2;
  float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
  return M3{{1.0f - (yy + zz), xy + wz, xz - wy}, {xy - wz, 1.0f - (xx + zz), yz + wx}, {xz + wy, yz - wx, 1.0f - (xx + yy)}};
}

// ---- Aff ------------------------------------------------------------------------------------------------------------------
RS_FN Aff aff_identity() { return Aff{m3_identity(), {0.0f, 0.0f, 0.0f}}; }
RS_FN V3 xform_point(const Aff& a, V3 p) { return a.m * p + a.t; }  // affine3a.rs:214
RS_FN Aff operator*(const Aff& a, const Aff& b) { return Aff{a.m * b.m, a.m * b.t + a.t}; }
// V3 AffineExt::transpose and inv_x_form (glam.cuh:260-264)
RS_FN Aff aff_transpose(const Aff& a) {
  M3 m = transpose(a.m);
  return Aff{m, m * (-a.t)};
}
RS_FN V3 inv_xform(const Aff& a, V3 v) { return mul_t(a.m, v - a.t); }

// ---- Quat -----------------------------------------------------------------------------------------------------------------
RS_FN V4 q4(Quat q) { return V4{q.x, q.y, q.z, q.w}; }
// quat.rs:549 normalize: divide by the 4-lane length
RS_FN Quat qnormalize(Quat q) {
  float len = __builtin_sqrtf(dot4(q4(q), q4(q)));
  return Quat{q.x / len, q.y / len, q.z / len, q.w / len};
}
// quat.rs:794 Quat * Quat (rtm mul_quat): two partial products per lane, then their sum (glam.cuh:250-261). The +-1
// multipliers of rtm are exact and appear here as negations.
RS_FN Quat operator*(Quat l, Quat r) {
  float ax = l.x * r.w + l.w * r.x;
  float ay = fneg(l.x * r.z) + l.w * r.y;
  float az = l.x * r.y + l.w * r.z;
  float aw = fneg(l.x * r.x) + l.w * r.w;
  float bx = fneg(l.z * r.y) + l.y * r.z;
  float by = l.z * r.x + l.y * r.w;
  float bz = l.z * r.w + fneg(l.y * r.x);
  float bw = fneg(l.z * r.z) + fneg(l.y * r.y);
  return Quat{ax + bx, ay + by, az + bz, aw + bw};
}
// quat.rs:873 mul_vec3a: ((v (w^2 - |b|^2)) + b 2(v.b)) + (b x v) 2w
RS_FN V3 qrotate(Quat q, V3 v) {
  V3 b = V3{q.x, q.y, q.z};
  float b2 = dot(b, b);
  V3 t1 = v * splat(q.w * q.w - b2);
  V3 t2 = b * splat(dot(v, b) * 2.0f);
  V3 t3 = cross(b, v) * splat(q.w * 2.0f);
  return (t1 + t2) + t3;
}
// quat.rs:220 from_rotation_axes (the DirectXMath branch tree), glam.cuh:283-307
RS_FN Quat quat_from_m3(const M3& m) {
  float m00 = m.c0.x, m01 = m.c0.y, m02 = m.c0.z;
  float m10 = m.c1.x, m11 = m.c1.y, m12 = m.c1.z;
  float m20 = m.c2.x, m21 = m.c2.y, m22 = m.c2.z;
  if (m22 <= 0.0f) {
    float dif10 = m11 - m00, omm22 = 1.0f - m22;
    if (dif10 <= 0.0f) {
      float f = omm22 - dif10, inv = 0.5f / __builtin_sqrtf(f);
      return Quat{f * inv, (m01 + m10) * inv, (m02 + m20) * inv, (m12 - m21) * inv};
    }
    float f = omm22 + dif10, inv = 0.5f / __builtin_sqrtf(f);
    return Quat{(m01 + m10) * inv, f * inv, (m12 + m21) * inv, (m20 - m02) * inv};
  }
  float sum10 = m11 + m00, opm22 = 1.0f + m22;
  if (sum10 <= 0.0f) {
    float f = opm22 - sum10, inv = 0.5f / __builtin_sqrtf(f);
    return Quat{(m02 + m20) * inv, (m12 + m21) * inv, f * inv, (m01 - m10) * inv};
  }
  float f = opm22 + sum10, inv = 0.5f / __builtin_sqrtf(f);
  return Quat{(m12 - m21) * inv, (m20 - m02) * inv, (m01 - m10) * inv, f * inv};
}

// ---- boxes ------------------------------------------------------------------------------------------------------------------
struct Aabb { V3 min, max; };
// shared/aabb.rs:43 intersects (touching counts); narrowphase.cuh:65-67
RS_FN bool overlaps(const Aabb& a, const Aabb& b) { return all3(cmple(a.min, b.max)) && all3(cmpge(a.max, b.min)); }

}  // namespace rs
