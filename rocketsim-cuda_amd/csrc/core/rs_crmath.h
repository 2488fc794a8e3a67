// Correctly rounded binary32 sin, cos, tan, atan2 and pow, a correctly rounded binary64 hypot, and a double-double
// binary64 atan2: the transcendentals of the tick (the reference reaches them as lf::sin/cos/tan/atan2/powf/atan2_64/hypot64,
// cuda/common/libm.cuh:40-69, with the correctly rounded build as its default, libm.cuh:22-36).
//
// A correctly rounded result is unique, so any correct method gives the reference's bits. The method here is this project's:
//  * arguments of sin/cos/tan are reduced exactly: x * 2/pi is formed in integers from a 192-bit window of the bits of 2/pi
//    (the reduced argument is good to ~2^-128 quarter turns), then scaled by pi/2 in double-double;
//  * a fast path evaluates in binary64 with a relative error below 2^-49 and keeps its result only when no binary32 rounding
//    boundary lies within 2^-46 of it (rounding test on the 29 bits below binary32 precision);
//  * otherwise an accurate path evaluates in double-double (error ~2^-100) and rounds with dd_round_f32, which also handles
//    subnormal results and exact midpoints;
//  * powf detects exact midpoint results (x^y a 25-bit odd multiple of a power of two) in integers and rounds them to even.
// Double-double products use Dekker's split (no FMA), so host and gfx950 run the same exact operations.
// Special values follow C99 Annex F; every NaN result is the canonical quiet NaN 0x7FC00000 (binary64: 0x7FF8000000000000),
// except powf(x, 1) and 1^y, which return their operand as the reference does (libm_cr.cuh:728-737).
#pragma once
#include "rs_defs.h"
#include "rs_crmath_tables.h"

namespace rs {
namespace crm {

// ---- double-double ----------------------------------------------------------------------------------------------------------
struct DD { double hi, lo; };
RS_FN DD two_sum(double a, double b) {
  double s = a + b, bb = s - a;
  return DD{s, (a - (s - bb)) + (b - bb)};
}
RS_FN DD fast_two_sum(double a, double b) {  // |a| >= |b| or a == 0
  double s = a + b;
  return DD{s, b - (s - a)};
}
RS_FN void split(double a, double& hi, double& lo) {
  double c = 134217729.0 * a;  // 2^27 + 1
  hi = c - (c - a);
  lo = a - hi;
}
RS_FN DD two_prod(double a, double b) {
  double p = a * b, ah, al, bh, bl;
  split(a, ah, al);
  split(b, bh, bl);
  return DD{p, ((ah * bh - p) + ah * bl + al * bh) + al * bl};
}
RS_FN DD dd(double a) { return DD{a, 0.0}; }
RS_FN DD neg(DD a) { return DD{-a.hi, -a.lo}; }
RS_FN DD add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi), t = two_sum(a.lo, b.lo);
  s = fast_two_sum(s.hi, s.lo + t.hi);
  return fast_two_sum(s.hi, s.lo + t.lo);
}
RS_FN DD sub(DD a, DD b) { return add(a, neg(b)); }
RS_FN DD mul(DD a, DD b) {
  DD p = two_prod(a.hi, b.hi);
  return fast_two_sum(p.hi, p.lo + (a.hi * b.lo + a.lo * b.hi));
}
RS_FN DD mul_d(DD a, double b) {
  DD p = two_prod(a.hi, b);
  return fast_two_sum(p.hi, p.lo + a.lo * b);
}
RS_FN DD div(DD a, DD b) {
  double q1 = a.hi / b.hi;
  DD r = sub(a, mul_d(b, q1));
  double q2 = r.hi / b.hi;
  r = sub(r, mul_d(b, q2));
  double q3 = r.hi / b.hi;
  DD q = fast_two_sum(q1, q2);
  return add(q, dd(q3));
}
RS_FN DD tab(const double (&t)[2]) { return DD{t[0], t[1]}; }

RS_FN double dabs(double a) { return dfrom(dbits(a) & 0x7FFFFFFFFFFFFFFFull); }
RS_FN float qnan_f() { return ffrom(0x7FC00000u); }
RS_FN double qnan_d() { return dfrom(0x7FF8000000000000ull); }

// Binary32 round-to-nearest-even of hi + lo (a normalized double-double, |lo| <= ulp(hi)/2). (float)hi is already the
// answer unless hi is exactly a binary32 rounding boundary (a midpoint between two floats, subnormal ones included), in which
// case lo decides; lo == 0 there is an exact tie, left to the conversion's ties-to-even.
RS_FN float dd_round_f32(double hi, double lo) {
  uint32_t sign = dbits(hi) >> 63 ? 0x80000000u : 0u;
  double a = dabs(hi), b = sign ? -lo : lo;
  float f = (float)a;
  uint32_t fb = fbits(f);
  if (b == 0.0 || (double)f == a) return ffrom(fb | sign);
  if (fb == 0x7F800000u) {  // rounded to infinity: only the midpoint between FLT_MAX and 2^128 can go back
    if (a == 0x1.fffffep127 + 0x1p103 && b < 0.0) return ffrom(0x7F7FFFFFu | sign);
    return ffrom(fb | sign);
  }
  uint32_t gb = a > (double)f ? fb + 1u : fb - 1u;  // the neighbour on a's side
  double m = ((double)f + (double)ffrom(gb)) * 0.5;
  if (a != m) return ffrom(fb | sign);
  return ffrom(((b > 0.0) == (gb > fb) ? gb : fb) | sign);
}

// True when every value within 64 units of the last place of y rounds to (float)y: y in the binary32 normal range and the
// 29 bits below binary32 precision at least 65 units away from the midpoint pattern 0x10000000.
RS_FN bool f32_round_safe(double y) {
  uint64_t u = dbits(y) & 0x7FFFFFFFFFFFFFFFull;
  if (u < 0x3810000000000000ull || u >= 0x47EFFFFFE0000000ull) return false;  // below 2^-126 or near FLT_MAX
  int64_t t = (int64_t)(u & 0x1FFFFFFFull) - 0x10000000ll;
  return t > 64 || t < -64;
}

// ---- exact reduction by pi/2 -----------------------------------------------------------------------------------------------
RS_FN void mul_24x64(uint64_t m, uint64_t w, uint64_t* hi, uint64_t* lo) {  // m < 2^24
  uint64_t a = m * (w & 0xFFFFFFFFull), b = m * (w >> 32);
  uint64_t l = a + (b << 32);
  *hi = (b >> 32) + (l < a ? 1u : 0u);
  *lo = l;
}
RS_FN uint64_t table_bits(int first) {  // 64 bits of 2/pi starting at fraction bit `first` (1-based)
  int wi = (first - 1) >> 6, sh = (first - 1) & 63;
  uint64_t w = crt::two_over_pi[wi] << sh;
  if (sh) w |= crt::two_over_pi[wi + 1] >> (64 - sh);
  return w;
}
RS_FN uint64_t bits_at(const uint64_t p[4], int pos) {  // bits [pos, pos + 64) of a 256-bit little-endian word array
  int wi = pos >> 6, sh = pos & 63;
  uint64_t w = wi < 4 ? p[wi] >> sh : 0u;
  if (sh && wi + 1 < 4) w |= p[wi + 1] << (64 - sh);
  return w;
}

// |x| = q pi/2 + r (mod 2 pi) for finite |x| >= pi/4: q mod 4, and r in radians as a double-double, |r| <= pi/4 + tiny.
// Out of line on the device (as are the accurate paths below): the stepper inlines every call site of the transcendentals,
// and a copy of these long integer and double-double sequences at each one would bloat the hot phases' code for branches
// that are rarely or never taken there.
struct Reduced { DD r; int q; };
RS_NOINL Reduced reduce_pio2(float x) {
  uint32_t u = fbits(x) & 0x7FFFFFFFu;
  uint64_t m = (u & 0x7FFFFFu) | 0x800000u;
  int E = (int)(u >> 23) - 150;  // |x| = m 2^E
  int i0 = E - 1 > 1 ? E - 1 : 1;
  uint64_t w2 = table_bits(i0), w1 = table_bits(i0 + 64), w0 = table_bits(i0 + 128);
  uint64_t h0, l0, h1, l1, h2, l2;
  mul_24x64(m, w0, &h0, &l0);
  mul_24x64(m, w1, &h1, &l1);
  mul_24x64(m, w2, &h2, &l2);
  uint64_t p[4];
  p[0] = l0;
  p[1] = l1 + h0;
  uint64_t c1 = p[1] < l1 ? 1u : 0u;
  uint64_t t2 = l2 + h1;
  uint64_t c2 = t2 < l2 ? 1u : 0u;
  p[2] = t2 + c1;
  c2 += p[2] < t2 ? 1u : 0u;
  p[3] = h2 + c2;
  int pt = i0 + 191 - E;  // bit position of 2^0 in p
  int q = (int)(bits_at(p, pt) & 3u);
  uint64_t fh = bits_at(p, pt - 64), fl = bits_at(p, pt - 128);  // 128 fraction bits
  bool negative = (fh >> 63) != 0;
  if (negative) {  // fraction >= 1/2: next quadrant, r = fraction - 1
    q = (q + 1) & 3;
    fl = ~fl + 1u;
    fh = ~fh + (fl == 0 ? 1u : 0u);
  }
  // |fraction| = (fh 2^64 + fl) 2^-128 quarter turns, to double-double
  double d1, d2;
  if (fh != 0) {
    int lz = __builtin_clzll(fh);
    uint64_t nh = lz ? (fh << lz) | (fl >> (64 - lz)) : fh, nl = fl << lz;
    d1 = (double)(nh >> 11) * dfrom((uint64_t)(1023 - 53 - lz) << 52);
    d2 = ((double)(nh & 0x7FFu) * 0x1p64 + (double)nl) * dfrom((uint64_t)(1023 - 128 - lz) << 52);
  } else {
    int lz = fl ? __builtin_clzll(fl) : 0;
    uint64_t nl = fl << lz;
    d1 = (double)(nl >> 11) * dfrom((uint64_t)(1023 - 117 - lz) << 52);
    d2 = (double)(nl & 0x7FFu) * dfrom((uint64_t)(1023 - 128 - lz) << 52);
  }
  DD f = fast_two_sum(d1, d2);
  if (negative) f = neg(f);
  return Reduced{mul(f, DD{crt::pio2_hi, crt::pio2_lo}), q};
}

// ---- sin, cos, tan ----------------------------------------------------------------------------------------------------------
// Taylor series; the fast path in binary64 (|r| <= pi/4: truncation < 2^-60), the accurate one in double-double (< 2^-104).
RS_FN double sin_poly(double r) {
  double z = r * r, p = crt::inv_fact[17][0];
  p = -crt::inv_fact[15][0] + z * p;
  p = crt::inv_fact[13][0] + z * p;
  p = -crt::inv_fact[11][0] + z * p;
  p = crt::inv_fact[9][0] + z * p;
  p = -crt::inv_fact[7][0] + z * p;
  p = crt::inv_fact[5][0] + z * p;
  p = -crt::inv_fact[3][0] + z * p;
  return r + (r * z) * p;
}
RS_FN double cos_poly(double r) {
  double z = r * r, p = crt::inv_fact[18][0];
  p = -crt::inv_fact[16][0] + z * p;
  p = crt::inv_fact[14][0] + z * p;
  p = -crt::inv_fact[12][0] + z * p;
  p = crt::inv_fact[10][0] + z * p;
  p = -crt::inv_fact[8][0] + z * p;
  p = crt::inv_fact[6][0] + z * p;
  p = -crt::inv_fact[4][0] + z * p;
  p = crt::inv_fact[2][0] + z * p;  // 1/2! - z/4! + z^2/6! ...
  return 1.0 - z * p;
}
RS_FN DD sin_dd(DD r) {
  DD z = mul(r, r), p = tab(crt::inv_fact[29]);
  for (int k = 13; k >= 0; --k) p = add(tab(crt::inv_fact[2 * k + 1]), mul(neg(z), p));
  return mul(r, p);  // p = 1/1! - z/3! + z^2/5! ...
}
RS_FN DD cos_dd(DD r) {
  DD z = mul(r, r), p = tab(crt::inv_fact[28]);
  for (int k = 13; k >= 0; --k) p = add(tab(crt::inv_fact[2 * k]), mul(neg(z), p));
  return p;  // 1/0! - z/2! + z^2/4! ...
}

// The accurate path of trig_cr: the double-double value of sin, cos or tan at q pi/2 + r, sign-flipped when `flip`, rounded.
RS_NOINL float trig_accurate(DD r, int q, int which, bool flip) {
  DD v;
  if (which == 2) {
    DD s = sin_dd(r), c = cos_dd(r);
    v = (q & 1) ? neg(div(c, s)) : div(s, c);
  } else {
    int k = which == 0 ? q : q + 1;
    v = (k & 1) ? cos_dd(r) : sin_dd(r);
    if (k & 2) v = neg(v);
  }
  if (flip) v = neg(v);
  return dd_round_f32(v.hi, v.lo);
}

// which: 0 sin, 1 cos, 2 tan
RS_FN float trig_cr(float x, int which) {
  uint32_t ax = fbits(x) & 0x7FFFFFFFu;
  bool xneg = (fbits(x) >> 31) != 0;
  if (ax >= 0x7F800000u) return qnan_f();
  if (ax == 0) return which == 1 ? 1.0f : x;
  int q = 0;
  DD r;
  if (ax < 0x3F490FDBu) {  // |x| < pi/4 (the float below pi/4's float)
    r = DD{(double)ffrom(ax), 0.0};
  } else {
    Reduced red = reduce_pio2(x);
    r = red.r;
    q = red.q;
  }
  // fast path
  double y;
  if (which == 2) {
    double s = sin_poly(r.hi), c = cos_poly(r.hi);
    y = (q & 1) ? -(c / s) : s / c;
  } else {
    int k = which == 0 ? q : q + 1;  // cos(x) = sin(x + pi/2)
    double v = (k & 1) ? cos_poly(r.hi) : sin_poly(r.hi);
    y = (k & 2) ? -v : v;
  }
  bool flip = xneg && which != 1;  // sin and tan are odd, cos even
  if (f32_round_safe(y)) {
    float f = (float)y;
    return flip ? -f : f;
  }
  return trig_accurate(r, q, which, flip);
}
RS_FN float sinf_cr(float x) { return trig_cr(x, 0); }
RS_FN float cosf_cr(float x) { return trig_cr(x, 1); }
RS_FN float tanf_cr(float x) { return trig_cr(x, 2); }

// ---- atan2 ------------------------------------------------------------------------------------------------------------------
// atan(z) for z in [0, 1]: atan(j/16) + atan(t), t = (z - j/16) / (1 + z j/16), |t| <= 1/32.
RS_FN double atan_fast(double z) {
  int j = (int)(z * 16.0 + 0.5);
  double c = (double)j * 0.0625;
  double t = (z - c) / (1.0 + z * c);
  double t2 = t * t;
  double p = crt::inv_odd[6][0];  // atan t = t + t^3 (-1/3 + t^2/5 - ... + t^10/13)
  p = -crt::inv_odd[5][0] + t2 * p;
  p = crt::inv_odd[4][0] + t2 * p;
  p = -crt::inv_odd[3][0] + t2 * p;
  p = crt::inv_odd[2][0] + t2 * p;
  p = -crt::inv_odd[1][0] + t2 * p;
  return crt::atan_j16[j][0] + (t + (t * t2) * p);
}
RS_FN DD atan_dd(DD z) {
  int j = (int)(z.hi * 16.0 + 0.5);
  double c = (double)j * 0.0625;
  DD t = div(add(z, dd(-c)), add(mul_d(z, c), dd(1.0)));
  DD t2 = mul(t, t), p = tab(crt::inv_odd[12]);
  for (int k = 11; k >= 0; --k) p = add(tab(crt::inv_odd[k]), mul(neg(t2), p));  // 1 - t^2/3 + t^4/5 ...
  return add(tab(crt::atan_j16[j]), mul(t, p));
}
// The quotient num/den of two doubles as a double-double (num, den > 0, exponents far from the limits).
RS_FN DD quot_dd(double num, double den) {
  double q = num / den;
  DD pr = two_prod(q, den);
  double rem = (num - pr.hi) - pr.lo;
  return fast_two_sum(q, rem / den);
}

// The accurate path of atan2f_cr: atan(num/den) in double-double, reflected as the caller's fast path does, rounded.
RS_NOINL float atan2f_accurate(double num, double den, bool swap, bool xneg, bool yneg) {
  DD v = atan_dd(quot_dd(num, den));
  if (swap) v = sub(DD{crt::pio2_hi, crt::pio2_lo}, v);
  if (xneg) v = sub(DD{crt::pi_hi, crt::pi_lo}, v);
  if (yneg) v = neg(v);
  return dd_round_f32(v.hi, v.lo);
}

RS_FN float atan2f_cr(float y, float x) {
  uint32_t uy = fbits(y), ux = fbits(x), ay = uy & 0x7FFFFFFFu, ax = ux & 0x7FFFFFFFu;
  bool yneg = (uy >> 31) != 0, xneg = (ux >> 31) != 0;
  if (ay > 0x7F800000u || ax > 0x7F800000u) return qnan_f();
  float sy = yneg ? -1.0f : 1.0f;
  if (ay == 0x7F800000u || ax == 0x7F800000u) {
    if (ay == 0x7F800000u && ax == 0x7F800000u) return xneg ? sy * 0x1.2d97c8p+1f : sy * 0x1.921fb6p-1f;  // 3pi/4, pi/4
    if (ax == 0x7F800000u) return xneg ? sy * 0x1.921fb6p+1f : sy * 0.0f;
    return sy * 0x1.921fb6p+0f;
  }
  if (ay == 0) return xneg ? sy * 0x1.921fb6p+1f : sy * 0.0f;  // atan2(+-0, x): +-pi for x <= -0, +-0 for x >= +0
  if (ax == 0) return sy * 0x1.921fb6p+0f;
  double a = (double)ffrom(ay), b = (double)ffrom(ax);
  bool swap = a > b;
  double num = swap ? b : a, den = swap ? a : b;
  double at = atan_fast(num / den);
  double res = swap ? crt::pio2_hi - at : at;
  if (xneg) res = crt::pi_hi - res;
  if (f32_round_safe(res)) {
    float f = (float)res;
    return yneg ? -f : f;
  }
  return atan2f_accurate(num, den, swap, xneg, yneg);
}

// atan2 of binary64 arguments in double-double, one final rounding to binary64 (the reference's atan2_64 is likewise a
// double-double method, not correctly rounded: libm_cr.cuh:941-944). Used by the auto-flip roll sign, on binary32 values.
RS_NOINL double atan2_64(double y, double x) {
  uint64_t uy = dbits(y), ux = dbits(x), ay = uy & 0x7FFFFFFFFFFFFFFFull, ax = ux & 0x7FFFFFFFFFFFFFFFull;
  const uint64_t INF = 0x7FF0000000000000ull;
  bool yneg = (uy >> 63) != 0, xneg = (ux >> 63) != 0;
  double sy = yneg ? -1.0 : 1.0;
  if (ay > INF || ax > INF) return qnan_d();
  if (ay == INF || ax == INF) {
    if (ay == INF && ax == INF) return xneg ? sy * 0x1.2d97c7f3321d2p+1 : sy * 0x1.921fb54442d18p-1;
    if (ax == INF) return xneg ? sy * crt::pi_hi : sy * 0.0;
    return sy * crt::pio2_hi;
  }
  if (ay == 0) {
    if (ax == 0 || !xneg) return xneg ? sy * crt::pi_hi : sy * 0.0;
    return sy * crt::pi_hi;
  }
  if (ax == 0) return sy * crt::pio2_hi;
  double a = dabs(y), b = dabs(x);
  bool swap = a > b;
  double num = swap ? b : a, den = swap ? a : b;
  if (!swap && !xneg && num / den < 0x1p-300) return y / x;  // atan(z) = z to far below an ulp
  int e = (int)(dbits(den) >> 52) - 1023;  // bring the pair to the unit range (exact power-of-two scaling)
  if (e > 400 || e < -400) {
    // 2^-e as two factors, each a normal binary64 for every e of a finite den (2^-1023 alone is not); den lands in [1, 2)
    // exactly, num too unless it falls far below an ulp of the result (then it is replaced just below)
    const int e1 = e / 2, e2 = e - e1;
    const double s1 = dfrom((uint64_t)(1023 - e1) << 52), s2 = dfrom((uint64_t)(1023 - e2) << 52);
    num = num * s1 * s2;
    den = den * s1 * s2;
    if (num < 0x1p-600) num = 0x1p-600 * (num > 0.0 ? 1.0 : 0.0);  // far below an ulp of any result here
  }
  DD v = atan_dd(quot_dd(num, den));
  if (swap) v = sub(DD{crt::pio2_hi, crt::pio2_lo}, v);
  if (xneg) v = sub(DD{crt::pi_hi, crt::pi_lo}, v);
  double r = v.hi + v.lo;
  return yneg ? -r : r;
}

// ---- hypot (binary64, correctly rounded) ------------------------------------------------------------------------------------
// (R + f) 2^F with 0 <= f < 1 (f > 0 iff sticky) to the nearest binary64, ties to even, subnormal results included. The caller
// keeps at least one bit of R below the result's last place (R has 55 bits in hypot64).
RS_FN double round_to_f64(uint64_t R, int F, bool sticky) {
  int nb = 64 - __builtin_clzll(R);
  int lsb = F + nb - 53;  // weight of the result's last bit
  if (lsb < -1074) lsb = -1074;
  int drop = lsb - F;  // 1 <= drop <= 63 here
  uint64_t rest = R & ((1ull << drop) - 1u), keep = R >> drop, half = 1ull << (drop - 1);
  if (rest > half || (rest == half && (sticky || (keep & 1u)))) keep += 1u;
  if (keep >> 53) {
    keep >>= 1;
    lsb += 1;
  }
  if (keep < (1ull << 52)) return dfrom(keep);  // subnormal: keep 2^-1074
  int be = lsb + 52 + 1023;
  if (be >= 2047) return dfrom(0x7FF0000000000000ull);
  return dfrom(((uint64_t)be << 52) | (keep & 0xFFFFFFFFFFFFFull));
}
RS_FN void mul64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
  uint64_t a0 = a & 0xFFFFFFFFull, a1 = a >> 32, b0 = b & 0xFFFFFFFFull, b1 = b >> 32;
  uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
  uint64_t mid = (p00 >> 32) + (p01 & 0xFFFFFFFFull) + (p10 & 0xFFFFFFFFull);
  *lo = (p00 & 0xFFFFFFFFull) | (mid << 32);
  *hi = p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32);
}
RS_FN void decompose_f64(uint64_t u, uint64_t* m, int* e) {  // |v| = m 2^e, m in [2^52, 2^53), v finite nonzero
  int be = (int)(u >> 52);
  uint64_t f = u & 0xFFFFFFFFFFFFFull;
  if (be == 0) {
    int lz = __builtin_clzll(f) - 11;
    *m = f << lz;
    *e = -1074 - lz;
  } else {
    *m = f | (1ull << 52);
    *e = be - 1075;
  }
}

// sqrt(x^2 + y^2), correctly rounded: the sum of squares exactly in integers, its square root with remainder.
RS_NOINL double hypot64(double x, double y) {
  uint64_t ux = dbits(x) & 0x7FFFFFFFFFFFFFFFull, uy = dbits(y) & 0x7FFFFFFFFFFFFFFFull;
  const uint64_t INF = 0x7FF0000000000000ull;
  if (ux == INF || uy == INF) return dfrom(INF);
  if (ux > INF || uy > INF) return qnan_d();
  uint64_t ua = ux > uy ? ux : uy, ub = ux > uy ? uy : ux;
  if (ub == 0) return dfrom(ua);
  uint64_t ma, mb;
  int ea, eb;
  decompose_f64(ua, &ma, &ea);
  decompose_f64(ub, &mb, &eb);
  int d = ea - eb;
  if (d > 60) return dfrom(ua);  // b^2 / 2a is far below half an ulp of a
  uint64_t S[4] = {0, 0, 0, 0}, ah, al, bh, bl;  // S = ma^2 2^(2d) + mb^2, little-endian words
  mul64(ma, ma, &ah, &al);
  mul64(mb, mb, &bh, &bl);
  {
    int sh = 2 * d, wsh = sh >> 6, bsh = sh & 63;
    uint64_t w[3] = {al << bsh, bsh ? (ah << bsh) | (al >> (64 - bsh)) : ah, bsh ? ah >> (64 - bsh) : 0u};
    for (int i = 0; i < 3; ++i)
      if (i + wsh < 4) S[i + wsh] = w[i];
  }
  {
    uint64_t s0 = S[0] + bl, c0 = s0 < bl ? 1u : 0u;
    uint64_t s1 = S[1] + bh, c1 = s1 < bh ? 1u : 0u;
    uint64_t s1c = s1 + c0;
    c1 += s1c < s1 ? 1u : 0u;
    S[0] = s0;
    S[1] = s1c;
    for (int i = 2; i < 4 && c1; ++i) {
      S[i] += c1;
      c1 = S[i] == 0 ? 1u : 0u;
    }
  }
  int top = 255;  // S >= 2^104
  while (!((S[top >> 6] >> (top & 63)) & 1u)) --top;
  int t2 = top + 1 - 110;  // V = S 2^-t2 in [2^108, 2^110), t2 even
  if (t2 & 1) t2 += 1;
  uint64_t vh, vl;
  bool sticky = false;
  if (t2 >= 0) {
    vl = bits_at(S, t2);
    vh = bits_at(S, t2 + 64);
    for (int i = 0; i < t2 && !sticky; ++i) sticky = ((S[i >> 6] >> (i & 63)) & 1u) != 0;
  } else {
    int l = -t2;
    vh = (S[1] << l) | (S[0] >> (64 - l));
    vl = S[0] << l;
  }
  uint64_t R = (uint64_t)__builtin_sqrt((double)vh * 0x1p64 + (double)vl);  // within a few units of floor(sqrt V)
  for (int it = 0; it < 16; ++it) {
    uint64_t ph, pl, qh, ql;
    mul64(R, R, &ph, &pl);
    if (ph > vh || (ph == vh && pl > vl)) { --R; continue; }
    mul64(R + 1, R + 1, &qh, &ql);
    if (qh < vh || (qh == vh && ql <= vl)) { ++R; continue; }
    break;
  }
  uint64_t ph, pl;
  mul64(R, R, &ph, &pl);
  sticky = sticky || ph != vh || pl != vl;
  return round_to_f64(R, t2 / 2 + eb, sticky);  // hypot = sqrt(V 2^t2) 2^eb
}

// ---- pow (binary32) ---------------------------------------------------------------------------------------------------------
RS_FN bool f_isint(float y) {  // y a finite integer (zero included)
  uint32_t w = fbits(y) & 0x7FFFFFFFu;
  int e = (int)(w >> 23) - 127;
  if (w == 0) return true;
  if (e < 0) return false;
  if (e >= 23) return true;
  return (w & ((1u << (23 - e)) - 1u)) == 0;
}
RS_FN bool f_isodd(float y) {  // y an odd integer
  uint32_t w = fbits(y) & 0x7FFFFFFFu;
  int e = (int)(w >> 23) - 127;
  if (e < 0 || e > 23 || !f_isint(y)) return false;
  return ((((w & 0x7FFFFFu) | 0x800000u) >> (23 - e)) & 1u) != 0;
}
RS_FN bool f_is_signaling(float x) { return ((fbits(x) ^ 0x00400000u) & 0x7FFFFFFFu) > 0x7FC00000u; }

// log2(x) for a positive normal double x, double-double: x = m 2^e, m in [sqrt(1/2), sqrt(2)), ln m = 2 atanh((m-1)/(m+1)).
RS_FN DD log2_dd(double x) {
  uint64_t u = dbits(x);
  int e = (int)(u >> 52) - 1023;
  double m = dfrom((u & 0xFFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > 1.4142135623730951) {
    m *= 0.5;
    e += 1;
  }
  DD s = div(dd(m - 1.0), dd(m + 1.0));
  DD s2 = mul(s, s), p = tab(crt::inv_odd[24]);
  for (int k = 23; k >= 0; --k) p = add(tab(crt::inv_odd[k]), mul(s2, p));
  DD ln = mul_d(mul(s, p), 2.0);
  return add(dd((double)e), mul(ln, DD{crt::inv_ln2_hi, crt::inv_ln2_lo}));
}
// 2^t for |t| <= 400, double-double.
RS_FN DD exp2_dd(DD t) {
  double n = t.hi >= 0.0 ? (double)(int64_t)(t.hi + 0.5) : -(double)(int64_t)(0.5 - t.hi);
  DD f = add(t, dd(-n));  // |f| <= 1/2 + tiny
  DD u = mul(f, DD{crt::ln2_hi, crt::ln2_lo});
  DD p = tab(crt::inv_fact[24]);
  for (int k = 23; k >= 0; --k) p = add(tab(crt::inv_fact[k]), mul(u, p));
  int ni = (int)n;  // scale by 2^n in two exact steps (n can reach +-400)
  int n1 = ni / 2, n2 = ni - n1;
  double s1 = dfrom((uint64_t)(1023 + n1) << 52), s2 = dfrom((uint64_t)(1023 + n2) << 52);
  return DD{p.hi * s1 * s2, p.lo * s1 * s2};
}
// True when |x|^y is exactly M, a binary32 rounding midpoint (an odd multiple of 2^(e-24) with 25 significant bits, or
// (2j+1) 2^-150 below the normal range). x > 0 finite, y finite nonzero, M > 0.
RS_FN bool pow_is_exact_midpoint(float x, float y, double M) {
  uint32_t xb = fbits(x);
  uint32_t xm = xb & 0x7FFFFFu;
  int xe = (int)(xb >> 23);
  if (xe == 0) xe = 1; else xm |= 0x800000u;
  xe -= 150;  // x = xm 2^xe
  if (xm == 0) return false;
  while (!(xm & 1u)) { xm >>= 1; ++xe; }
  // M = Mo 2^Me, Mo odd
  uint64_t Mb = dbits(M);
  uint64_t Mm = (Mb & 0xFFFFFFFFFFFFFull) | (1ull << 52);
  int Me = (int)(Mb >> 52) - 1075;
  while (!(Mm & 1u)) { Mm >>= 1; ++Me; }
  // y = a 2^-k, a odd (k = 0 for an integer, then a may be even)
  uint32_t yb = fbits(y) & 0x7FFFFFFFu;
  int ye = (int)(yb >> 23);
  uint32_t ym = (yb & 0x7FFFFFu) | 0x800000u;
  ye -= 150;  // |y| = ym 2^ye
  bool yneg = (fbits(y) >> 31) != 0;
  while (!(ym & 1u) && ye < 0) { ym >>= 1; ++ye; }
  int k = ye < 0 ? -ye : 0;
  if (ye > 8) return false;  // |y| >= 2^32: x^y cannot be a 25-bit odd multiple (xm > 1), nor a midpoint power of two
  int64_t a = (int64_t)ym << (ye > 0 ? ye : 0);
  if (yneg) a = -a;
  if (xm == 1) {  // x a power of two: x^y = 2^(xe y); the only power-of-two midpoint is 2^-150
    if (Mm != 1) return false;
    // xe * a / 2^k == Me exactly
    int64_t num = (int64_t)xe * a;
    return k < 40 && (num % ((int64_t)1 << k)) == 0 && num / ((int64_t)1 << k) == Me;
  }
  if (yneg) return false;  // xm > 1 odd: x^y with y < 0 is not dyadic
  uint64_t w = xm;  // w = xm^(1/2^k), exactly, or no midpoint
  for (int i = 0; i < k; ++i) {
    uint64_t r = (uint64_t)__builtin_sqrt((double)w);
    while (r * r > w) --r;
    while ((r + 1) * (r + 1) <= w) ++r;
    if (r * r != w) return false;
    w = r;
  }
  uint64_t p = 1;  // w^a, bounded by 2^26
  for (int64_t i = 0; i < a; ++i) {
    p *= w;
    if (p > (1ull << 26)) return false;
  }
  if (p != Mm) return false;
  int64_t num = (int64_t)xe * a;  // exponent: xe a / 2^k must equal Me
  return (num % ((int64_t)1 << k)) == 0 && num / ((int64_t)1 << k) == Me;
}

RS_NOINL float powf_cr(float x, float y) {
  uint32_t ux = fbits(x), uy = fbits(y), ax = ux & 0x7FFFFFFFu, ay = uy & 0x7FFFFFFFu;
  bool xneg = (ux >> 31) != 0, yneg = (uy >> 31) != 0;
  if (ax == 0x3F800000u) {  // |x| = 1
    if (xneg) {
      if (ay > 0x7F800000u) return qnan_f();
      if (f_isint(y)) return f_isodd(y) ? x : -x;
      return qnan_f();
    }
    return f_is_signaling(y) ? qnan_f() : x;
  }
  if (ay == 0) return f_is_signaling(x) ? qnan_f() : 1.0f;
  if (uy == 0x3F800000u) return f_is_signaling(x) ? qnan_f() : x;
  if (ay >= 0x7F800000u) {  // y = +-inf or NaN
    if (ax > 0x7F800000u) return qnan_f();
    if (ay == 0x7F800000u) return ((ax < 0x3F800000u) != yneg) ? 0.0f : ffrom(0x7F800000u);
    return qnan_f();
  }
  if (ax > 0x7F800000u) return qnan_f();
  if (ax == 0x7F800000u) {  // x = +-inf
    float xi = f_isodd(y) ? x : ffrom(ax);
    return yneg ? 1.0f / xi : xi;
  }
  if (xneg && ax != 0 && !f_isint(y)) return qnan_f();
  if (ax == 0) {
    if (yneg) return f_isodd(y) ? 1.0f / x : ffrom(0x7F800000u);
    return f_isodd(y) ? x : 0.0f;
  }
  bool rneg = xneg && f_isodd(y);
  double xa = (double)ffrom(ax);
  DD t = mul_d(log2_dd(xa), (double)y);
  if (t.hi > 200.0) return rneg ? ffrom(0xFF800000u) : ffrom(0x7F800000u);
  if (t.hi < -200.0) return rneg ? -0.0f : 0.0f;
  DD v = exp2_dd(t);
  // An exact midpoint result must round to even; the double-double value sits within ~2^-95 of it on either side.
  float f = (float)v.hi;
  uint32_t fb = fbits(f);
  if (fb != 0x7F800000u) {
    uint32_t gb = v.hi > (double)f ? fb + 1u : fb - 1u;
    if ((double)f != v.hi && gb != 0x7F800000u) {
      double m = ((double)f + (double)ffrom(gb)) * 0.5;
      if (dabs(v.hi - m) <= m * 0x1p-80 && pow_is_exact_midpoint(ffrom(ax), y, m)) v = DD{m, 0.0};
    }
  }
  if (rneg) v = neg(v);
  return dd_round_f32(v.hi, v.lo);
}

}  // namespace crm
}  // namespace rs
