// TEST INFRASTRUCTURE ONLY. A C ABI over the reference's correctly rounded functions (cuda/common/libm_cr.cuh, included in
// place from the reference tree by `make -C oracle ref`), for caller-chosen inputs: the reference's own shim
// (crates/rocketsim-cuda/csrc/libm.cpp) only evaluates its built-in input streams. Output: _ref/libref_libm.so.
#include <stdint.h>
#include <string.h>

#include "common/libm_cr.cuh"

static uint32_t bits_of(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}
static float float_of(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}

extern "C" {
// fn 0 sinf_cr, 1 cosf_cr, 2 tanf_cr over the bit patterns x[i]
void refm_unary(int fn, const uint32_t* x, uint64_t n, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    float v = float_of(x[i]);
    out[i] = bits_of(fn == 0 ? lfcr::sinf_cr(v) : fn == 1 ? lfcr::cosf_cr(v) : lfcr::tanf_cr(v));
  }
}
// op 0 atan2f_cr(a, b), 1 powf_cr(a, b)
void refm_binary(int op, const uint32_t* a, const uint32_t* b, uint64_t n, uint32_t* out) {
  for (uint64_t i = 0; i < n; ++i) {
    float x = float_of(a[i]), y = float_of(b[i]);
    out[i] = bits_of(op == 0 ? lfcr::atan2f_cr(x, y) : lfcr::powf_cr(x, y));
  }
}
// op 0 atan2_64_cr(a, b), 1 hypot64_cr(a, b)
void refm_f64(int op, const double* a, const double* b, uint64_t n, double* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = op == 0 ? lfcr::atan2_64_cr(a[i], b[i]) : lfcr::hypot64_cr(a[i], b[i]);
}
}
