// Build policy of the single-source stepper core. The same headers compile with hipcc for gfx950 (the stepper kernels) and
// with g++ for the host (the world builder, and the CPU build the tests check against).
//
// Float policy (what makes the two builds agree bit for bit): every float and double operation is one IEEE operation with
// round-to-nearest-even, written as a plain C++ operator. Both builds must compile with -ffp-contract=off (no a*b+c fusion),
// without fast-math, and with fp32 denormals kept (gfx950: -fno-gpu-flush-denormals-to-zero); HIP's float `/` and sqrt are
// its correctly rounded sequences (the default -fhip-fp32-correctly-rounded-divide-sqrt). The host build must have no FMA
// target (x86-64 without -mfma/-march=native); rs_defs.h refuses to compile otherwise.
#pragma once
#include <stdint.h>
#include <stddef.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define RS_HD __host__ __device__
#define RS_FN __host__ __device__ inline
#define RS_NOINL __host__ __device__ __attribute__((noinline)) inline
#define RS_INLINE __host__ __device__ __attribute__((always_inline)) inline
#define RS_TABLE static __device__ __constant__ const
#else
#define RS_HD
#define RS_FN inline
#define RS_NOINL inline
#define RS_INLINE inline
#define RS_TABLE static const
#endif

// Address-space qualifiers for the gfx950 compiler. A non-inlined function that takes its lane-private state and its global
// StaticData through pointers qualified with these compiles their accesses to scratch_* / global_* instructions; through
// generic pointers they would be flat instructions, which wait on both the vector-memory and the LDS/scalar counters and
// need 64-bit VGPR addresses. Empty on the host.
// RS_PRIVATE(p) states the same for a pointer parameter that always points into the lane's private memory. The StaticData
// is passed as constant memory (RS_CONST_AS): no kernel writes it, and the compiler may then read it with scalar loads
// wherever the address is the same in every lane (it cannot prove that of global memory inside a non-kernel function).
#if defined(__HIP_DEVICE_COMPILE__)
#define RS_PRIVATE_AS __attribute__((address_space(5)))
#define RS_GLOBAL_AS __attribute__((address_space(1)))
#define RS_CONST_AS __attribute__((address_space(4)))
#define RS_PRIVATE(p) __builtin_assume(__builtin_amdgcn_is_private((const void*)(p)))
#else
#define RS_PRIVATE_AS
#define RS_GLOBAL_AS
#define RS_CONST_AS
#define RS_PRIVATE(p) ((void)0)
#endif

// Wave-uniform values across a call. A non-inlined function receives its arguments in vector registers, so the compiler
// treats them as lane-varying: a world pointer that every lane shares would then be read with vector loads, and a phase or
// car index would steer divergent branches. RS_UNIFORM_INDEX(i) states that the active lanes hold the same index (true of
// the phase and car indices: every lane runs the same phases and cars in the same order); rs_uniform_ptr(p) does the same
// for a pointer, for callers whose lanes all step in one world, so reads at uniform offsets become scalar loads.
#if defined(__HIP_DEVICE_COMPILE__)
#define RS_UNIFORM_INDEX(i) ((i) = __builtin_amdgcn_readfirstlane(i))
template <class T>
__device__ inline T* rs_uniform_ptr(T* p) {
  const unsigned long long v = (unsigned long long)p;
  const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)v);
  const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(unsigned)(v >> 32));
  return (T*)(((unsigned long long)hi << 32) | lo);
}
#else
#define RS_UNIFORM_INDEX(i) ((void)0)
template <class T>
inline T* rs_uniform_ptr(T* p) { return p; }
#endif

#if defined(__clang__)
#pragma clang fp contract(off)
#define RS_UNROLL _Pragma("unroll")
#else
#define RS_UNROLL
#endif

#if !defined(__HIP_DEVICE_COMPILE__) && (defined(__FMA__) || defined(__FP_FAST_FMAF))
#error "the host build of the stepper core must not target FMA (it would allow fused multiply-adds)"
#endif

namespace rs {

RS_FN uint32_t fbits(float a) { return __builtin_bit_cast(uint32_t, a); }
RS_FN float ffrom(uint32_t u) { return __builtin_bit_cast(float, u); }
RS_FN uint64_t dbits(double a) { return __builtin_bit_cast(uint64_t, a); }
RS_FN double dfrom(uint64_t u) { return __builtin_bit_cast(double, u); }

}  // namespace rs
