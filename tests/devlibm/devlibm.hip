// TEST INFRASTRUCTURE ONLY (not linked into librsim.so): the stepper core's libm entry points (rs::fsin, fcos, ftan,
// fatan2, fpow, datan2, dhypot of core/rs_consts.h, the functions every tick calls) evaluated on the GPU over
// caller-chosen inputs, so the tests can check the device build of the correctly rounded functions against the CPU
// build and the reference's. Compiled with the product's flags (rsim/build.py: -ffp-contract=off, denormals kept,
// correctly rounded divide and sqrt, no packed fp32) into tests/devlibm/libdevlibm.so.
//
// Entry points (extern "C", device pointers, asynchronous on `stream`, return a hipError_t as int):
//   devlibm_unary_digests  fn 0 sin / 1 cos / 2 tan over every float bit pattern of chunks first..first+count-1, chunk c
//                          being the 2^24 patterns c * 2^24 ...; digest[c] = sum (mod 2^64) over the chunk of
//                          mix64(x << 32 | f(x)) with mix64 SplitMix64's finaliser, so the digest does not depend on the
//                          order the lanes add in
//   devlibm_unary          the same functions over a list of bit patterns
//   devlibm_binary         op 0 atan2f(a[i], b[i]), op 1 powf(a[i], b[i])
//   devlibm_f64            op 0 atan2_64(a[i], b[i]), op 1 hypot64(a[i], b[i])
#include <hip/hip_runtime.h>

#include "core/rs_consts.h"

namespace {

constexpr uint32_t CHUNK_BITS = 24;
constexpr uint32_t PER_THREAD = 16;  // consecutive patterns per lane
constexpr uint32_t BLOCK = 256;

__device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__device__ inline uint32_t unary(int fn, uint32_t xb) {
  const float x = rs::ffrom(xb);
  return rs::fbits(fn == 0 ? rs::fsin(x) : fn == 1 ? rs::fcos(x) : rs::ftan(x));
}

__global__ void __launch_bounds__(BLOCK) k_clear(unsigned long long* d, uint32_t n) {
  uint32_t i = blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) d[i] = 0;
}
// grid: count * 2^24 / (PER_THREAD * BLOCK) blocks; each block lies inside one chunk
__global__ void __launch_bounds__(BLOCK) k_unary_digests(int fn, uint32_t first, unsigned long long* digest) {
  const uint64_t t = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  const uint64_t x0 = ((uint64_t)first << CHUNK_BITS) + t * PER_THREAD;  // < 2^32: the grid covers chunks below 256
  uint64_t sum = 0;
  for (uint32_t j = 0; j < PER_THREAD; ++j) {
    const uint32_t x = (uint32_t)(x0 + j);
    sum += mix64(((uint64_t)x << 32) | unary(fn, x));
  }
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);  // wave64 sum
  if ((threadIdx.x & 63) == 0) atomicAdd(&digest[(x0 >> CHUNK_BITS) - first], (unsigned long long)sum);
}
__global__ void __launch_bounds__(BLOCK) k_unary(int fn, const uint32_t* x, uint64_t n, uint32_t* out) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) out[i] = unary(fn, x[i]);
}
__global__ void __launch_bounds__(BLOCK) k_binary(int op, const float* a, const float* b, uint64_t n, uint32_t* out) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) out[i] = rs::fbits(op == 0 ? rs::fatan2(a[i], b[i]) : rs::fpow(a[i], b[i]));
}
__global__ void __launch_bounds__(BLOCK) k_f64(int op, const double* a, const double* b, uint64_t n, double* out) {
  const uint64_t i = (uint64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (i < n) out[i] = op == 0 ? rs::datan2(a[i], b[i]) : rs::dhypot(a[i], b[i]);
}
unsigned blocks(uint64_t n) { return (unsigned)((n + BLOCK - 1) / BLOCK); }

}  // namespace

extern "C" {
int devlibm_unary_digests(void* stream, int fn, uint32_t first, uint32_t count, uint64_t* digest) {
  if (fn < 0 || fn > 2 || first >= 256 || count > 256 - first || !digest) return (int)hipErrorInvalidValue;
  hipStream_t st = (hipStream_t)stream;
  if (!count) return 0;
  k_clear<<<blocks(count), BLOCK, 0, st>>>(reinterpret_cast<unsigned long long*>(digest), count);
  const uint64_t threads = ((uint64_t)count << CHUNK_BITS) / PER_THREAD;
  k_unary_digests<<<(unsigned)(threads / BLOCK), BLOCK, 0, st>>>(fn, first, reinterpret_cast<unsigned long long*>(digest));
  return (int)hipGetLastError();
}
int devlibm_unary(void* stream, int fn, const uint32_t* x, uint64_t n, uint32_t* out) {
  if (fn < 0 || fn > 2 || n >= (1ull << 32)) return (int)hipErrorInvalidValue;
  if (!n) return 0;
  k_unary<<<blocks(n), BLOCK, 0, (hipStream_t)stream>>>(fn, x, n, out);
  return (int)hipGetLastError();
}
int devlibm_binary(void* stream, int op, const float* a, const float* b, uint64_t n, uint32_t* out) {
  if (op < 0 || op > 1 || n >= (1ull << 32)) return (int)hipErrorInvalidValue;
  if (!n) return 0;
  k_binary<<<blocks(n), BLOCK, 0, (hipStream_t)stream>>>(op, a, b, n, out);
  return (int)hipGetLastError();
}
int devlibm_f64(void* stream, int op, const double* a, const double* b, uint64_t n, double* out) {
  if (op < 0 || op > 1 || n >= (1ull << 32)) return (int)hipErrorInvalidValue;
  if (!n) return 0;
  k_f64<<<blocks(n), BLOCK, 0, (hipStream_t)stream>>>(op, a, b, n, out);
  return (int)hipGetLastError();
}
}
