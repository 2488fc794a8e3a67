// The learned cost model: a 48 -> 64 -> 64 -> 2 MLP that predicts, from an arena's state, the log of its decision's cost
// (wall-clock ticks of a wave running that arena alone) and a contact logit; and the kernels that regroup arenas by it.
//
//   k_features     48 features per arena from the AoSoA-64 states (positions, velocities, wall and gap distances, counts)
//   k_mlp_infer    inference: bf16 weights and activations, fp32 accumulation, on v_mfma_f32_16x16x16_bf16 (one wave, 64
//                  arenas as four 16-row tiles; weights staged transposed in LDS, activations re-laid through LDS)
//   k_mlp_train    training forward + backward per arena in fp32 (the same bf16 rounding points as inference; the rounding
//                  is treated as the identity in the backward pass), writing activations and their gradients
//   k_mlp_reduce   parameter gradients: sums over the batch of outer products of activations and gradients
//   k_adam         Adam update of the fp32 master parameters
//   k_gather_words / k_scatter_words / k_gather_states   the regrouping permutations (states stay AoSoA-64)
//
// Parameter vector (fp32): W1[48][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2] (W[i][j]: input i -> output j).
#include "device/cost_model.h"

namespace rs {
namespace dev {
namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ inline uint16_t bf16_bits(float x) {  // round to nearest even (finite inputs)
  uint32_t u = __builtin_bit_cast(uint32_t, x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf16_round(float x) { return __builtin_bit_cast(float, (uint32_t)bf16_bits(x) << 16); }

__device__ inline uint32_t word_of(const uint32_t* buf, uint32_t a, size_t w) {
  return buf[(size_t)(a / AOSOA_LANES) * ARENA_WORDS * AOSOA_LANES + w * AOSOA_LANES + (a % AOSOA_LANES)];
}
__device__ inline float fword(const uint32_t* buf, uint32_t a, size_t byte_off) { return __builtin_bit_cast(float, word_of(buf, a, byte_off / 4)); }
__device__ inline V3 vword(const uint32_t* buf, uint32_t a, size_t byte_off) {
  return V3{fword(buf, a, byte_off), fword(buf, a, byte_off + 4), fword(buf, a, byte_off + 8)};
}
__device__ inline float clamp1(float x) { return x < -4.0f ? -4.0f : (x > 4.0f ? 4.0f : x); }

}  // namespace

// ---- features ------------------------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_features(const uint32_t* __restrict__ states, uint32_t n, float* __restrict__ feat) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  float f[COST_FEATURES];
  const size_t ball = offsetof(ArenaState, ball);
  V3 bp = vword(states, a, ball + offsetof(Ball, pos)), bv = vword(states, a, ball + offsetof(Ball, vel));
  V3 bw = vword(states, a, ball + offsetof(Ball, ang_vel));
  const V3 field = V3{4096.0f, 5120.0f, 2044.0f};
  f[0] = bp.x / field.x, f[1] = bp.y / field.y, f[2] = bp.z / field.z;
  f[3] = bv.x / 6000.0f, f[4] = bv.y / 6000.0f, f[5] = bv.z / 6000.0f;
  f[6] = bw.x / 6.0f, f[7] = bw.y / 6.0f, f[8] = bw.z / 6.0f;
  V3 cp[NUM_CARS];
  for (int c = 0; c < NUM_CARS; ++c) {
    const size_t car = offsetof(ArenaState, cars) + (size_t)c * sizeof(Car);
    V3 p = vword(states, a, car + offsetof(Car, pos)), v = vword(states, a, car + offsetof(Car, vel));
    V3 w = vword(states, a, car + offsetof(Car, ang_vel));
    float up_z = fword(states, a, car + offsetof(Car, rot) + offsetof(M3, c2) + 8);
    float boost = fword(states, a, car + offsetof(Car, boost));
    uint32_t ground = word_of(states, a, (car + offsetof(Car, is_on_ground)) / 4);
    uint32_t demo = word_of(states, a, (car + offsetof(Car, is_demoed)) / 4);
    float* g = f + 9 + 13 * c;
    g[0] = p.x / field.x, g[1] = p.y / field.y, g[2] = p.z / field.z;
    g[3] = v.x / 2300.0f, g[4] = v.y / 2300.0f, g[5] = v.z / 2300.0f;
    g[6] = w.x / 5.5f, g[7] = w.y / 5.5f, g[8] = w.z / 5.5f;
    g[9] = boost / 100.0f, g[10] = ground ? 1.0f : 0.0f, g[11] = up_z, g[12] = demo ? 1.0f : 0.0f;
    cp[c] = p;
  }
  f[35] = (field.x - fabsf(bp.x)) / field.x, f[36] = (field.y - fabsf(bp.y)) / field.y, f[37] = (field.z - bp.z) / field.z;
  f[38] = (field.x - fabsf(cp[0].x)) / field.x, f[39] = (field.y - fabsf(cp[0].y)) / field.y;
  f[40] = (field.x - fabsf(cp[1].x)) / field.x, f[41] = (field.y - fabsf(cp[1].y)) / field.y;
  f[42] = length(cp[0] - cp[1]) / 1000.0f, f[43] = length(cp[0] - bp) / 1000.0f, f[44] = length(cp[1] - bp) / 1000.0f;
  f[45] = (float)word_of(states, a, offsetof(ArenaState, n_pairs) / 4) / 16.0f;
  f[46] = (float)word_of(states, a, offsetof(ArenaState, n_manifolds) / 4) / 16.0f;
  f[47] = (float)word_of(states, a, offsetof(ArenaState, n_dyn_cells) / 4) / 64.0f;
  for (int i = 0; i < COST_FEATURES; ++i) feat[(size_t)a * COST_FEATURES + i] = clamp1(f[i]);
}

// ---- inference on MFMA ---------------------------------------------------------------------------------------------------------
// One wave, 64 arenas. v_mfma_f32_16x16x16_bf16 lane layout: A (16x16, rows x k): lane l holds row l % 16, k = 4 (l / 16)
// + 0..3; B (k x cols): col l % 16, k = 4 (l / 16) + 0..3; D (rows x cols): col l % 16, rows 4 (l / 16) + 0..3. Weights sit
// transposed in LDS (wt[out][in], bf16) so that both operands are four consecutive bf16 per lane.
constexpr int KP = 64;  // padded input width of every layer
__global__ void __launch_bounds__(64) k_mlp_infer(const float* __restrict__ params, const float* __restrict__ feat, uint32_t n,
                                                  float* __restrict__ out) {
  __shared__ uint16_t w1t[64 * KP], w2t[64 * KP], w3t[16 * KP];
  __shared__ uint16_t act[64 * KP];  // the 64 arenas' current activations, bf16, row-major [arena][k]
  __shared__ float bias[64 + 64 + 16];
  const int l = threadIdx.x;
  const float *W1 = params + COST_W1, *B1 = params + COST_B1, *W2 = params + COST_W2, *B2 = params + COST_B2,
              *W3 = params + COST_W3, *B3 = params + COST_B3;
  for (int e = l; e < 64 * KP; e += 64) {  // e = out * KP + in
    int o = e / KP, i = e % KP;
    w1t[e] = i < COST_FEATURES ? bf16_bits(W1[i * 64 + o]) : (uint16_t)0;
    w2t[e] = bf16_bits(W2[i * 64 + o]);
  }
  for (int e = l; e < 16 * KP; e += 64) {
    int o = e / KP, i = e % KP;
    w3t[e] = o < 2 ? bf16_bits(W3[i * 2 + o]) : (uint16_t)0;
  }
  bias[l] = B1[l];
  bias[64 + l] = B2[l];
  if (l < 16) bias[128 + l] = l < 2 ? B3[l] : 0.0f;
  const uint32_t base = blockIdx.x * 64;
  for (int r = 0; r < 64; ++r) {  // the block's 64 feature rows, padded with zeros (and zero rows past n)
    uint32_t a = base + r;
    float x = (a < n && l < COST_FEATURES) ? feat[(size_t)a * COST_FEATURES + l] : 0.0f;
    act[r * KP + l] = bf16_bits(x);
  }
  __syncthreads();
  const int lr = l % 16, lk = 4 * (l / 16);
  for (int layer = 0; layer < 3; ++layer) {
    const uint16_t* wt = layer == 0 ? w1t : layer == 1 ? w2t : w3t;
    const int nt = layer == 2 ? 1 : 4;  // 16-column output tiles
    v4f acc[4][4];
    for (int t = 0; t < 4; ++t)
      for (int j = 0; j < nt; ++j) {
        acc[t][j] = v4f{0.0f, 0.0f, 0.0f, 0.0f};
        for (int ks = 0; ks < KP / 16; ++ks) {
          const uint16_t* pa = act + (t * 16 + lr) * KP + ks * 16 + lk;
          const uint16_t* pb = wt + (j * 16 + lr) * KP + ks * 16 + lk;
          v4s av = v4s{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
          v4s bv = v4s{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
          acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc[t][j], 0, 0, 0);
        }
      }
    __syncthreads();  // every lane has read this layer's input
    for (int t = 0; t < 4; ++t)
      for (int j = 0; j < nt; ++j)
        for (int i = 0; i < 4; ++i) {
          int row = t * 16 + lk + i, col = j * 16 + lr;
          float v = acc[t][j][i] + bias[layer * 64 + col];
          if (layer < 2) {
            act[row * KP + col] = bf16_bits(v > 0.0f ? v : 0.0f);
          } else if (col < 2 && base + row < n) {
            out[(size_t)(base + row) * 2 + col] = v;
          }
        }
    __syncthreads();
  }
}

// ---- training ------------------------------------------------------------------------------------------------------------------
// Per arena: the forward pass in fp32 with bf16-rounded weights and activations (the inference kernel's rounding points,
// its sums in a different order), the losses (squared error on the normalised log cost, logistic loss on the contact
// flag, both averaged over the batch) and the backward pass down to the first layer's pre-activations.
__global__ void __launch_bounds__(128) k_mlp_train(const float* __restrict__ params, const float* __restrict__ feat,
                                                   const float* __restrict__ target, const float* __restrict__ contact,
                                                   uint32_t b, float* __restrict__ acts, float* __restrict__ loss) {
  uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= b) return;
  const float *W1 = params + COST_W1, *B1 = params + COST_B1, *W2 = params + COST_W2, *B2 = params + COST_B2,
              *W3 = params + COST_W3, *B3 = params + COST_B3;
  float* row = acts + (size_t)a * COST_ACT_WORDS;  // x[48] h1[64] h2[64] dy[2] dh1[64] dh2[64]
  float *x = row, *h1 = row + 48, *h2 = row + 112, *dy = row + 176, *dh1 = row + 178, *dh2 = row + 242;
  for (int i = 0; i < COST_FEATURES; ++i) x[i] = bf16_round(feat[(size_t)a * COST_FEATURES + i]);
  for (int j = 0; j < 64; ++j) {
    float s = 0.0f;
    for (int i = 0; i < COST_FEATURES; ++i) s += bf16_round(W1[i * 64 + j]) * x[i];
    s += B1[j];
    h1[j] = bf16_round(s > 0.0f ? s : 0.0f);
  }
  for (int j = 0; j < 64; ++j) {
    float s = 0.0f;
    for (int i = 0; i < 64; ++i) s += bf16_round(W2[i * 64 + j]) * h1[i];
    s += B2[j];
    h2[j] = bf16_round(s > 0.0f ? s : 0.0f);
  }
  float y[2];
  for (int k = 0; k < 2; ++k) {
    float s = 0.0f;
    for (int i = 0; i < 64; ++i) s += bf16_round(W3[i * 2 + k]) * h2[i];
    y[k] = s + B3[k];
  }
  const float inv_b = 1.0f / (float)b;
  float e = y[0] - target[a], c = contact[a];
  float sig = 1.0f / (1.0f + expf(-y[1]));
  dy[0] = 2.0f * e * inv_b;
  dy[1] = (sig - c) * inv_b;
  float sp = y[1] > 0.0f ? y[1] + log1pf(expf(-y[1])) : log1pf(expf(y[1]));  // softplus
  loss[a] = e * e + (sp - c * y[1]);
  for (int j = 0; j < 64; ++j) {
    float g = bf16_round(W3[j * 2 + 0]) * dy[0] + bf16_round(W3[j * 2 + 1]) * dy[1];
    dh2[j] = h2[j] > 0.0f ? g : 0.0f;
  }
  for (int i = 0; i < 64; ++i) {
    float g = 0.0f;
    for (int j = 0; j < 64; ++j) g += bf16_round(W2[i * 64 + j]) * dh2[j];
    dh1[i] = h1[i] > 0.0f ? g : 0.0f;
  }
}

// grad[p] = sum over the batch of (input activation) x (output gradient) for the weights, (output gradient) for the biases.
__global__ void __launch_bounds__(256) k_mlp_reduce(const float* __restrict__ acts, uint32_t b, float* __restrict__ grad) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= COST_PARAMS) return;
  int in_off = -1, out_off, i = 0, j;
  if (p < (uint32_t)COST_B1) { i = p / 64; j = p % 64; in_off = 0; out_off = 178; }
  else if (p < (uint32_t)COST_W2) { j = p - COST_B1; out_off = 178; }
  else if (p < (uint32_t)COST_B2) { uint32_t q = p - COST_W2; i = q / 64; j = q % 64; in_off = 48; out_off = 242; }
  else if (p < (uint32_t)COST_W3) { j = p - COST_B2; out_off = 242; }
  else if (p < (uint32_t)COST_B3) { uint32_t q = p - COST_W3; i = q / 2; j = q % 2; in_off = 112; out_off = 176; }
  else { j = p - COST_B3; out_off = 176; }
  float s = 0.0f;
  for (uint32_t a = 0; a < b; ++a) {
    const float* row = acts + (size_t)a * COST_ACT_WORDS;
    float g = row[out_off + j];
    s += in_off >= 0 ? row[in_off + i] * g : g;
  }
  grad[p] = s;
}

__global__ void __launch_bounds__(256) k_adam(float* __restrict__ params, float* __restrict__ m, float* __restrict__ v,
                                              const float* __restrict__ grad, float lr, float bc1, float bc2) {
  uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= COST_PARAMS) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float g = grad[p];
  float mp = b1 * m[p] + (1.0f - b1) * g, vp = b2 * v[p] + (1.0f - b2) * g * g;
  m[p] = mp;
  v[p] = vp;
  params[p] -= lr * (mp / bc1) / (sqrtf(vp / bc2) + eps);
}

// ---- regrouping ----------------------------------------------------------------------------------------------------------------
// dst item i = src item perm[i] (words per item contiguous) and its inverse.
__global__ void __launch_bounds__(256) k_gather_words(const uint32_t* __restrict__ src, const uint32_t* __restrict__ perm,
                                                      uint32_t* __restrict__ dst, uint32_t n, uint32_t words) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * words) return;
  size_t i = e / words, w = e % words;
  uint32_t j = perm[i];
  if (j < n) dst[e] = src[(size_t)j * words + w];  // a permutation never fails this; a bad index reads nothing
}
__global__ void __launch_bounds__(256) k_scatter_words(const uint32_t* __restrict__ src, const uint32_t* __restrict__ perm,
                                                       uint32_t* __restrict__ dst, uint32_t n, uint32_t words) {
  size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)n * words) return;
  size_t i = e / words, w = e % words;
  uint32_t j = perm[i];
  if (j < n) dst[(size_t)j * words + w] = src[e];
}
// AoSoA-64 states: arena i of dst (n arenas) = arena perm[i] of src (n_src arenas), the persistent words (the tail is zero at a tick boundary). The
// grid's y dimension walks the words, so all arenas of one word are copied together and the source lines they share
// are read from HBM once.
__global__ void __launch_bounds__(256) k_gather_states(const uint32_t* __restrict__ src, uint32_t n_src,
                                                       const uint32_t* __restrict__ perm, uint32_t* __restrict__ dst, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t w = blockIdx.y;
  if (i >= n || w >= (uint32_t)ARENA_WORDS) return;
  uint32_t a = perm[i];
  if (a < n_src) dst[aosoa_index(i, w)] = src[aosoa_index(a, w)];
}

// ---- the order by predicted cost ---------------------------------------------------------------------------------------------
// A stable counting sort on the device: each arena's predicted normalised log cost falls into one of ORDER_BUCKETS
// buckets (bucket 0 the most expensive); ranks are stable within a bucket (slot order), so the order is deterministic.
// k_order_count: per-block bucket histograms, stored bucket-major (counts[b * blocks + block]) so that one exclusive scan
// of the whole array (k_order_scan) gives every (bucket, block) pair its first rank. k_order_place: each arena's rank =
// that base + its rank among the block's earlier arenas of the same bucket (wave ballots, then the earlier waves' counts),
// mapped to a slot by the mode.
__device__ inline uint32_t order_bucket(float p) {
  float q = (ORDER_HI - p) * ORDER_SCALE;
  return q != q ? ORDER_BUCKETS - 1u : q <= 0.0f ? 0u : q >= (float)(ORDER_BUCKETS - 1) ? ORDER_BUCKETS - 1u : (uint32_t)q;
}

__device__ inline uint32_t order_slot(uint32_t s, uint32_t n, int mode) {
  if (mode != ORDER_BALANCED) return s;
  // rank s is dealt round-robin over the W waves; the last wave holds r lanes, so after r rounds it is skipped
  const uint32_t W = (n + AOSOA_LANES - 1) / AOSOA_LANES, r = n - AOSOA_LANES * (W - 1);
  if (s < r * W) return (s % W) * AOSOA_LANES + s / W;
  const uint32_t t = s - r * W;  // W > 1 here: with one wave, r * W = n > s
  return (t % (W - 1)) * AOSOA_LANES + r + t / (W - 1);
}

__global__ void __launch_bounds__(ORDER_BLOCK) k_order_count(const float* __restrict__ pred, uint32_t n, uint32_t* __restrict__ counts) {
  __shared__ uint32_t h[ORDER_BUCKETS];
  const uint32_t t = threadIdx.x, blocks = gridDim.x;
  for (uint32_t b = t; b < ORDER_BUCKETS; b += ORDER_BLOCK) h[b] = 0;
  __syncthreads();
  const uint32_t a = blockIdx.x * ORDER_BLOCK + t;
  if (a < n) atomicAdd(&h[order_bucket(pred[2 * (size_t)a])], 1u);
  __syncthreads();
  for (uint32_t b = t; b < ORDER_BUCKETS; b += ORDER_BLOCK) counts[(size_t)b * blocks + blockIdx.x] = h[b];
}

// Exclusive scan of m words in place, one work-group of ORDER_SCAN threads: each thread sums a contiguous chunk, the
// chunk sums are scanned in LDS, then each thread rewrites its chunk.
__global__ void __launch_bounds__(ORDER_SCAN) k_order_scan(uint32_t* __restrict__ v, uint32_t m) {
  __shared__ uint32_t s[ORDER_SCAN];
  const uint32_t t = threadIdx.x, chunk = (m + ORDER_SCAN - 1) / ORDER_SCAN;
  const uint32_t lo = t * chunk < m ? t * chunk : m, hi = lo + chunk < m ? lo + chunk : m;
  uint32_t sum = 0;
  for (uint32_t i = lo; i < hi; ++i) sum += v[i];
  s[t] = sum;
  __syncthreads();
  for (uint32_t d = 1; d < ORDER_SCAN; d <<= 1) {  // Hillis-Steele inclusive scan
    uint32_t x = t >= d ? s[t - d] : 0u;
    __syncthreads();
    s[t] += x;
    __syncthreads();
  }
  uint32_t run = s[t] - sum;  // exclusive prefix of this chunk
  for (uint32_t i = lo; i < hi; ++i) {
    uint32_t c = v[i];
    v[i] = run;
    run += c;
  }
}

__global__ void __launch_bounds__(ORDER_BLOCK) k_order_place(const float* __restrict__ pred, uint32_t n, const uint32_t* __restrict__ base,
                                                             int mode, uint32_t* __restrict__ perm) {
  constexpr uint32_t WAVES = ORDER_BLOCK / AOSOA_LANES;
  __shared__ uint32_t wc[WAVES][ORDER_BUCKETS];
  const uint32_t t = threadIdx.x, lane = t % AOSOA_LANES, wave = t / AOSOA_LANES;
  for (uint32_t e = t; e < WAVES * ORDER_BUCKETS; e += ORDER_BLOCK) wc[e / ORDER_BUCKETS][e % ORDER_BUCKETS] = 0;
  __syncthreads();
  const uint32_t a = blockIdx.x * ORDER_BLOCK + t;
  const bool valid = a < n;
  const uint32_t b = valid ? order_bucket(pred[2 * (size_t)a]) : 0u;
  uint64_t peers = __ballot(valid);  // every lane of the wave takes part in each ballot
  for (uint32_t bit = 0; (1u << bit) < ORDER_BUCKETS; ++bit) {
    const bool on = (b >> bit) & 1u;
    const uint64_t m = __ballot(on);
    peers &= on ? m : ~m;
  }
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t rank_in_wave = (uint32_t)__popcll(peers & below);
  if (valid && rank_in_wave == 0) wc[wave][b] = (uint32_t)__popcll(peers);  // the bucket's first lane records its count
  __syncthreads();
  if (!valid) return;
  uint32_t rank = rank_in_wave;
  for (uint32_t w = 0; w < wave; ++w) rank += wc[w][b];
  const uint32_t s = base[(size_t)b * gridDim.x + blockIdx.x] + rank;
  if (s < n) perm[order_slot(s, n, mode)] = a;  // s < n always (the counts sum to n); the check keeps a stray rank out
}

// The cost of an AoSoA block of 64 arenas for the block queue (k_run_queue): the largest predicted cost among its arenas
// (a wave lasts about as long as its slowest lane), written as a prediction row (cost, 0) so the block order is the
// arenas' counting sort applied to blocks. NaN predictions are skipped (a block of NaNs keeps NaN, which sorts last).
__global__ void __launch_bounds__(256) k_block_cost(const float* __restrict__ pred, uint32_t n, float* __restrict__ out) {
  const uint32_t b = blockIdx.x * 256 + threadIdx.x, nb = (n + AOSOA_LANES - 1) / AOSOA_LANES;
  if (b >= nb) return;
  float m = __builtin_nanf("");
  for (uint32_t a = b * AOSOA_LANES; a < n && a < (b + 1) * AOSOA_LANES; ++a) {
    const float c = pred[2 * (size_t)a];
    if (c == c && !(c <= m)) m = c;  // m NaN: the first non-NaN cost; afterwards the larger
  }
  out[2 * (size_t)b] = m;
  out[2 * (size_t)b + 1] = 0.0f;
}

// ---- launchers -----------------------------------------------------------------------------------------------------------------
static unsigned blocks_of(size_t n, unsigned b) { return (unsigned)((n + b - 1) / b > 0 ? (n + b - 1) / b : 1); }
uint64_t order_work_words(uint32_t n) { return (uint64_t)ORDER_BUCKETS * blocks_of(n, ORDER_BLOCK); }
// The fixed workspace of launch_cost_order for callers that cannot allocate: the bucket counts of up to ORDER_SCRATCH_MAX_N
// arenas (1 MiB). Each GPU holds its own copy for the life of the process; host code must not pass this name to a kernel
// (on the host it is a stand-in, not a device address) but the address order_scratch_address returns for the current
// device. The counts are overwritten whole by every order (k_order_count writes every (bucket, block) word it scans), so
// nothing needs clearing between calls; two orders in flight on one GPU at once would race on it.
static __device__ __align__(256) uint32_t scratch[ORDER_SCRATCH_WORDS];
hipError_t order_scratch_address(uint32_t** out) { return hipGetSymbolAddress((void**)out, HIP_SYMBOL(scratch)); }

hipError_t launch_cost_order(hipStream_t st, const float* pred, uint32_t n, int mode, uint32_t* work, uint32_t* perm) {
  const unsigned blocks = blocks_of(n, ORDER_BLOCK);
  k_order_count<<<blocks, ORDER_BLOCK, 0, st>>>(pred, n, work);
  k_order_scan<<<1, ORDER_SCAN, 0, st>>>(work, (uint32_t)order_work_words(n));
  k_order_place<<<blocks, ORDER_BLOCK, 0, st>>>(pred, n, work, mode, perm);
  return hipGetLastError();
}
uint64_t block_order_work_words(uint32_t n) {
  const uint32_t nb = (n + AOSOA_LANES - 1) / AOSOA_LANES;
  return 2ull * nb + order_work_words(nb);
}
hipError_t launch_block_order(hipStream_t st, const float* pred, uint32_t n, uint32_t* work, uint32_t* order) {
  const uint32_t nb = (n + AOSOA_LANES - 1) / AOSOA_LANES;
  float* block_pred = reinterpret_cast<float*>(work);
  k_block_cost<<<blocks_of(nb, 256), 256, 0, st>>>(pred, n, block_pred);
  return launch_cost_order(st, block_pred, nb, ORDER_SORTED, work + 2ull * nb, order);
}
hipError_t launch_features(hipStream_t st, const uint32_t* states, uint32_t n, float* feat) {
  k_features<<<blocks_of(n, 256), 256, 0, st>>>(states, n, feat);
  return hipGetLastError();
}
hipError_t launch_mlp_infer(hipStream_t st, const float* params, const float* feat, uint32_t n, float* out) {
  k_mlp_infer<<<blocks_of(n, 64), 64, 0, st>>>(params, feat, n, out);
  return hipGetLastError();
}
hipError_t launch_mlp_train(hipStream_t st, const float* params, const float* feat, const float* target, const float* contact,
                            uint32_t b, float* acts, float* loss, float* grad) {
  k_mlp_train<<<blocks_of(b, 128), 128, 0, st>>>(params, feat, target, contact, b, acts, loss);
  k_mlp_reduce<<<blocks_of(COST_PARAMS, 256), 256, 0, st>>>(acts, b, grad);
  return hipGetLastError();
}
hipError_t launch_adam(hipStream_t st, float* params, float* m, float* v, const float* grad, float lr, float bc1, float bc2) {
  k_adam<<<blocks_of(COST_PARAMS, 256), 256, 0, st>>>(params, m, v, grad, lr, bc1, bc2);
  return hipGetLastError();
}
hipError_t launch_gather_words(hipStream_t st, const uint32_t* src, const uint32_t* perm, uint32_t* dst, uint32_t n, uint32_t words,
                               bool scatter) {
  if (scatter) k_scatter_words<<<blocks_of((size_t)n * words, 256), 256, 0, st>>>(src, perm, dst, n, words);
  else k_gather_words<<<blocks_of((size_t)n * words, 256), 256, 0, st>>>(src, perm, dst, n, words);
  return hipGetLastError();
}
hipError_t launch_gather_states(hipStream_t st, const uint32_t* src, uint32_t n_src, const uint32_t* perm, uint32_t* dst,
                                uint32_t n) {
  k_gather_states<<<dim3(blocks_of(n, 256), ARENA_WORDS), 256, 0, st>>>(src, n_src, perm, dst, n);
  return hipGetLastError();
}

}  // namespace dev
}  // namespace rs
