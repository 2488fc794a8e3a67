// Launchers of the cost-model and regrouping kernels (cost_model.hip). All asynchronous on `st`.
#pragma once
#include "device/launch.h"

namespace rs {
namespace dev {

constexpr int COST_FEATURES = 48;
// offsets into the fp32 parameter vector: W1[48][64], b1[64], W2[64][64], b2[64], W3[64][2], b3[2]
constexpr int COST_W1 = 0, COST_B1 = 48 * 64, COST_W2 = COST_B1 + 64, COST_B2 = COST_W2 + 64 * 64, COST_W3 = COST_B2 + 64,
              COST_B3 = COST_W3 + 64 * 2, COST_PARAMS = COST_B3 + 2;
constexpr int COST_ACT_WORDS = 48 + 64 + 64 + 2 + 64 + 64;  // per training arena: x, h1, h2, dy, dh1, dh2

// The order by predicted cost: ORDER_BUCKETS buckets of the normalised log cost over [ORDER_HI - ORDER_BUCKETS /
// ORDER_SCALE, ORDER_HI), bucket 0 the most expensive (above the range), NaN in the last.
constexpr uint32_t ORDER_BUCKETS = 256, ORDER_BLOCK = 256, ORDER_SCAN = 1024;
constexpr float ORDER_HI = 3.0f, ORDER_SCALE = 64.0f;
constexpr int ORDER_SORTED = 0, ORDER_BALANCED = 1;
uint64_t order_work_words(uint32_t n);
// The allocation-free order's fixed workspace: order_work_words(ORDER_SCRATCH_MAX_N) words in a device array of this
// library, one per GPU; order_scratch_address gives the current device's copy.
constexpr uint32_t ORDER_SCRATCH_MAX_N = 262144;
constexpr uint32_t ORDER_SCRATCH_WORDS = ORDER_BUCKETS * ((ORDER_SCRATCH_MAX_N + ORDER_BLOCK - 1) / ORDER_BLOCK);
hipError_t order_scratch_address(uint32_t** out);
hipError_t launch_cost_order(hipStream_t st, const float* pred, uint32_t n, int mode, uint32_t* work, uint32_t* perm);

// The order of the ceil(n / 64) AoSoA blocks by their largest predicted arena cost, most expensive first (stable):
// work holds block_order_work_words(n) words.
uint64_t block_order_work_words(uint32_t n);
hipError_t launch_block_order(hipStream_t st, const float* pred, uint32_t n, uint32_t* work, uint32_t* order);
hipError_t launch_features(hipStream_t st, const uint32_t* states, uint32_t n, float* feat);
hipError_t launch_mlp_infer(hipStream_t st, const float* params, const float* feat, uint32_t n, float* out);
hipError_t launch_mlp_train(hipStream_t st, const float* params, const float* feat, const float* target, const float* contact,
                            uint32_t b, float* acts, float* loss, float* grad);
hipError_t launch_adam(hipStream_t st, float* params, float* m, float* v, const float* grad, float lr, float bc1, float bc2);
hipError_t launch_gather_words(hipStream_t st, const uint32_t* src, const uint32_t* perm, uint32_t* dst, uint32_t n, uint32_t words,
                               bool scatter);
hipError_t launch_gather_states(hipStream_t st, const uint32_t* src, uint32_t n_src, const uint32_t* perm, uint32_t* dst,
                                uint32_t n);

}  // namespace dev
}  // namespace rs
