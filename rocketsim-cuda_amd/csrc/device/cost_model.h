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

hipError_t launch_features(hipStream_t st, const uint32_t* states, uint32_t n, float* feat);
hipError_t launch_mlp_infer(hipStream_t st, const float* params, const float* feat, uint32_t n, float* out);
hipError_t launch_mlp_train(hipStream_t st, const float* params, const float* feat, const float* target, const float* contact,
                            uint32_t b, float* acts, float* loss, float* grad);
hipError_t launch_adam(hipStream_t st, float* params, float* m, float* v, const float* grad, float lr, float bc1, float bc2);
hipError_t launch_gather_words(hipStream_t st, const uint32_t* src, const uint32_t* perm, uint32_t* dst, uint32_t n, uint32_t words,
                               bool scatter);
hipError_t launch_gather_states(hipStream_t st, const uint32_t* src, const uint32_t* perm, uint32_t* dst, uint32_t n);

}  // namespace dev
}  // namespace rs
