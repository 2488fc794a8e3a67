// A user kernel on the device API (include/rsim_device.h): one decision of every arena with its controls, then an fp16
// observation and a reward per arena, without the states leaving the GPU. The counterpart of an RL loop's
// step-observe-reward kernel written against the reference's step_ticks.
//
// Observation (OBS_WORDS halves per arena): the ball's position / 4096 and velocity / 6000, then per car its position /
// 4096, velocity / 2300, boost / 100 and wheel contact (0 or 1). Reward: the ball's progress along +y this decision, / 100.
#include "rsim_device.h"

constexpr int OBS_WORDS = 6 + 8 * rs::NUM_CARS;

__global__ void __launch_bounds__(256) k_step_observe(uint32_t* states, const rs::StaticData* __restrict__ world,
                                                      const rs::Controls* __restrict__ controls, uint32_t n, uint32_t* result,
                                                      uint16_t* obs, float* reward) {
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  rs::ArenaState s;
  rsim_dev::load_arena(s, states, a);
  const float y0 = s.ball.pos.y;
  result[a] = rsim_dev::step_ticks(s, *world, controls + (size_t)a * rs::NUM_CARS, rs::TICKS_PER_DECISION);
  rsim_dev::store_arena(states, a, s);
  uint16_t* o = obs + (size_t)a * OBS_WORDS;
  const rs::V3 bp = s.ball.pos, bv = s.ball.vel;
  o[0] = rsim_dev::f2h(bp.x / 4096.0f), o[1] = rsim_dev::f2h(bp.y / 4096.0f), o[2] = rsim_dev::f2h(bp.z / 4096.0f);
  o[3] = rsim_dev::f2h(bv.x / 6000.0f), o[4] = rsim_dev::f2h(bv.y / 6000.0f), o[5] = rsim_dev::f2h(bv.z / 6000.0f);
  for (int c = 0; c < rs::NUM_CARS; ++c) {
    const rs::Car& car = s.cars[c];
    uint16_t* q = o + 6 + 8 * c;
    q[0] = rsim_dev::f2h(car.pos.x / 4096.0f), q[1] = rsim_dev::f2h(car.pos.y / 4096.0f), q[2] = rsim_dev::f2h(car.pos.z / 4096.0f);
    q[3] = rsim_dev::f2h(car.vel.x / 2300.0f), q[4] = rsim_dev::f2h(car.vel.y / 2300.0f), q[5] = rsim_dev::f2h(car.vel.z / 2300.0f);
    q[6] = rsim_dev::f2h(car.boost / 100.0f), q[7] = rsim_dev::f2h(car.is_on_ground ? 1.0f : 0.0f);
  }
  reward[a] = (s.ball.pos.y - y0) / 100.0f;
}

extern "C" {
int observe_words() { return OBS_WORDS; }
// states, world, controls, result, obs, reward: device pointers; stream: the rsim device's stream (rsim_stream).
int observe_step(void* stream, uint32_t* states, const void* world, const void* controls, uint32_t n, uint32_t* result,
                 uint16_t* obs, float* reward) {
  if (!n) return 0;
  const uint32_t block = 64;
  k_step_observe<<<(n + block - 1) / block, block, rsim_dev::lds_bytes(block), (hipStream_t)stream>>>(
      states, (const rs::StaticData*)world, (const rs::Controls*)controls, n, result, obs, reward);
  return (int)hipGetLastError();
}
// Word offsets of what the observation reads (for checking on the host): ball pos, ball vel, then per car pos, vel, boost,
// is_on_ground.
void observe_layout(uint32_t* out) {
  out[0] = offsetof(rs::ArenaState, ball) / 4 + offsetof(rs::Ball, pos) / 4;
  out[1] = offsetof(rs::ArenaState, ball) / 4 + offsetof(rs::Ball, vel) / 4;
  for (int c = 0; c < rs::NUM_CARS; ++c) {
    const size_t car = (offsetof(rs::ArenaState, cars) + c * sizeof(rs::Car)) / 4;
    out[2 + 4 * c] = (uint32_t)(car + offsetof(rs::Car, pos) / 4);
    out[3 + 4 * c] = (uint32_t)(car + offsetof(rs::Car, vel) / 4);
    out[4 + 4 * c] = (uint32_t)(car + offsetof(rs::Car, boost) / 4);
    out[5 + 4 * c] = (uint32_t)(car + offsetof(rs::Car, is_on_ground) / 4);
  }
}
// The fp16 conversion on the host, for checking (the same function the kernel uses).
void observe_f2h_host(const float* in, uint64_t n, uint16_t* out) {
  for (uint64_t i = 0; i < n; ++i) out[i] = rsim_dev::f2h(in[i]);
}
}
