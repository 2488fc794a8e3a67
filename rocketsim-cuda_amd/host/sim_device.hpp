// C++ host side of the decision stepper: the reference crate's device API (crates/rocketsim-cuda/src/device.rs:30-172,
// SimDevice / DeviceStatic / DeviceStates) over librsim.so's C ABI (include/rsim.h). RAII owns every device buffer, as
// cudarc's CudaSlice does in the reference; failures throw rsim::Error carrying the C ABI's status and message, where the
// reference returns DriverError or panics on misuse.
//
// States, StaticData and controls are the reference's word formats, held here as std::vector<uint32_t> (one ArenaState is
// rsim_arena_words() words, one StaticData rsim_static_words(), one arena's controls 16 words).
#pragma once

#include <stdint.h>

#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rsim.h"

namespace rsim {

constexpr uint32_t TICKS_PER_DECISION = RSIM_TICKS_PER_DECISION;
constexpr uint32_t SIM8_SCORED_MASK = RSIM_SIM8_SCORED_MASK;
constexpr uint32_t SIM8_UNIMPLEMENTED = RSIM_SIM8_UNIMPLEMENTED;
constexpr uint32_t SIM8_OVERFLOW = RSIM_SIM8_OVERFLOW;
constexpr uint32_t SIM8_BLOCK = RSIM_SIM8_BLOCK;
constexpr uint32_t CONTROL_WORDS = 16;  // per arena: two cars x (throttle, steer, pitch, yaw, roll, jump, boost, handbrake)

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& what) : std::runtime_error(what), code(c) {}
};
inline void check(int rc, const char* what) {
  if (rc != RSIM_OK) throw Error(rc, std::string(what) + ": " + rsim_last_error());
}

// StaticData on the device (device.rs:47-50).
class DeviceStatic {
 public:
  explicit DeviceStatic(rsim_world* w = nullptr) : w_(w) {}
  DeviceStatic(DeviceStatic&& o) noexcept : w_(std::exchange(o.w_, nullptr)) {}
  DeviceStatic& operator=(DeviceStatic&& o) noexcept {
    std::swap(w_, o.w_);
    return *this;
  }
  ~DeviceStatic() { rsim_world_free(w_); }
  const rsim_world* get() const { return w_; }

 private:
  rsim_world* w_;
};

// n arenas in one AoSoA-64 device buffer (device.rs:41-45).
class DeviceStates {
 public:
  explicit DeviceStates(rsim_states* s = nullptr) : s_(s) {}
  DeviceStates(DeviceStates&& o) noexcept : s_(std::exchange(o.s_, nullptr)) {}
  DeviceStates& operator=(DeviceStates&& o) noexcept {
    std::swap(s_, o.s_);
    return *this;
  }
  ~DeviceStates() { rsim_states_free(s_); }
  uint32_t n() const { return rsim_states_count(s_); }
  rsim_states* get() { return s_; }
  const rsim_states* get() const { return s_; }

 private:
  rsim_states* s_;
};

// A u32 device buffer (the reference's CudaSlice<u32>): controls, results, traces.
class DeviceWords {
 public:
  explicit DeviceWords(rsim_words* b = nullptr) : b_(b) {}
  DeviceWords(DeviceWords&& o) noexcept : b_(std::exchange(o.b_, nullptr)) {}
  DeviceWords& operator=(DeviceWords&& o) noexcept {
    std::swap(b_, o.b_);
    return *this;
  }
  ~DeviceWords() { rsim_words_free(b_); }
  uint64_t len() const { return rsim_words_count(b_); }
  rsim_words* get() { return b_; }
  const rsim_words* get() const { return b_; }

 private:
  rsim_words* b_;
};

class SimDevice {
 public:
  // SimDevice::new (device.rs:53-67)
  explicit SimDevice(int ordinal = 0) { check(rsim_open(ordinal, &d_), "rsim_open"); }
  SimDevice(const SimDevice&) = delete;
  SimDevice& operator=(const SimDevice&) = delete;
  ~SimDevice() { rsim_close(d_); }

  DeviceStatic upload_static(const std::vector<uint32_t>& sd) {
    if (sd.size() != rsim_static_words()) throw Error(RSIM_ERR_INVALID, "upload_static: not a StaticData");
    rsim_world* w = nullptr;
    check(rsim_upload_world(d_, sd.data(), &w), "upload_static");
    return DeviceStatic(w);
  }
  // A table of `worlds.size() / rsim_static_words()` StaticData for sim8_variants.
  DeviceStatic upload_worlds(const std::vector<uint32_t>& worlds) {
    const uint64_t sw = rsim_static_words();
    if (worlds.empty() || worlds.size() % sw) throw Error(RSIM_ERR_INVALID, "upload_worlds: not whole StaticData");
    rsim_world* w = nullptr;
    check(rsim_upload_worlds(d_, worlds.data(), (uint32_t)(worlds.size() / sw), &w), "upload_worlds");
    return DeviceStatic(w);
  }
  // `states` holds n ArenaStates back to back (n * rsim_arena_words() words); packed to AoSoA-64 on upload.
  DeviceStates upload_states(const std::vector<uint32_t>& states) {
    const uint64_t aw = rsim_arena_words();
    if (states.size() % aw) throw Error(RSIM_ERR_INVALID, "upload_states: not whole ArenaStates");
    rsim_states* s = nullptr;
    check(rsim_upload_states(d_, states.data(), (uint32_t)(states.size() / aw), &s), "upload_states");
    return DeviceStates(s);
  }
  DeviceStates alloc_states(uint32_t n) {
    rsim_states* s = nullptr;
    check(rsim_alloc_states(d_, n, &s), "alloc_states");
    return DeviceStates(s);
  }
  std::vector<uint32_t> download_words(const DeviceStates& s) {
    std::vector<uint32_t> out(rsim_aosoa_words(s.n()));
    check(rsim_download_words(d_, s.get(), out.data()), "download_words");
    return out;
  }
  std::vector<uint32_t> download_states(const DeviceStates& s) {
    std::vector<uint32_t> out((size_t)s.n() * rsim_arena_words());
    check(rsim_download_states(d_, s.get(), out.data()), "download_states");
    return out;
  }
  // controls: n arenas x 16 words (Controls[n][2]).
  DeviceWords upload_controls(const std::vector<uint32_t>& controls) {
    if (controls.size() % CONTROL_WORDS) throw Error(RSIM_ERR_INVALID, "upload_controls: not whole arenas");
    rsim_words* b = nullptr;
    check(rsim_upload_controls(d_, controls.data(), (uint32_t)(controls.size() / CONTROL_WORDS), &b), "upload_controls");
    return DeviceWords(b);
  }
  DeviceWords alloc_words(uint64_t n) {
    rsim_words* b = nullptr;
    check(rsim_alloc_words(d_, n, &b), "alloc_words");
    return DeviceWords(b);
  }
  std::vector<uint32_t> read_words(const DeviceWords& b) {
    std::vector<uint32_t> out(b.len());
    check(rsim_read_words(d_, b.get(), out.data(), out.size()), "read_words");
    return out;
  }
  // sim8 (device.rs:151-168): one decision of every arena, in place, asynchronous on the device's stream.
  void sim8(DeviceStates& states, const DeviceStatic& sd, const DeviceWords& controls, DeviceWords& result,
            uint32_t block = SIM8_BLOCK, DeviceWords* trace = nullptr) {
    check(rsim_sim8(d_, states.get(), sd.get(), controls.get(), result.get(), block, trace ? trace->get() : nullptr), "sim8");
  }
  // sim8 with arena a stepped in world variant[a] of a table (no counterpart in the reference).
  void sim8_variants(DeviceStates& states, const DeviceStatic& worlds, const DeviceWords* variant, const DeviceWords& controls,
                     DeviceWords& result, uint32_t block = SIM8_BLOCK, DeviceWords* trace = nullptr) {
    check(rsim_sim8_variants(d_, states.get(), worlds.get(), variant ? variant->get() : nullptr, controls.get(), result.get(),
                             block, trace ? trace->get() : nullptr),
          "sim8_variants");
  }
  // The parity kernels (device.rs:95-141), synchronous; one PhaseStatus per arena.
  void identity(const DeviceStates& in, DeviceStates& out) { check(rsim_identity(d_, in.get(), out.get()), "identity"); }
  void identity_tick(const DeviceStates& in, DeviceStates& out) {
    check(rsim_identity_tick(d_, in.get(), out.get()), "identity_tick");
  }
  std::vector<int32_t> phase(int k, DeviceStates& states, const DeviceStatic& sd) {
    std::vector<int32_t> st(states.n());
    check(rsim_phase(d_, k, states.get(), sd.get(), st.data()), "phase");
    return st;
  }
  std::vector<int32_t> step_tick(DeviceStates& states, const DeviceStatic& sd) {
    std::vector<int32_t> st(states.n());
    check(rsim_step_tick(d_, states.get(), sd.get(), st.data()), "step_tick");
    return st;
  }
  void synchronize() { check(rsim_sync(d_), "synchronize"); }
  rsim_device* raw() { return d_; }

 private:
  rsim_device* d_ = nullptr;
};

}  // namespace rsim
