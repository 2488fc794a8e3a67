// C++ host side of the V3 wire format: crates/rocketsim-cuda/src/dump.rs's public functions (parse_static :976,
// parse_arena :885, write_arena :903) over librsim.so's C ABI (include/rsim_v3.h). Failures throw rsim::Error with the
// reader's message, where the reference returns Err(String).
#pragma once

#include <stdint.h>
#include <stdio.h>

#include <string>
#include <vector>

#include "rsim_v3.h"
#include "sim_device.hpp"

namespace rsim {
namespace v3 {

using Template = rsim_v3_template;

// StaticData words of a V3 static dump.
inline std::vector<uint32_t> parse_static(const std::vector<uint8_t>& bytes) {
  std::vector<uint32_t> sd(rsim_static_words());
  check(rsim_v3_parse_static(bytes.data(), bytes.size(), sd.data()), "parse_static");
  return sd;
}
// ArenaState words of a V3 arena dump, and the template of what the state leaves out.
inline std::vector<uint32_t> parse_arena(const std::vector<uint8_t>& bytes, const std::vector<uint32_t>& sd, Template* tpl) {
  std::vector<uint32_t> s(rsim_arena_words());
  check(rsim_v3_parse_arena(bytes.data(), bytes.size(), sd.data(), s.data(), tpl), "parse_arena");
  return s;
}
// A V3 arena dump of one ArenaState (`state` points at rsim_arena_words() words).
inline std::vector<uint8_t> write_arena(const uint32_t* state, const std::vector<uint32_t>& sd, const Template& tpl) {
  uint64_t len = 0;
  int rc = rsim_v3_write_arena(state, sd.data(), &tpl, nullptr, 0, &len);
  if (rc != RSIM_ERR_SHORT) check(rc, "write_arena");
  std::vector<uint8_t> out(len);
  check(rsim_v3_write_arena(state, sd.data(), &tpl, out.data(), out.size(), &len), "write_arena");
  return out;
}

inline std::vector<uint8_t> read_file(const char* path) {
  FILE* f = fopen(path, "rb");
  if (!f) throw Error(RSIM_ERR_INVALID, std::string("cannot open ") + path);
  std::vector<uint8_t> b;
  uint8_t buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) b.insert(b.end(), buf, buf + n);
  fclose(f);
  return b;
}
inline void write_file(const char* path, const std::vector<uint8_t>& b) {
  FILE* f = fopen(path, "wb");
  if (!f || fwrite(b.data(), 1, b.size(), f) != b.size()) {
    if (f) fclose(f);
    throw Error(RSIM_ERR_INVALID, std::string("cannot write ") + path);
  }
  fclose(f);
}

}  // namespace v3
}  // namespace rsim
