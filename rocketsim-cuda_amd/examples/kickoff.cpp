// Steps many copies of a 1v1 kickoff on the GPU (the C++ counterpart of the reference's examples/kickoff.rs): an
// Octane-vs-Octane Soccar arena and a kickoff with the blue car on the centre line; the state is replicated to N arenas
// and advanced by the decision kernel (8 ticks per launch) with both cars driving and boosting straight ahead.
//   kickoff [arenas] [decisions] [--static FILE --arena FILE] [--out FILE]
// Defaults: 16384 arenas, 300 decisions (300 x 8 ticks = 20 s of game time). The world and the kickoff come from the world
// builder, or, as in the reference, from V3 dumps (--static: V3's export_static, --arena: its export_state); --out writes
// arena 0's final state as a V3 arena dump V3 can import. Built by rsim/build.py next to librsim.so.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "../host/sim_device.hpp"
#include "../host/v3_dump.hpp"

int main(int argc, char** argv) {
  const char *static_path = nullptr, *arena_path = nullptr, *out_path = nullptr;
  std::vector<const char*> pos;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--static") && i + 1 < argc) static_path = argv[++i];
    else if (!strcmp(argv[i], "--arena") && i + 1 < argc) arena_path = argv[++i];
    else if (!strcmp(argv[i], "--out") && i + 1 < argc) out_path = argv[++i];
    else pos.push_back(argv[i]);
  }
  const uint32_t arenas = pos.size() > 0 ? (uint32_t)atoi(pos[0]) : 16384;
  const uint32_t decisions = pos.size() > 1 ? (uint32_t)atoi(pos[1]) : 300;
  if ((static_path == nullptr) != (arena_path == nullptr)) {
    fprintf(stderr, "kickoff: --static and --arena go together\n");
    return 2;
  }
  try {
    const uint64_t aw = rsim_arena_words();
    std::vector<uint32_t> sd, one;
    rsim::v3::Template tpl = {};
    if (static_path) {  // the reference's path: V3's exports through the dump parser
      sd = rsim::v3::parse_static(rsim::v3::read_file(static_path));
      one = rsim::v3::parse_arena(rsim::v3::read_file(arena_path), sd, &tpl);
    } else {
      sd.assign(rsim_static_words(), 0u);
      rsim::check(rsim_build_world(0, 0, -650.0f, 1.0f, 91.25f, sd.data()), "build_world");
      // kickoff seeds pick the spawn spot; take one with the blue car on the centre line (x = 0), driving at the ball
      one.assign(aw, 0u);
      const uint32_t car0_pos_x = 24 + 5;  // ArenaState word of cars[0].pos.x
      uint64_t seed = 0;
      for (;; ++seed) {
        rsim::check(rsim_sample_arenas(sd.data(), RSIM_SCEN_KICKOFF, seed, 0, 1, one.data()), "sample_arenas");
        float x;
        memcpy(&x, &one[car0_pos_x], 4);
        if (x == 0.0f) break;
      }
    }
    std::vector<uint32_t> states((size_t)arenas * aw);
    for (uint32_t a = 0; a < arenas; ++a) memcpy(&states[(size_t)a * aw], one.data(), aw * 4);

    rsim::SimDevice dev(0);
    rsim::DeviceStatic dsd = dev.upload_static(sd);
    rsim::DeviceStates d = dev.upload_states(states);
    float one_f = 1.0f;
    uint32_t drive[8] = {0, 0, 0, 0, 0, 0, 1, 0};  // throttle 1, boost 1
    memcpy(&drive[0], &one_f, 4);
    std::vector<uint32_t> ctl((size_t)arenas * rsim::CONTROL_WORDS);
    for (uint32_t a = 0; a < arenas; ++a)
      for (int c = 0; c < 2; ++c) memcpy(&ctl[(size_t)a * 16 + c * 8], drive, sizeof drive);
    rsim::DeviceWords controls = dev.upload_controls(ctl);
    rsim::DeviceWords result = dev.alloc_words(arenas);

    dev.sim8(d, dsd, controls, result);  // warm-up
    dev.synchronize();
    auto t0 = std::chrono::steady_clock::now();
    for (uint32_t k = 0; k < decisions; ++k) dev.sim8(d, dsd, controls, result);
    dev.synchronize();
    double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    std::vector<uint32_t> out = dev.download_states(d);
    float ball[3];
    memcpy(ball, &out[4], 12);  // ArenaState.ball.pos (uu) follows tick_count and rng (two words each)
    printf("%u arenas, %u decisions of %u ticks, %.1f game seconds\n", arenas, decisions, rsim::TICKS_PER_DECISION,
           (decisions + 1) * (double)rsim::TICKS_PER_DECISION / 120.0);
    printf("arena 0 ball position (uu): (%.1f, %.1f, %.1f)\n", ball[0], ball[1], ball[2]);
    printf("%.2f M arena-ticks/s (%.2f s)\n", arenas * (double)decisions * rsim::TICKS_PER_DECISION / secs / 1e6, secs);
    if (out_path) {
      if (!static_path) throw rsim::Error(RSIM_ERR_INVALID, "--out needs the template of a parsed dump (--static, --arena)");
      rsim::v3::write_file(out_path, rsim::v3::write_arena(out.data(), sd, tpl));
      printf("arena 0 written as a V3 arena dump: %s\n", out_path);
    }
  } catch (const rsim::Error& e) {
    fprintf(stderr, "kickoff: %s (status %d)\n", e.what(), e.code);
    return 1;
  }
  return 0;
}
