// The synthetic world and start states: the stand-in for what the reference loads from RocketSim V3's exports
// (crates/rocketsim-cuda/src/dump.rs parse_static / parse_arena), which need V3 and the game's meshes. Host-side C++.
//
//  * build_world: a procedural Soccar arena of ~7.4k triangles in 16 meshes (floor/ceiling fillets of 256 uu along the
//    walls, 1024 uu rounded corners, back walls with 1786 x 642.775 goal mouths, 880 uu deep goals) plus 4 planes (floor,
//    ceiling, side walls); a BVH4 per mesh with one leaf per triangle; Bullet-style internal-edge angles; the ray grid with
//    its static cell lists; sweep-and-prune constants; body, car (7 hitbox presets), wheel, pad and mutator constants.
//  * build_arena: a consistent ArenaState for given car and ball poses: bodies, broadphase built by inserting every handle
//    as Bullet's AxisSweep3::addHandle does (sorted edges, the pair list, the dynamic grid cells).
//  * the scenario samplers of the configurations (kickoff, open play, contact-heavy) and the per-decision controls, all
//    from seeded SplitMix64 streams.
#pragma once
#include "core/rs_tick.h"

namespace rs {
namespace world {

constexpr int NUM_PRESETS = 7;  // Octane, Dominus, Plank, Breakout, Hybrid, Merc, Psyclops (three wheels)

struct Mutators {  // the per-arena build parameters (defaults: Soccar)
  float gravity_z = -650.0f;
  float boost_accel_scale = 1.0f;
  float ball_radius = 91.25f;
};

// The static world of one car pairing: car 0 (blue) uses preset p0, car 1 (orange) p1.
void build_world(StaticData* w, int p0, int p1, const Mutators& mut);

struct CarInit {
  V3 pos;             // uu
  float yaw, pitch, roll;
  V3 vel, ang_vel;    // uu/s, rad/s
  float boost;
  bool on_ground;
  bool demoed;
  float demo_timer;
};
struct BallInit { V3 pos, vel, ang_vel; };
// A start state; pad_active[i] false marks pad i as recently used (its cooldown still running), tick_count the arena tick.
void build_arena(ArenaState* s, const StaticData& w, const CarInit cars[NUM_CARS], const BallInit& ball, const bool* pad_active,
                 uint64_t tick_count, uint64_t rng_seed);

// SplitMix64
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }  // [0, 1)
  float uniform(float a, float b) { return (float)(a + (b - a) * unit()); }
  double gauss() {  // Box-Muller, one value
    double u1 = unit(), u2 = unit();
    if (u1 < 1e-300) u1 = 1e-300;
    return __builtin_sqrt(-2.0 * __builtin_log(u1)) * __builtin_cos(6.283185307179586 * u2);
  }
};

enum Scenario { SCEN_KICKOFF = 0, SCEN_OPEN = 1, SCEN_PINCH = 2, SCEN_HEADON = 3, SCEN_GOALFRAME = 4 };
// A start state of the given kind, drawn from rng (the world fixes the presets). Returns the kind drawn.
void sample_arena(ArenaState* s, const StaticData& w, int kind, Rng& rng);
// The controls of one decision for one car: axes uniform in {-1, 0, 1}, jump / boost / handbrake Bernoulli(0.3).
Controls sample_controls(Rng& rng);

}  // namespace world
}  // namespace rs
