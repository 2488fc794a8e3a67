// The V3 wire format (include/rsim_v3.h): V3 arena and static dumps <-> ArenaState / StaticData. Host code.
//
// The format is a flat little-endian stream: a header, then tagged sections whose fields follow V3's own dump traversal
// (v3/rocketsim-dump.patch, rocketsim/src/dump.rs there) and the reader of the reference (crates/rocketsim-cuda/src/
// dump.rs). Each section is written here once, as a template over the direction (Writer or Reader), so that what
// rsim_v3_write_arena emits is by construction what rsim_v3_parse_arena reads. Sizes in the stream: a bool or V3 enum
// is one byte, an index or count a u32, an optional a flag byte then its payload at full size (zero when absent), a
// sequence a u32 count then its elements.
//
// What the reader refuses (it never drops state silently), in dump.rs's words: a list over its cap, Heatseeker /
// Dropshot / Snowday fields off their defaults, a ball-world contact record list, a non-empty free list, thresholds,
// pair keys or lookup tables that are not the ones the pool implies, static broadphase proxies that differ from the
// world's, and any truncation or trailing bytes.
#include "rsim_v3.h"

#include <math.h>
#include <string.h>

#include <memory>
#include <string>
#include <vector>

#include "core/rs_state.h"
#include "host/errors.h"

static_assert(__BYTE_ORDER__ == __ORDER_LITTLE_ENDIAN__, "the wire format is little-endian, copied as such");
static_assert(RSIM_V3_BODIES == rs::NUM_BODIES, "one template record per body");
static_assert(RSIM_V3_RECORD_WORDS * 4 == sizeof(rs::ContactRecord), "record words");

namespace rs {
namespace wire {
namespace {

constexpr uint8_t ARENA_MAGIC[8] = {'V', '3', 'A', 'R', 'E', 'N', 'A', 0};
constexpr uint8_t STATIC_MAGIC[8] = {'V', '3', 'S', 'T', 'A', 'T', 'I', 'C'};
constexpr size_t BODY_RECORD = RSIM_V3_BODY_RECORD_BYTES;
constexpr size_t COMPANION_AT = 1, COMPANION_BYTES = 5;  // the companion_id option inside a body record (dump.rs:22-23)
constexpr size_t ISLAND_WIDTHS[6] = {4, 4, 8, 8, 4, 4};   // element bytes of the island manager's six sequences

// UserInfoType (user_info_types.rs) of StaticData.bodies[].user_idx
constexpr uint32_t UI_NONE = 0, UI_CAR = 1, UI_BALL = 2;

uint32_t fbits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

// ---- the two directions ----------------------------------------------------------------------------------------------------
struct Writer {
  static constexpr bool READ = false;
  std::vector<uint8_t> out;
  std::string err;
  void raw(void* p, size_t n) {
    const uint8_t* b = (const uint8_t*)p;
    out.insert(out.end(), b, b + n);
  }
  // a sequence's element count; returns what the traversal iterates over
  size_t count(size_t n) {
    uint32_t x = (uint32_t)n;
    raw(&x, 4);
    return n;
  }
  void fail(const std::string& m) {
    if (err.empty()) err = m;
  }
};

// Reads from a byte range. The first failure is kept and ends the stream: every later read yields zeros and every
// count 0, so a traversal that carries on after a failure touches nothing.
struct Reader {
  static constexpr bool READ = true;
  const uint8_t* b;
  size_t n, p = 0;
  std::string err;
  Reader(const uint8_t* bytes, size_t len) : b(bytes), n(len) {}
  size_t remaining() const { return n - p; }
  void raw(void* dst, size_t k) {
    if (!err.empty() || remaining() < k) {
      if (err.empty()) err = "dump truncated";
      p = n;
      memset(dst, 0, k);
      return;
    }
    memcpy(dst, b + p, k);
    p += k;
  }
  void skip(size_t k) {
    if (!err.empty() || remaining() < k) {
      fail("dump truncated");
      return;
    }
    p += k;
  }
  size_t count(size_t) {
    uint32_t x = 0;
    raw(&x, 4);
    if (x > remaining()) {  // every element takes at least a byte
      fail("count exceeds the bytes left");
      return 0;
    }
    return x;
  }
  void fail(const std::string& m) {
    if (err.empty()) err = m;
    p = n;
  }
};

// ---- field codecs ------------------------------------------------------------------------------------------------------------
template <class S> void u8f(S& s, uint8_t& v) { s.raw(&v, 1); }
template <class S> void u16f(S& s, uint16_t& v) { s.raw(&v, 2); }
template <class S> void u32f(S& s, uint32_t& v) { s.raw(&v, 4); }
template <class S> void u64f(S& s, uint64_t& v) { s.raw(&v, 8); }
template <class S> void f32f(S& s, float& v) { s.raw(&v, 4); }
// a bool as one byte: written 0/1, read back as the byte's value
template <class S> void flag(S& s, uint32_t& v) {
  uint8_t x = v != 0;
  s.raw(&x, 1);
  v = x;
}
// a V3 enum or small integer held in one byte
template <class S> void byte(S& s, uint32_t& v) {
  uint8_t x = (uint8_t)v;
  s.raw(&x, 1);
  v = x;
}
// a u16 held in a u32 word
template <class S> void half(S& s, uint32_t& v) {
  uint16_t x = (uint16_t)v;
  s.raw(&x, 2);
  v = x;
}
template <class S> void vec3(S& s, V3& v) {
  f32f(s, v.x);
  f32f(s, v.y);
  f32f(s, v.z);
}
template <class S> void mat3(S& s, M3& m) {
  vec3(s, m.c0);
  vec3(s, m.c1);
  vec3(s, m.c2);
}
template <class S> void aff3(S& s, Aff& a) {
  mat3(s, a.m);
  vec3(s, a.t);
}
template <class S> void quat(S& s, Quat& q) {
  f32f(s, q.x);
  f32f(s, q.y);
  f32f(s, q.z);
  f32f(s, q.w);
}
template <class S> void aabb(S& s, Aabb& a) {
  vec3(s, a.min);
  vec3(s, a.max);
}
template <class S> void u64w(S& s, U64w& v) {
  uint64_t x = u64(v);
  u64f(s, x);
  v = rs::u64w(x);
}
template <class S, class T, class F> void opt(S& s, Opt<T>& o, F f) {
  flag(s, o.some);
  f(s, o.v);
}
template <class S> void tag(S& s, const char* t) {
  uint32_t want, got;
  memcpy(&want, t, 4);
  got = want;
  u32f(s, got);
  if (got != want) s.fail(std::string("section tag ") + std::string(t, 4) + " expected");
}
template <class S> bool expect_count(S& s, size_t got, size_t want, const char* what) {
  if (got == want) return true;
  s.fail(std::string(what) + ": " + std::to_string(got) + ", StaticData holds " + std::to_string(want));
  return false;
}

// ---- arena pieces (dump.rs:241-371) ----------------------------------------------------------------------------------------
template <class S> void controls(S& s, Controls& c) {
  f32f(s, c.throttle);
  f32f(s, c.steer);
  f32f(s, c.pitch);
  f32f(s, c.yaw);
  f32f(s, c.roll);
  flag(s, c.jump);
  flag(s, c.boost);
  flag(s, c.handbrake);
}

template <class S> void hit(S& s, HitInfo& h) {
  vec3(s, h.hit_point);
  vec3(s, h.hit_normal);
  f32f(s, h.hit_fraction);
  byte(s, h.user_info);
}

// A manifold point; V3's applied_impulse is never carried across a tick (zero in every dump) and ArenaState has no slot.
template <class S> void point(S& s, ManifoldPoint& p) {
  vec3(s, p.local_a);
  vec3(s, p.local_b);
  vec3(s, p.world_on_b);
  vec3(s, p.world_on_a);
  vec3(s, p.normal_on_b);
  f32f(s, p.distance);
  f32f(s, p.friction);
  f32f(s, p.restitution);
  float applied = 0.0f;
  f32f(s, applied);
  if (S::READ && fbits(applied) != 0) s.fail("manifold point with a non-zero applied_impulse");
  flag(s, p.is_special);
}

template <class S> void wheel(S& s, Wheel& w) {
  opt(s, w.ray, [](S& s, RaycastInfo& r) {
    vec3(s, r.contact_normal);
    vec3(s, r.contact_point);
    u32f(s, r.ground_body);
    f32f(s, r.suspension_length);
    vec3(s, r.impulse);
    opt(s, r.ground_stick, [](S& s, V3& v) { vec3(s, v); });
    flag(s, r.in_contact_with_world);
    f32f(s, r.clipped_inv_contact_dot_suspension);
    f32f(s, r.suspension_relative_vel);
  });
  vec3(s, w.hard_point);
  vec3(s, w.axle_dir);
  f32f(s, w.engine_force);
  f32f(s, w.brake);
  f32f(s, w.steer_angle);
  vec3(s, w.vel_at_contact_point);
  f32f(s, w.lat_friction);
  f32f(s, w.long_friction);
  f32f(s, w.extra_pushback);
}

template <class S> void car(S& s, Car& c) {
  f32f(s, c.dodge_deadzone);
  vec3(s, c.vel_impulse_cache);
  flag(s, c.sticky_gate_prev);
  vec3(s, c.pos);
  mat3(s, c.rot);
  vec3(s, c.vel);
  vec3(s, c.ang_vel);
  controls(s, c.controls);
  flag(s, c.has_jumped);
  flag(s, c.has_double_jumped);
  flag(s, c.has_flipped);
  vec3(s, c.flip_rel_torque);
  u32f(s, c.jump_ticks);
  f32f(s, c.flip_time);
  flag(s, c.is_flipping);
  flag(s, c.is_jumping);
  flag(s, c.jump_hold_broken);
  f32f(s, c.air_time);
  f32f(s, c.air_time_since_jump);
  f32f(s, c.boost);
  f32f(s, c.handbrake_val);
  flag(s, c.is_supersonic);
  f32f(s, c.supersonic_grace_timer);
  f32f(s, c.time_since_boosted);
  f32f(s, c.boosting_time);
  flag(s, c.is_boosting);
  flag(s, c.is_auto_flipping);
  f32f(s, c.auto_flip_timer);
  f32f(s, c.auto_flip_torque_scale);
  f32f(s, c.bump_cooldown_timer);
  opt(s, c.bump_cooldown_victim, [](S& s, uint32_t& v) { u32f(s, v); });
  opt(s, c.last_extra_hit_tick, [](S& s, U64w& v) { u64w(s, v); });
  flag(s, c.is_demoed);
  f32f(s, c.demo_respawn_timer);
  flag(s, c.is_on_ground);
  for (auto& h : c.wheels_with_contact) opt(s, h, [](S& s, HitInfo& x) { hit(s, x); });
  opt(s, c.world_contact_normal, [](S& s, V3& v) { vec3(s, v); });
  controls(s, c.prev_controls);
  for (auto& w : c.wheels) wheel(s, w);
}

// One BODY record of a dynamic body: exactly RSIM_V3_BODY_RECORD_BYTES bytes.
template <class S> void body(S& s, Body& b) {
  byte(s, b.collision_flags);
  // companion_id, the solver-body link: reassigned by the solver's setup before any read, so not state (written none)
  Opt<uint32_t> companion{0, 0};
  opt(s, companion, [](S& s, uint32_t& v) { u32f(s, v); });
  byte(s, b.activation);
  aff3(s, b.world);
  quat(s, b.world_quat);
  aff3(s, b.interp_world);
  f32f(s, b.deactivation_time);
  vec3(s, b.lin_vel);
  vec3(s, b.ang_vel);
  mat3(s, b.inv_inertia_world);
  vec3(s, b.accum_lin_vel);
  vec3(s, b.accum_ang_vel);
  f32f(s, b.cached_linear_damping);
  f32f(s, b.cached_damping_time_step);
  f32f(s, b.cached_linear_damping_value);
}

// Index of the triangular pair table for a proxy pair (overlapping_pair_cache.rs).
uint64_t pair_table_index(uint32_t a, uint32_t b) {
  uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
  return hi * (hi + 1) / 2 + lo;
}

int dyn_slot(const StaticData& w, uint32_t body) {
  for (int d = 0; d < NUM_DYN; ++d)
    if (w.dyn_body[d] == body) return d;
  return -1;
}
int dyn_handle_slot(const StaticData& w, uint32_t handle) {
  for (int d = 0; d < NUM_DYN; ++d)
    if (w.dyn_handle[d] == handle) return d;
  return -1;
}

// ---- the record fold (dump.rs:395-447) -------------------------------------------------------------------------------------
// V3 lists every contact record of a tick in call order; ArenaState keeps what the tick's P7 reads of them. Car-ball and
// car-car records (user a Car, user b Ball or Car) are stored in order; a car-world record sets bit EV_CAR_HIT_WORLD + car
// of fold_events and makes its normal the car's fold_car_normal (the last one wins); a ball-world record (user a Ball,
// user b None) sets EV_BALL_HIT_WORLD, and EV_BALL_WALL as well when |normal.z| < 0.8; anything else is a no-op in P7.
std::string fold(ArenaState& a, const StaticData& w, const ContactRecord* recs, size_t n) {
  a.fold_events = 0;
  for (auto& v : a.fold_car_normal) v = V3{0, 0, 0};
  a.n_records = 0;
  memset(a.records, 0, sizeof a.records);
  for (size_t i = 0; i < n; ++i) {
    const ContactRecord& r = recs[i];
    if (r.body_a >= (uint32_t)NUM_BODIES || r.body_b >= (uint32_t)NUM_BODIES) return "contact record body index out of range";
    const BodyConst& ba = w.bodies[r.body_a];
    const BodyConst& bb = w.bodies[r.body_b];
    if (ba.user_idx == UI_CAR && (bb.user_idx == UI_BALL || bb.user_idx == UI_CAR)) {
      if (a.n_records >= (uint32_t)MAX_RECORDS) return "stored contact records over the cap";
      a.records[a.n_records++] = r;
    } else if (ba.user_idx == UI_CAR) {
      if (ba.user_pointer >= (uint32_t)NUM_CARS) return "car body with a car index out of range";
      a.fold_events |= 1u << (EV_CAR_HIT_WORLD + ba.user_pointer);
      a.fold_car_normal[ba.user_pointer] = r.pt.normal_on_b;
    } else if (ba.user_idx == UI_BALL && bb.user_idx == UI_NONE) {
      a.fold_events |= 1u << EV_BALL_HIT_WORLD;
      if (fabsf(r.pt.normal_on_b.z) < 0.8f) a.fold_events |= 1u << EV_BALL_WALL;
    }
  }
  return std::string();
}

// Records that fold back to the state's fold fields: one ball-world record whose normal gives the same two bits, one
// car-world record per car flagged, with its normal, then the stored records (P7's result does not depend on the order
// of the folded kinds). Empty with an error when the world has no static body to be the "world" side.
std::string unfold(const ArenaState& a, const StaticData& w, std::vector<ContactRecord>& out) {
  out.clear();
  int world_body = -1;
  for (int i = 0; i < NUM_BODIES && world_body < 0; ++i)
    if (w.bodies[i].user_idx == UI_NONE) world_body = i;
  if (world_body < 0) return "the world has no static body";
  if (a.n_records > (uint32_t)MAX_RECORDS) return "stored contact records over the cap";
  auto rec = [&](uint32_t body_a, V3 n) {
    ContactRecord r;
    memset(&r, 0, sizeof r);
    r.body_a = body_a;
    r.body_b = (uint32_t)world_body;
    r.pt.normal_on_b = n;
    out.push_back(r);
  };
  if (a.fold_events >> EV_BALL_HIT_WORLD & 1)
    rec(w.dyn_body[0], (a.fold_events >> EV_BALL_WALL & 1) ? V3{1, 0, 0} : V3{0, 0, 1});
  for (int c = 0; c < NUM_CARS; ++c)
    if (a.fold_events >> (EV_CAR_HIT_WORLD + c) & 1) rec(w.dyn_body[1 + c], a.fold_car_normal[c]);
  for (uint32_t i = 0; i < a.n_records; ++i) out.push_back(a.records[i]);
  return std::string();
}

// ---- the arena traversal (dump.rs:449-869) ---------------------------------------------------------------------------------
template <class S> void arena_io(S& s, ArenaState& a, const StaticData& w, rsim_v3_template& t) {
  // ARNA
  tag(s, "ARNA");
  u64w(s, a.tick_count);
  u64w(s, a.rng);
  uint32_t ball_only = 0;  // always 0 in 1v1
  flag(s, ball_only);
  if (S::READ && ball_only) s.fail("a ball-only arena");

  // BALL; then Heatseeker's and Dropshot's ball info (24 bytes, carried by the template) and Snowday's ground stick
  tag(s, "BALL");
  vec3(s, a.ball.pos);
  mat3(s, a.ball.rot);
  vec3(s, a.ball.vel);
  vec3(s, a.ball.ang_vel);
  uint8_t other[24];
  memcpy(other, t.ball_other, 24);
  s.raw(other, 24);
  if (S::READ) {
    if (!t.ball_other_set) {
      memcpy(t.ball_other, other, 24);
      t.ball_other_set = 1;
    } else if (memcmp(other, t.ball_other, 24) != 0) {
      s.fail("Heatseeker or Dropshot ball info changed between dumps");
    }
  }
  u64w(s, a.ball.tick_count_since_kickoff);
  uint8_t ground_stick = 0;
  u8f(s, ground_stick);
  if (S::READ && ground_stick) s.fail("Snowday ground_stick_applied is set");

  // CARS
  tag(s, "CARS");
  if (s.count(NUM_CARS) != (size_t)NUM_CARS) {
    s.fail("not a 2-car arena");
    return;
  }
  for (auto& c : a.cars) car(s, c);

  // PADS: per pad the tick it last gave boost (an i64 option) and the cars touching it, in arrival order
  tag(s, "PADS");
  if (s.count(NUM_PADS) != (size_t)NUM_PADS) {
    s.fail("not a 34-pad arena");
    return;
  }
  for (int i = 0; i < NUM_PADS; ++i) {
    opt(s, a.pad_gave_boost[i], [](S& s, U64w& v) { u64w(s, v); });
    PadTouching& pt = a.pad_touching[i];
    size_t n = s.count(pt.n);
    if (n > (size_t)NUM_CARS) {
      s.fail("a pad touched by more cars than the arena has");
      return;
    }
    pt.n = (uint32_t)n;
    for (size_t k = 0; k < n; ++k) u32f(s, pt.car_idx[k]);
  }

  // CONT: every contact record of the tick, in call order (only the two world points and the normal are kept)
  tag(s, "CONT");
  std::vector<ContactRecord> list;
  if (!S::READ) {
    std::string e = unfold(a, w, list);
    if (!e.empty()) {
      s.fail(e);
      return;
    }
  }
  size_t n_rec = s.count(list.size());
  if (S::READ) list.assign(n_rec, ContactRecord{});
  for (auto& r : list) {
    flag(s, r.is_swap);
    u32f(s, r.body_a);
    u32f(s, r.body_b);
    ManifoldPoint p;
    memset(&p, 0, sizeof p);
    p.world_on_b = r.pt.world_on_b;
    p.world_on_a = r.pt.world_on_a;
    p.normal_on_b = r.pt.normal_on_b;
    point(s, p);
    r.pt = RecordPoint{p.world_on_b, p.world_on_a, p.normal_on_b};
  }
  if (S::READ) {
    std::string e = fold(a, w, list.data(), list.size());
    if (!e.empty()) {
      a.overflow[OVF_RECORDS] += 1;
      s.fail(e);
      return;
    }
  }
  if (s.count(0) != 0) {  // ball-world records: ball-only arenas only
    s.fail("ball-world contact records present");
    return;
  }
  uint32_t ball_only_tracker = 0;
  flag(s, ball_only_tracker);

  // BODY: the dynamic bodies are state; the static bodies' records are constants, carried by the template
  tag(s, "BODY");
  if (s.count(NUM_BODIES) != (size_t)NUM_BODIES) {
    s.fail("not a 23-body arena");
    return;
  }
  for (int i = 0; i < NUM_BODIES; ++i) {
    int d = dyn_slot(w, (uint32_t)i);
    if (d >= 0) {
      body(s, a.bodies[d]);
    } else {
      uint8_t rec[BODY_RECORD];
      memcpy(rec, t.static_body[i], BODY_RECORD);
      s.raw(rec, BODY_RECORD);
      if (S::READ) {
        memset(rec + COMPANION_AT, 0, COMPANION_BYTES);
        bool empty = true;
        for (size_t k = 0; k < BODY_RECORD && empty; ++k) empty = t.static_body[i][k] == 0;
        if (empty) memcpy(t.static_body[i], rec, BODY_RECORD);  // static bodies never change: keep the first seen
      }
    }
  }

  // DISP: the retained manifolds, the active list, and the dense manifold table (a search of the pool)
  tag(s, "DISP");
  size_t nm = s.count(a.n_manifolds);
  if (S::READ) {
    if (nm > (size_t)MAX_MANIFOLDS) {
      a.overflow[OVF_MANIFOLDS] += 1;
      s.fail("retained manifolds over the cap");
      return;
    }
    a.n_manifolds = (uint32_t)nm;
  }
  for (size_t i = 0; i < nm; ++i) {
    Manifold& m = a.manifolds[i];
    u32f(s, m.body0);
    u32f(s, m.body1);
    uint32_t lo = m.body0 < m.body1 ? m.body0 : m.body1, hi = m.body0 < m.body1 ? m.body1 : m.body0;
    uint64_t key = (uint64_t)lo << 32 | hi, got = key;
    u64f(s, got);
    if (S::READ && got != key) s.fail("manifold pair_key is not min << 32 | max");
    if (m.body0 >= (uint32_t)NUM_BODIES || m.body1 >= (uint32_t)NUM_BODIES) {
      s.fail("manifold body index out of range");
      return;
    }
    // both thresholds are the min over the manifold's two bodies (persistent_manifold.rs:60-65)
    const BodyConst &b0 = w.bodies[m.body0], &b1 = w.bodies[m.body1];
    float cbt = fminf(b0.contact_breaking_threshold, b1.contact_breaking_threshold);
    float cpt = fminf(b0.contact_processing_threshold, b1.contact_processing_threshold);
    const uint32_t want_cbt = fbits(cbt), want_cpt = fbits(cpt);
    f32f(s, cbt);
    f32f(s, cpt);
    if (S::READ && (fbits(cbt) != want_cbt || fbits(cpt) != want_cpt)) {
      s.fail("a manifold threshold is not the min over its two bodies");
      return;
    }
    u32f(s, m.n_points);
    if (m.n_points > (uint32_t)MAX_POINTS) {
      s.fail("manifold with more than 4 points");
      return;
    }
    for (auto& p : m.pts) point(s, p);
  }
  size_t na = s.count(a.n_active);
  if (S::READ) {
    if (na > (size_t)MAX_ACTIVE) {
      a.overflow[OVF_ACTIVE] += 1;
      s.fail("active manifolds over the cap");
      return;
    }
    a.n_active = (uint32_t)na;
  }
  for (size_t i = 0; i < na; ++i) u32f(s, a.active[i]);
  // manifold_stride and manifold_table: the table is rebuilt from the pool, so it is only checked on read
  uint32_t stride = S::READ ? t.manifold_stride : (nm > 0 ? 32u : t.manifold_stride);
  u32f(s, stride);
  const uint64_t cells = (uint64_t)stride * stride;
  size_t tl = s.count((size_t)cells);
  if (S::READ) {
    t.manifold_stride = stride;
    if ((uint64_t)tl != cells) {
      s.fail("manifold table length is not stride squared");
      return;
    }
  }
  std::vector<uint32_t> table(tl, 0u);
  for (size_t i = 0; i < nm; ++i) {
    const Manifold& m = a.manifolds[i];
    uint32_t lo = m.body0 < m.body1 ? m.body0 : m.body1, hi = m.body0 < m.body1 ? m.body1 : m.body0;
    if (hi >= stride) {
      s.fail("manifold body index outside the table");
      return;
    }
    table[(size_t)lo * stride + hi] = (uint32_t)i + 1;
  }
  for (size_t i = 0; i < tl; ++i) {
    uint32_t e = table[i];
    u32f(s, e);
    if (S::READ && e != table[i]) {
      s.fail("manifold table is not the search of the pool");
      return;
    }
  }

  // BROA: the dynamic ray-cell lists (cells ascending; each: the cell, then its handles in list order)
  tag(s, "BROA");
  if (S::READ) {
    size_t ncells = s.count(0);
    std::vector<DynCell> entries;
    bool have_last = false;
    uint32_t last = 0;
    for (size_t c = 0; c < ncells; ++c) {
      uint32_t cell = 0;
      u32f(s, cell);
      if (have_last && last >= cell) {
        s.fail("dynamic ray cells not ascending");
        return;
      }
      have_last = true;
      last = cell;
      size_t nh = s.count(0);
      for (size_t k = 0; k < nh; ++k) {
        uint32_t h = 0;
        u32f(s, h);
        entries.push_back(DynCell{cell, h});
      }
    }
    if (entries.size() > (size_t)MAX_DYN_CELLS) {
      a.overflow[OVF_DYN_CELLS] += 1;
      s.fail("dynamic ray-cell entries over the cap");
      return;
    }
    a.n_dyn_cells = (uint32_t)entries.size();
    memset(a.dyn_cells, 0, sizeof a.dyn_cells);
    if (!entries.empty()) memcpy(a.dyn_cells, entries.data(), entries.size() * sizeof(DynCell));
  } else {
    size_t groups = 0;
    for (uint32_t i = 0; i < a.n_dyn_cells; ++i) groups += (i == 0 || a.dyn_cells[i].cell != a.dyn_cells[i - 1].cell);
    s.count(groups);
    for (uint32_t i = 0; i < a.n_dyn_cells;) {
      uint32_t j = i;
      while (j < a.n_dyn_cells && a.dyn_cells[j].cell == a.dyn_cells[i].cell) ++j;
      uint32_t cell = a.dyn_cells[i].cell;
      u32f(s, cell);
      s.count(j - i);
      for (uint32_t k = i; k < j; ++k) {
        uint32_t h = a.dyn_cells[k].handle;
        u32f(s, h);
      }
      i = j;
    }
  }
  // the sweep-and-prune handles: the static proxies are the world's constants, the dynamic ones are state
  if (s.count(NUM_HANDLES) != (size_t)NUM_HANDLES) {
    s.fail("not a 24-handle broadphase");
    return;
  }
  for (int h = 0; h < NUM_HANDLES; ++h) {
    const ProxyConst& k = w.proxies[h];
    ProxyConst p = k;
    u32f(s, p.client_body);
    byte(s, p.filter_group);
    byte(s, p.filter_mask);
    u32f(s, p.unique_id);
    int d = dyn_handle_slot(w, (uint32_t)h);
    if (d >= 0) {
      DynProxy& dp = a.dyn_proxy[d];
      aabb(s, dp.aabb);
      u32f(s, dp.cell_idx);
      for (auto& i : dp.indices) u32f(s, i);
    } else {
      ProxyConst tmp = k;
      aabb(s, tmp.aabb);
      u32f(s, tmp.cell_idx);
      for (auto& i : tmp.indices) u32f(s, i);
      const bool same_box = tmp.aabb.min.x == k.aabb.min.x && tmp.aabb.min.y == k.aabb.min.y &&
                            tmp.aabb.min.z == k.aabb.min.z && tmp.aabb.max.x == k.aabb.max.x &&
                            tmp.aabb.max.y == k.aabb.max.y && tmp.aabb.max.z == k.aabb.max.z;
      if (S::READ && (!same_box || tmp.cell_idx != k.cell_idx || memcmp(tmp.indices, k.indices, sizeof k.indices) != 0)) {
        s.fail("a static broadphase proxy differs from StaticData.proxies");
        return;
      }
    }
    if (S::READ && (p.client_body != k.client_body || p.filter_group != k.filter_group || p.filter_mask != k.filter_mask ||
                    p.unique_id != k.unique_id)) {
      s.fail("a broadphase proxy constant differs from StaticData.proxies");
      return;
    }
  }
  if (s.count(NUM_HANDLES) != (size_t)NUM_HANDLES) {
    s.fail("handle_edges length");
    return;
  }
  for (auto& e : a.handle_edges) {
    for (auto& v : e.min) u16f(s, v);
    for (auto& v : e.max) u16f(s, v);
  }
  for (auto& axis : a.edges) {
    if (s.count(NUM_EDGES) != (size_t)NUM_EDGES) {
      s.fail("edge array length");
      return;
    }
    for (auto& e : axis) {
      u16f(s, e.pos);
      u16f(s, e.handle);
    }
  }
  uint16_t num_handles = t.num_handles, next_new_handle = t.next_new_handle;
  u16f(s, num_handles);
  if (s.count(0) != 0) {
    s.fail("free_list is not empty");
    return;
  }
  u16f(s, next_new_handle);
  if (S::READ) {
    t.num_handles = num_handles;
    t.next_new_handle = next_new_handle;
  }
  size_t np = s.count(a.n_pairs);
  if (S::READ) {
    if (np > (size_t)MAX_PAIRS) {
      a.overflow[OVF_PAIRS] += 1;
      s.fail("retained pairs over the cap");
      return;
    }
    a.n_pairs = (uint32_t)np;
  }
  for (size_t i = 0; i < np; ++i) {
    u32f(s, a.pairs[i].proxy0);
    u32f(s, a.pairs[i].proxy1);
  }
  // pair_indices: the triangular table is a search of the pair array (length: the largest index in use + 1)
  if (S::READ) {
    size_t len = s.count(0);
    size_t seen = 0;
    for (size_t i = 0; i < len; ++i) {
      uint32_t g = 0;
      u32f(s, g);
      if (!g) continue;
      ++seen;
      bool found = false;
      for (size_t k = 0; k < np && !found; ++k)
        found = pair_table_index(a.pairs[k].proxy0, a.pairs[k].proxy1) == i && g == k + 1;
      if (!found) {
        s.fail("pair index table is not the search of the pair array");
        return;
      }
    }
    if (seen != np) {
      s.fail("pair index table and pair array disagree");
      return;
    }
  } else {
    uint64_t len = 0;
    for (size_t k = 0; k < np; ++k) {
      if (a.pairs[k].proxy0 >= (uint32_t)NUM_HANDLES || a.pairs[k].proxy1 >= (uint32_t)NUM_HANDLES) {
        s.fail("pair proxy index out of range");
        return;
      }
      uint64_t i = pair_table_index(a.pairs[k].proxy0, a.pairs[k].proxy1) + 1;
      if (i > len) len = i;
    }
    std::vector<uint32_t> pt(len, 0u);
    for (size_t k = 0; k < np; ++k) pt[pair_table_index(a.pairs[k].proxy0, a.pairs[k].proxy1)] = (uint32_t)k + 1;
    s.count(len);
    for (auto& e : pt) u32f(s, e);
  }
}

// ISLD: the island manager is a memo of the manifold sort: skipped on read (two u32 sequences, two i64 sequences, two
// u32 caches, one bool), written empty (dump.rs:851-869).
template <class S> void island_io(S& s) {
  tag(s, "ISLD");
  for (size_t width : ISLAND_WIDTHS) {
    size_t n = s.count(0);
    if constexpr (S::READ) s.skip(n * width);
  }
  uint32_t ball_only = 0;
  flag(s, ball_only);
}

// ---- the static traversal (dump.rs:1024-1300) ------------------------------------------------------------------------------
// `totals` (read only): the triangle, node, leaf and cell-entry totals the header announces.
struct Totals {
  size_t tris, nodes, leaves, cell_entries;
};

template <class S> void static_io(S& s, StaticData& w, const Totals& want) {
  // WRLD
  tag(s, "WRLD");
  byte(s, w.game_mode);
  flag(s, w.no_ball_rot);
  u32f(s, w.ball_body);
  vec3(s, w.gravity_uu);
  for (auto& f : w.mut_f) f32f(s, f);
  flag(s, w.unlimited_flips);
  flag(s, w.unlimited_double_jumps);
  flag(s, w.recharge_boost_enabled);
  f32f(s, w.recharge_boost_per_second);
  f32f(s, w.recharge_boost_delay);
  byte(s, w.demo_mode);
  flag(s, w.enable_team_demos);
  f32f(s, w.goal_base_threshold_y);
  flag(s, w.boost_pads_enabled);
  vec3(s, w.gravity_bt);
  if (!expect_count(s, s.count(NUM_DYN), NUM_DYN, "dynamic bodies")) return;
  for (auto& d : w.dyn_body) u32f(s, d);
  vec3(s, w.grid_min);
  vec3(s, w.grid_max);
  f32f(s, w.grid_cell_size);
  f32f(s, w.grid_cell_size_sq);
  for (auto& v : w.grid_num_cells) u32f(s, v);
  vec3(s, w.sap_world_min);
  vec3(s, w.sap_quantize);
  half(s, w.sap_handle_mask);
  half(s, w.sap_sentinel);
  half(s, w.sap_max_handles);
  f32f(s, w.sap_world_half_extent);

  // BODY
  tag(s, "BODY");
  if (!expect_count(s, s.count(NUM_BODIES), NUM_BODIES, "bodies")) return;
  uint32_t mesh_tris[NUM_BODIES] = {};
  for (int i = 0; i < NUM_BODIES; ++i) {
    BodyConst& b = w.bodies[i];
    u32f(s, b.world_array_idx);
    byte(s, b.user_idx);
    u32f(s, b.user_pointer);
    u32f(s, b.broadphase_handle);
    byte(s, b.filter_group);
    byte(s, b.filter_mask);
    f32f(s, b.mass);
    f32f(s, b.inv_mass);
    vec3(s, b.inv_inertia_local);
    vec3(s, b.inv_mass_splat);
    f32f(s, b.friction);
    f32f(s, b.restitution);
    f32f(s, b.contact_processing_threshold);
    f32f(s, b.contact_breaking_threshold);
    f32f(s, b.linear_damping);
    f32f(s, b.linear_sleeping_threshold);
    f32f(s, b.angular_sleeping_threshold);
    aff3(s, b.pose);
    byte(s, b.shape_kind);
    switch (b.shape_kind) {
      case 0:  // triangle mesh
        u32f(s, b.mesh_tri_count);
        mesh_tris[i] = b.mesh_tri_count;
        break;
      case 1:  // plane
        break;
      case 2:  // sphere
        vec3(s, b.sphere_implicit_dim);
        f32f(s, b.sphere_margin);
        break;
      case 3:  // compound of one box (a car)
        vec3(s, b.box_half_extents);
        f32f(s, b.box_margin);
        aff3(s, b.child_trans);
        aabb(s, b.compound_local_aabb);
        break;
      default:
        s.fail("unknown shape kind");
        return;
    }
  }
  for (int d = 0; d < NUM_DYN; ++d) {
    if (w.dyn_body[d] >= (uint32_t)NUM_BODIES) {
      s.fail("dynamic body index out of range");
      return;
    }
    if (S::READ) w.dyn_handle[d] = w.bodies[w.dyn_body[d]].broadphase_handle;
  }

  // MESH: per mesh its triangles, wide BVH nodes, leaves and triangle infos, packed one mesh after another
  tag(s, "MESH");
  if (!expect_count(s, s.count(NUM_MESHES), NUM_MESHES, "meshes")) return;
  size_t tf = 0, nf = 0, lf = 0;
  for (int mi = 0; mi < NUM_MESHES; ++mi) {
    MeshInfo& m = w.meshes[mi];
    u32f(s, m.body);
    aabb(s, m.aabb_ident);
    // a writer takes each array from where the mesh has it; a reader packs them in mesh order
    size_t nt = s.count(m.tri_count);
    size_t t0 = S::READ ? tf : m.tri_first;
    if (t0 + nt > (size_t)MAX_TRIS) {
      s.fail("triangles over the StaticData size");
      return;
    }
    if (S::READ) {
      m.tri_first = (uint32_t)tf;
      m.tri_count = (uint32_t)nt;
    }
    for (size_t k = t0; k < t0 + nt; ++k) {
      vec3(s, w.tris[k].p0);
      vec3(s, w.tris[k].p1);
      vec3(s, w.tris[k].p2);
      vec3(s, w.tris[k].normal);
    }
    aabb(s, m.root_aabb);
    size_t nn = s.count(m.node_count);
    size_t n0 = S::READ ? nf : m.node_first;
    if (n0 + nn > (size_t)MAX_NODES) {
      s.fail("BVH nodes over the StaticData size");
      return;
    }
    if (S::READ) {
      m.node_first = (uint32_t)nf;
      m.node_count = (uint32_t)nn;
    }
    for (size_t k = n0; k < n0 + nn; ++k) {
      WideNode& node = w.nodes[k];
      for (float* lanes : {node.min_x, node.min_y, node.min_z, node.max_x, node.max_y, node.max_z})
        for (int l = 0; l < 4; ++l) f32f(s, lanes[l]);
      for (auto& c : node.children) u32f(s, c);
      byte(s, node.child_count);
      u32f(s, node.valid_mask);
    }
    size_t nl = s.count(m.leaf_count);
    size_t l0 = S::READ ? lf : m.leaf_first;
    if (l0 + nl > (size_t)MAX_LEAVES) {
      s.fail("BVH leaves over the StaticData size");
      return;
    }
    if (S::READ) {
      m.leaf_first = (uint32_t)lf;
      m.leaf_count = (uint32_t)nl;
    }
    for (size_t k = l0; k < l0 + nl; ++k) {
      aabb(s, w.leaves[k].bounds);
      u32f(s, w.leaves[k].tri);
    }
    size_t ni = s.count(m.tri_count);
    if (ni != nt) {
      s.fail("one triangle info per triangle expected");
      return;
    }
    for (size_t k = t0; k < t0 + ni; ++k) {
      TriInfo& info = w.infos[k];
      byte(s, info.flags);
      f32f(s, info.edge_v0_v1_angle);
      f32f(s, info.edge_v1_v2_angle);
      f32f(s, info.edge_v2_v0_angle);
    }
    tf += nt;
    nf += nn;
    lf += nl;
  }
  if (S::READ && (tf != want.tris || nf != want.nodes || lf != want.leaves)) {
    s.fail("mesh totals differ from the header counts");
    return;
  }
  for (const MeshInfo& m : w.meshes) {
    if (m.body >= (uint32_t)NUM_BODIES) {
      s.fail("mesh body index out of range");
      return;
    }
    if (m.tri_count != mesh_tris[m.body]) {
      s.fail("mesh triangle count differs from its body record");
      return;
    }
  }

  // PLAN
  tag(s, "PLAN");
  if (!expect_count(s, s.count(NUM_PLANES), NUM_PLANES, "planes")) return;
  for (Plane& p : w.planes) {
    u32f(s, p.body);
    vec3(s, p.normal);
    flag(s, p.is_single_axis);
    u32f(s, p.single_axis_idx);
    flag(s, p.single_axis_backwards);
    aabb(s, p.aabb_ident);
    aabb(s, p.aabb);
  }

  // CELL: the static ray-cell lists of the occupied cells, ascending
  tag(s, "CELL");
  if (S::READ) {
    size_t m = s.count(0), first = 0;
    for (size_t c = 0; c < m; ++c) {
      uint32_t cell = 0, cnt = 0;
      u32f(s, cell);
      u32f(s, cnt);
      if (cell >= (uint32_t)NUM_GRID_CELLS || first + cnt > (size_t)MAX_CELL_ENTRIES) {
        s.fail("static ray cells over the StaticData size");
        return;
      }
      w.cells[cell] = CellRange{(uint32_t)first, cnt};
      for (size_t k = first; k < first + cnt; ++k) u32f(s, w.cell_handles[k]);
      first += cnt;
    }
    if (first != want.cell_entries) {
      s.fail("static ray-cell entries differ from the header count");
      return;
    }
  } else {
    size_t occupied = 0;
    for (const CellRange& c : w.cells) occupied += c.count != 0;
    s.count(occupied);
    for (uint32_t cell = 0; cell < (uint32_t)NUM_GRID_CELLS; ++cell) {
      CellRange c = w.cells[cell];
      if (!c.count) continue;
      if ((size_t)c.first + c.count > (size_t)MAX_CELL_ENTRIES) {
        s.fail("static ray cells over the StaticData size");
        return;
      }
      u32f(s, cell);
      u32f(s, c.count);
      for (uint32_t k = c.first; k < c.first + c.count; ++k) u32f(s, w.cell_handles[k]);
    }
  }

  // CARS
  tag(s, "CARS");
  if (!expect_count(s, s.count(NUM_CARS), NUM_CARS, "cars")) return;
  for (CarConst& c : w.cars) {
    u32f(s, c.car_idx);
    byte(s, c.team);
    u32f(s, c.body_idx);
    flag(s, c.three_wheels);
    vec3(s, c.hitbox_size);
    vec3(s, c.hitbox_offset);
    for (WheelConst& wc : c.wheels) {
      vec3(s, wc.connection_point);
      f32f(s, wc.suspension_rest_length);
      f32f(s, wc.radius);
      f32f(s, wc.suspension_force_scale);
      f32f(s, wc.ray_length);
    }
  }

  // PADS
  tag(s, "PADS");
  if (!expect_count(s, s.count(NUM_PADS), NUM_PADS, "pads")) return;
  f32f(s, w.max_pad_z);
  for (PadConst& p : w.pads) {
    vec3(s, p.pos);
    flag(s, p.is_big);
    f32f(s, p.box_radius);
    f32f(s, p.cyl_radius);
    f32f(s, p.max_cooldown);
    f32f(s, p.boost_amount);
    aabb(s, p.aabb);
  }

  // PROX: the sweep-and-prune proxies (index 0 the sentinel); a dynamic body's box and cells are state (zero here)
  tag(s, "PROX");
  if (!expect_count(s, s.count(NUM_HANDLES), NUM_HANDLES, "proxies")) return;
  for (ProxyConst& p : w.proxies) {
    u32f(s, p.client_body);
    byte(s, p.filter_group);
    byte(s, p.filter_mask);
    u32f(s, p.unique_id);
    aabb(s, p.aabb);
    u32f(s, p.cell_idx);
    for (auto& i : p.indices) u32f(s, i);
  }
}

// V3's table_sin (car/base.rs:101): the f64 sine of i * 2 pi / 16384, rounded to f32, as V3 computes it on the host.
void fill_table_sin(StaticData& w) {
  for (int i = 0; i < TABLE_SIN_LEN; ++i) w.table_sin[i] = (float)sin((double)i * 6.283185307179586476925286766559 / 16384.0);
}

uint64_t fnv1a64(const uint8_t* b, size_t n) {
  uint64_t h = 0xcbf29ce484222325ull;
  for (size_t i = 0; i < n; ++i) h = (h ^ b[i]) * 0x00000100000001b3ull;
  return h;
}

uint32_t rd32(const uint8_t* p) {
  uint32_t v;
  memcpy(&v, p, 4);
  return v;
}
uint64_t rd64(const uint8_t* p) {
  uint64_t v;
  memcpy(&v, p, 8);
  return v;
}
void put32(std::vector<uint8_t>& o, uint32_t v) { o.insert(o.end(), (uint8_t*)&v, (uint8_t*)&v + 4); }
void put64(std::vector<uint8_t>& o, uint64_t v) { o.insert(o.end(), (uint8_t*)&v, (uint8_t*)&v + 8); }

int format_error(const std::string& what) { return host::set_error(RSIM_ERR_FORMAT, what.c_str()); }

int arena_header(const uint8_t* b, size_t len, uint64_t* hash, uint32_t* counts) {
  if (len < RSIM_V3_ARENA_HEADER_BYTES || memcmp(b, ARENA_MAGIC, 8) != 0) return format_error("not an arena dump (magic)");
  uint32_t version = rd32(b + 8);
  if (version != RSIM_V3_ARENA_VERSION) return format_error("unsupported arena dump version " + std::to_string(version));
  if (hash) *hash = rd64(b + 12);
  if (counts)
    for (int i = 0; i < RSIM_V3_COUNTS; ++i) counts[i] = rd32(b + 20 + 4 * i);
  return RSIM_OK;
}

int static_header(const uint8_t* b, size_t len, uint64_t* hash, uint32_t* counts) {
  if (len < RSIM_V3_STATIC_HEADER_BYTES || memcmp(b, STATIC_MAGIC, 8) != 0) return format_error("not a static dump (magic)");
  uint32_t version = rd32(b + 8);
  if (version != RSIM_V3_STATIC_VERSION)
    return format_error("unsupported static dump version " + std::to_string(version) + " (need 3: re-vendor V3)");
  uint64_t h = rd64(b + 12), payload = rd64(b + 20);
  if (payload != len - RSIM_V3_STATIC_HEADER_BYTES) return format_error("static dump payload length mismatch");
  if (fnv1a64(b + RSIM_V3_STATIC_HEADER_BYTES, payload) != h) return format_error("static dump payload hash mismatch");
  if (hash) *hash = h;
  if (counts)
    for (int i = 0; i < RSIM_V3_COUNTS; ++i) counts[i] = rd32(b + 28 + 4 * i);
  return RSIM_OK;
}

// Hands `bytes` to the caller's buffer, or reports the size it needs.
int deliver(const std::vector<uint8_t>& bytes, void* out, uint64_t cap, uint64_t* len_out, const char* what) {
  if (len_out) *len_out = bytes.size();
  if (!out || cap < bytes.size()) return host::set_error(RSIM_ERR_SHORT, what);
  memcpy(out, bytes.data(), bytes.size());
  return RSIM_OK;
}

}  // namespace
}  // namespace wire
}  // namespace rs

using namespace rs;
using namespace rs::wire;

extern "C" {

uint64_t rsim_v3_template_bytes(void) { return sizeof(rsim_v3_template); }

uint64_t rsim_fnv1a64(const void* bytes, uint64_t len) {
  if (!bytes) return fnv1a64(nullptr, 0);
  return fnv1a64((const uint8_t*)bytes, len);
}

int rsim_v3_static_header(const void* bytes, uint64_t len, uint64_t* payload_hash, uint32_t counts[RSIM_V3_COUNTS]) {
  if (!bytes) return host::set_error(RSIM_ERR_INVALID, "rsim_v3_static_header: null bytes");
  return static_header((const uint8_t*)bytes, len, payload_hash, counts);
}

int rsim_v3_parse_static(const void* bytes, uint64_t len, void* static_out) {
  if (!bytes || !static_out) return host::set_error(RSIM_ERR_INVALID, "rsim_v3_parse_static: null argument");
  const uint8_t* b = (const uint8_t*)bytes;
  uint32_t c[RSIM_V3_COUNTS];
  int rc = static_header(b, len, nullptr, c);
  if (rc) return rc;
  // exact: the arena's fixed body, mesh, plane, car, pad and grid counts; at most: the mesh and ray-cell arrays
  const struct {
    int i;
    uint32_t size;
    bool exact;
  } want[RSIM_V3_COUNTS] = {{0, NUM_BODIES, true}, {1, NUM_DYN, true},       {2, NUM_MESHES, true},
                            {3, MAX_TRIS, false},  {4, MAX_NODES, false},    {5, MAX_LEAVES, false},
                            {6, MAX_TRIS, false},  {7, NUM_PLANES, true},    {8, NUM_CARS, true},
                            {9, NUM_PADS, true},   {10, NUM_GRID_CELLS, true}, {11, MAX_CELL_ENTRIES, false}};
  for (const auto& x : want) {
    uint32_t got = c[x.i];
    if (x.exact ? got != x.size : got > x.size)
      return format_error("static dump count " + std::to_string(x.i) + " is " + std::to_string(got) + "; StaticData is sized " +
                          (x.exact ? "for " : "for at most ") + std::to_string(x.size));
  }
  if (c[6] != c[3]) return format_error("static dump count 6 (triangle infos) differs from count 3 (triangles)");
  std::unique_ptr<StaticData> w(new StaticData());
  Reader r(b + RSIM_V3_STATIC_HEADER_BYTES, len - RSIM_V3_STATIC_HEADER_BYTES);
  static_io(r, *w, Totals{c[3], c[4], c[5], c[11]});
  if (!r.err.empty()) return format_error(r.err);
  if (r.remaining()) return format_error(std::to_string(r.remaining()) + " trailing bytes in the static dump");
  fill_table_sin(*w);
  memcpy(static_out, w.get(), sizeof(StaticData));
  return RSIM_OK;
}

int rsim_v3_write_static(const void* static_data, void* out, uint64_t cap, uint64_t* len_out) {
  if (!static_data) return host::set_error(RSIM_ERR_INVALID, "rsim_v3_write_static: null static data");
  std::unique_ptr<StaticData> w(new StaticData());
  memcpy(w.get(), static_data, sizeof(StaticData));
  Writer wr;
  static_io(wr, *w, Totals{0, 0, 0, 0});
  if (!wr.err.empty()) return format_error(wr.err);
  uint32_t tris = 0, nodes = 0, leaves = 0, entries = 0;
  for (const MeshInfo& m : w->meshes) {
    tris += m.tri_count;
    nodes += m.node_count;
    leaves += m.leaf_count;
  }
  for (const CellRange& c : w->cells) entries += c.count;
  const uint32_t counts[RSIM_V3_COUNTS] = {NUM_BODIES, NUM_DYN, NUM_MESHES, tris,     nodes,          leaves,
                                           tris,       NUM_PLANES, NUM_CARS, NUM_PADS, NUM_GRID_CELLS, entries};
  std::vector<uint8_t> o(STATIC_MAGIC, STATIC_MAGIC + 8);
  put32(o, RSIM_V3_STATIC_VERSION);
  put64(o, fnv1a64(wr.out.data(), wr.out.size()));
  put64(o, wr.out.size());
  for (uint32_t c : counts) put32(o, c);
  o.insert(o.end(), wr.out.begin(), wr.out.end());
  return deliver(o, out, cap, len_out, "rsim_v3_write_static: output buffer too small");
}

int rsim_v3_arena_header(const void* bytes, uint64_t len, uint64_t* config_hash, uint32_t counts[RSIM_V3_COUNTS]) {
  if (!bytes) return host::set_error(RSIM_ERR_INVALID, "rsim_v3_arena_header: null bytes");
  return arena_header((const uint8_t*)bytes, len, config_hash, counts);
}

int rsim_v3_parse_arena(const void* bytes, uint64_t len, const void* static_data, void* state_out, rsim_v3_template* tpl_out) {
  if (!bytes || !static_data || !state_out) return host::set_error(RSIM_ERR_INVALID, "rsim_v3_parse_arena: null argument");
  const uint8_t* b = (const uint8_t*)bytes;
  uint64_t hash = 0;
  int rc = arena_header(b, len, &hash, nullptr);
  if (rc) return rc;
  std::unique_ptr<ArenaState> a(new ArenaState());
  rsim_v3_template t;
  memset(&t, 0, sizeof t);
  t.config_hash = hash;
  Reader r(b + RSIM_V3_ARENA_HEADER_BYTES, len - RSIM_V3_ARENA_HEADER_BYTES);
  arena_io(r, *a, *(const StaticData*)static_data, t);
  island_io(r);
  if (!r.err.empty()) return format_error(r.err);
  if (r.remaining()) return format_error(std::to_string(r.remaining()) + " trailing bytes");
  memcpy(state_out, a.get(), sizeof(ArenaState));
  if (tpl_out) *tpl_out = t;
  return RSIM_OK;
}

int rsim_v3_write_arena(const void* state, const void* static_data, const rsim_v3_template* tpl, void* out, uint64_t cap,
                        uint64_t* len_out) {
  if (!state || !static_data || !tpl) return host::set_error(RSIM_ERR_INVALID, "rsim_v3_write_arena: null argument");
  std::unique_ptr<ArenaState> a(new ArenaState());
  memcpy(a.get(), state, sizeof(ArenaState));
  const StaticData& w = *(const StaticData*)static_data;
  // the caps a written state must keep (a parsed or stepped one always does)
  if (a->n_manifolds > (uint32_t)MAX_MANIFOLDS || a->n_active > (uint32_t)MAX_ACTIVE || a->n_pairs > (uint32_t)MAX_PAIRS ||
      a->n_dyn_cells > (uint32_t)MAX_DYN_CELLS || a->n_records > (uint32_t)MAX_RECORDS)
    return format_error("a list of the state is over its cap");
  for (const PadTouching& p : a->pad_touching)
    if (p.n > (uint32_t)NUM_CARS) return format_error("a pad touched by more cars than the arena has");
  rsim_v3_template t = *tpl;
  Writer wr;
  arena_io(wr, *a, w, t);
  island_io(wr);
  if (!wr.err.empty()) return format_error(wr.err);
  std::vector<ContactRecord> recs;
  unfold(*a, w, recs);
  uint32_t points = 0;
  for (uint32_t i = 0; i < a->n_manifolds; ++i) points += a->manifolds[i].n_points;
  const uint32_t stride = a->n_manifolds ? 32u : t.manifold_stride;
  const uint32_t counts[RSIM_V3_COUNTS] = {NUM_BODIES, a->n_manifolds, a->n_active, points,   stride * stride,
                                           NUM_HANDLES, a->n_pairs,     a->n_dyn_cells, NUM_CARS, NUM_PADS,
                                           (uint32_t)recs.size(), 0};
  std::vector<uint8_t> o(ARENA_MAGIC, ARENA_MAGIC + 8);
  put32(o, RSIM_V3_ARENA_VERSION);
  put64(o, t.config_hash);
  for (uint32_t c : counts) put32(o, c);
  o.insert(o.end(), wr.out.begin(), wr.out.end());
  return deliver(o, out, cap, len_out, "rsim_v3_write_arena: output buffer too small");
}

int rsim_v3_fold_records(void* state, const void* static_data, const void* records, uint32_t n) {
  if (!state || !static_data || (n && !records)) return host::set_error(RSIM_ERR_INVALID, "rsim_v3_fold_records: null argument");
  std::string e = fold(*(ArenaState*)state, *(const StaticData*)static_data, (const ContactRecord*)records, n);
  if (!e.empty()) return format_error(e);
  return RSIM_OK;
}

int rsim_v3_unfold_records(const void* state, const void* static_data, void* records_out, uint32_t cap, uint32_t* n_out) {
  if (!state || !static_data) return host::set_error(RSIM_ERR_INVALID, "rsim_v3_unfold_records: null argument");
  std::vector<ContactRecord> recs;
  std::string e = unfold(*(const ArenaState*)state, *(const StaticData*)static_data, recs);
  if (!e.empty()) return format_error(e);
  if (n_out) *n_out = (uint32_t)recs.size();
  if (!records_out || cap < recs.size()) return host::set_error(RSIM_ERR_SHORT, "rsim_v3_unfold_records: output too small");
  if (!recs.empty()) memcpy(records_out, recs.data(), recs.size() * sizeof(ContactRecord));
  return RSIM_OK;
}

}  // extern "C"
