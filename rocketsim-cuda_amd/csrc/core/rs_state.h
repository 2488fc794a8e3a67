// The arena state and the static world, word for word in the reference's format (cuda/sim/arena_state.cuh:38-342). The
// layout is a data format shared with the reference's CPU build (the golden vectors), so the field order and sizes are
// fixed; the offsets below are asserted against the values the reference's header gives under g++.
//
// Rules of the format: every field is a 4-byte word or a struct of them (no padding); bools are u32 0/1; an optional is
// {u32 some; T v} with v zero when absent; a u64 is {lo, hi}; a list is a count plus a capped array whose unused slots are
// zero, with an overflow counter bumped instead of writing past the cap.
#pragma once
#include "rs_vec.h"

namespace rs {

constexpr int NUM_BODIES = 23;  // 16 arena meshes, 4 planes, the ball (20), car 0 (21), car 1 (22)
constexpr int NUM_DYN = 3;      // dynamic bodies, in state order: ball, car 0, car 1
constexpr int NUM_CARS = 2;
constexpr int NUM_WHEELS = 4;
constexpr int NUM_PADS = 34;
constexpr int NUM_HANDLES = 24;  // sweep-and-prune handles: sentinel + 23 bodies
constexpr int NUM_EDGES = 2 * NUM_HANDLES;
constexpr int NUM_MESHES = 16;
constexpr int NUM_PLANES = 4;

constexpr int MAX_MANIFOLDS = 16;
constexpr int MAX_POINTS = 4;
constexpr int MAX_ACTIVE = 16;
constexpr int MAX_RECORDS = 16;
constexpr int MAX_PAIRS = 16;
constexpr int MAX_DYN_CELLS = 64;

enum Overflow { OVF_MANIFOLDS, OVF_ACTIVE, OVF_RECORDS, OVF_PAIRS, OVF_DYN_CELLS, OVF_PAD_TOUCHING, OVF_COUNT };

// ArenaState.events bits (arena_state.cuh:65-75); per-car events use bit + car index.
enum : uint32_t {
  EV_CAR_HIT_BALL = 0, EV_BUMP = 2, EV_DEMO = 4, EV_PICKUP = 6, EV_LANDED = 8, EV_CAR_HIT_WORLD = 10, EV_BALL_HIT_WORLD = 12,
  EV_BALL_WALL = 14,
};

struct U64w { uint32_t lo, hi; };
template <class T> struct Opt { uint32_t some; T v; };
RS_FN uint64_t u64(U64w x) { return ((uint64_t)x.hi << 32) | x.lo; }
RS_FN U64w u64w(uint64_t v) { return U64w{(uint32_t)v, (uint32_t)(v >> 32)}; }

struct Controls { float throttle, steer, pitch, yaw, roll; uint32_t jump, boost, handbrake; };
struct HitInfo { V3 hit_point, hit_normal; float hit_fraction; uint32_t user_info; };
struct RaycastInfo {
  V3 contact_normal, contact_point;
  uint32_t ground_body;
  float suspension_length;
  V3 impulse;
  Opt<V3> ground_stick;
  uint32_t in_contact_with_world;
  float clipped_inv_contact_dot_suspension;
  float suspension_relative_vel;
};
struct Wheel {
  Opt<RaycastInfo> ray;
  V3 hard_point, axle_dir;
  float engine_force, brake, steer_angle;
  V3 vel_at_contact_point;
  float lat_friction, long_friction;
  float extra_pushback;
};
struct Car {
  float dodge_deadzone;
  V3 vel_impulse_cache;
  uint32_t sticky_gate_prev;
  V3 pos;
  M3 rot;
  V3 vel, ang_vel;
  Controls controls;
  uint32_t has_jumped, has_double_jumped, has_flipped;
  V3 flip_rel_torque;
  uint32_t jump_ticks;
  float flip_time;
  uint32_t is_flipping, is_jumping, jump_hold_broken;
  float air_time, air_time_since_jump, boost, handbrake_val;
  uint32_t is_supersonic;
  float supersonic_grace_timer, time_since_boosted, boosting_time;
  uint32_t is_boosting, is_auto_flipping;
  float auto_flip_timer, auto_flip_torque_scale, bump_cooldown_timer;
  Opt<uint32_t> bump_cooldown_victim;
  Opt<U64w> last_extra_hit_tick;
  uint32_t is_demoed;
  float demo_respawn_timer;
  uint32_t is_on_ground;
  Opt<HitInfo> wheels_with_contact[NUM_WHEELS];
  Opt<V3> world_contact_normal;
  Controls prev_controls;
  Wheel wheels[NUM_WHEELS];
};
struct PadTouching { uint32_t n; uint32_t car_idx[NUM_CARS]; };
struct Ball { V3 pos; M3 rot; V3 vel, ang_vel; U64w tick_count_since_kickoff; };
struct Body {
  uint32_t collision_flags;
  uint32_t activation;  // 0 active, 1 sleeping, 2 disabled, 3 wants deactivation
  float deactivation_time;
  Aff world;
  Quat world_quat;
  Aff interp_world;
  V3 lin_vel, ang_vel;
  M3 inv_inertia_world;
  V3 accum_lin_vel, accum_ang_vel;
  float cached_linear_damping, cached_damping_time_step, cached_linear_damping_value;
};
struct ManifoldPoint {
  V3 local_a, local_b, world_on_b, world_on_a, normal_on_b;
  float distance;
  float friction, restitution;
  uint32_t is_special;
};
struct Manifold { uint32_t body0, body1; uint32_t n_points; ManifoldPoint pts[MAX_POINTS]; };
struct RecordPoint { V3 world_on_b, world_on_a, normal_on_b; };
struct ContactRecord { uint32_t is_swap, body_a, body_b; RecordPoint pt; };
struct DynProxy { Aabb aabb; uint32_t cell_idx; uint32_t indices[3]; };
struct HandleEdges { uint16_t min[3], max[3]; };
struct Edge { uint16_t pos, handle; };
struct DynCell { uint32_t cell, handle; };
struct Pair { uint32_t proxy0, proxy1; };

struct ArenaState {
  U64w tick_count;
  U64w rng;
  Ball ball;
  Car cars[NUM_CARS];
  Opt<U64w> pad_gave_boost[NUM_PADS];
  PadTouching pad_touching[NUM_PADS];
  Body bodies[NUM_DYN];
  uint32_t n_manifolds;
  Manifold manifolds[MAX_MANIFOLDS];
  uint32_t n_dyn_cells;
  DynCell dyn_cells[MAX_DYN_CELLS];
  DynProxy dyn_proxy[NUM_DYN];
  HandleEdges handle_edges[NUM_HANDLES];
  Edge edges[3][NUM_EDGES];
  uint32_t n_pairs;
  Pair pairs[MAX_PAIRS];
  uint32_t overflow[OVF_COUNT];
  uint32_t events;
  // mid-tick part: zero at every tick boundary
  uint32_t n_active;
  uint32_t active[MAX_ACTIVE];
  uint32_t fold_events;
  V3 fold_car_normal[NUM_CARS];
  uint32_t n_records;
  ContactRecord records[MAX_RECORDS];
};

constexpr int ARENA_WORDS = (int)(sizeof(ArenaState) / 4);
constexpr int TICK_WORDS = (int)(offsetof(ArenaState, n_active) / 4);

// ---- the static world (arena_state.cuh:236-339) ----------------------------------------------------------------------
constexpr int MAX_TRIS = 8020;
constexpr int MAX_NODES = 3754;
constexpr int MAX_LEAVES = 8020;
constexpr int TABLE_SIN_LEN = 16384;
constexpr int NUM_GRID_CELLS = 242;
constexpr int MAX_CELL_ENTRIES = 1210;
constexpr int NUM_MUT = 20;

struct MeshInfo { uint32_t body, tri_first, tri_count, node_first, node_count, leaf_first, leaf_count; Aabb aabb_ident, root_aabb; };
struct Tri { V3 p0, p1, p2, normal; };
struct WideNode {
  float min_x[4], min_y[4], min_z[4], max_x[4], max_y[4], max_z[4];
  uint32_t children[4];  // bit 31 set: a leaf index; else a node index (both local to the mesh)
  uint32_t child_count;
  uint32_t valid_mask;
};
struct Leaf { Aabb bounds; uint32_t tri; };
struct TriInfo { uint32_t flags; float edge_v0_v1_angle, edge_v1_v2_angle, edge_v2_v0_angle; };
struct Plane {
  uint32_t body;
  V3 normal;
  uint32_t is_single_axis, single_axis_idx, single_axis_backwards;
  Aabb aabb_ident, aabb;
};
struct BodyConst {
  uint32_t world_array_idx, user_idx, user_pointer, broadphase_handle, filter_group, filter_mask;
  float mass, inv_mass;
  V3 inv_inertia_local, inv_mass_splat;
  float friction, restitution, contact_processing_threshold, contact_breaking_threshold;
  float linear_damping, linear_sleeping_threshold, angular_sleeping_threshold;
  Aff pose;
  uint32_t shape_kind;  // 0 triangle mesh, 1 plane, 2 sphere, 3 compound (a car: one box child)
  V3 sphere_implicit_dim;
  float sphere_margin;  // the sphere's radius
  V3 box_half_extents;
  float box_margin;
  Aff child_trans;
  Aabb compound_local_aabb;
  uint32_t mesh_tri_count;
};
struct WheelConst { V3 connection_point; float suspension_rest_length, radius, suspension_force_scale, ray_length; };
struct CarConst {
  uint32_t car_idx, team, body_idx, three_wheels;
  V3 hitbox_size, hitbox_offset;
  WheelConst wheels[NUM_WHEELS];
};
struct PadConst { V3 pos; uint32_t is_big; float box_radius, cyl_radius, max_cooldown, boost_amount; Aabb aabb; };
struct ProxyConst { uint32_t client_body, filter_group, filter_mask, unique_id; Aabb aabb; uint32_t cell_idx, indices[3]; };
struct CellRange { uint32_t first, count; };

struct StaticData {
  uint32_t game_mode, no_ball_rot, ball_body;
  V3 gravity_uu;
  float mut_f[NUM_MUT];
  uint32_t unlimited_flips, unlimited_double_jumps, recharge_boost_enabled;
  float recharge_boost_per_second, recharge_boost_delay;
  uint32_t demo_mode, enable_team_demos;
  float goal_base_threshold_y;
  uint32_t boost_pads_enabled;
  V3 gravity_bt;
  uint32_t dyn_body[NUM_DYN];
  V3 grid_min, grid_max;
  float grid_cell_size, grid_cell_size_sq;
  uint32_t grid_num_cells[3];
  V3 sap_world_min, sap_quantize;
  uint32_t sap_handle_mask, sap_sentinel, sap_max_handles;
  float sap_world_half_extent;
  float max_pad_z;
  uint32_t dyn_handle[NUM_DYN];
  BodyConst bodies[NUM_BODIES];
  Plane planes[NUM_PLANES];
  MeshInfo meshes[NUM_MESHES];
  CarConst cars[NUM_CARS];
  PadConst pads[NUM_PADS];
  ProxyConst proxies[NUM_HANDLES];
  CellRange cells[NUM_GRID_CELLS];
  uint32_t cell_handles[MAX_CELL_ENTRIES];
  TriInfo infos[MAX_TRIS];
  Tri tris[MAX_TRIS];
  WideNode nodes[MAX_NODES];
  Leaf leaves[MAX_LEAVES];
  float table_sin[TABLE_SIN_LEN];
};
constexpr int STATIC_WORDS = (int)(sizeof(StaticData) / 4);

// The format's sizes and landmark offsets, as the reference's header yields them.
static_assert(sizeof(ArenaState) == 11136, "ArenaState is 2,784 words");
static_assert(TICK_WORDS == 2567, "2,567 words persist across a tick boundary");
static_assert(offsetof(ArenaState, cars) == 96 && offsetof(ArenaState, bodies) == 2888, "ArenaState layout");
static_assert(offsetof(ArenaState, manifolds) == 3552 && offsetof(ArenaState, edges) == 9532, "ArenaState layout");
static_assert(offsetof(ArenaState, events) == 10264, "ArenaState layout");
static_assert(sizeof(StaticData) == 1271580, "StaticData is 1,271,580 bytes");
static_assert(offsetof(StaticData, tris) == 146044 && offsetof(StaticData, nodes) == 531004, "StaticData layout");
static_assert(offsetof(StaticData, leaves) == 981484 && offsetof(StaticData, table_sin) == 1206044, "StaticData layout");

// Mutator floats of StaticData.mut_f, in order (consts.cuh:42-45).
enum MutF {
  MF_CAR_MASS, MF_BALL_MASS, MF_BALL_MAX_SPEED, MF_BALL_DRAG, MF_JUMP_ACCEL, MF_JUMP_IMMEDIATE_FORCE, MF_BOOST_ACCEL_GROUND,
  MF_BOOST_ACCEL_AIR, MF_BOOST_USED_PER_SECOND, MF_RESPAWN_DELAY, MF_BUMP_COOLDOWN_TIME, MF_CAR_MAX_BOOST_AMOUNT,
  MF_CAR_SPAWN_BOOST_AMOUNT, MF_PAD_AMOUNT_SMALL, MF_PAD_AMOUNT_BIG, MF_PAD_COOLDOWN_BIG, MF_PAD_COOLDOWN_SMALL,
  MF_BALL_HIT_EXTRA_FORCE_SCALE, MF_BUMP_FORCE_SCALE, MF_BALL_RADIUS
};

// Result word of one decision (arena_step.cuh:16-19).
constexpr uint32_t TICKS_PER_DECISION = 8;
constexpr uint32_t SIM8_SCORED_MASK = 0xFFu;
constexpr uint32_t SIM8_UNIMPLEMENTED = 0x80000000u;
constexpr uint32_t SIM8_OVERFLOW = 0x40000000u;

}  // namespace rs
