"""Named access to the ArenaState and StaticData words: numpy structured dtypes with the field names and order of the
reference's Rust mirror (crates/rocketsim-cuda/src/layout.rs:178-461), so a state array can be read and edited by name.

    s = rsim.sample_arenas(world, rsim.SCEN_KICKOFF, 1, 8)      # uint32[8, ARENA_WORDS]
    a = layout.arena_view(s)                                    # zero-copy record view, shape (8,)
    a["ball"]["pos"]["z"], a["cars"][:, 0]["boost"]
    layout.arena_leaves()[w]                                    # (field path, "f" | "i") of word w (layout.rs:490-496)

Every field is a 4-byte word or a struct of them (u16 pairs only in the sweep-and-prune edges), bools are u32 0/1, an
optional is {some, v}, a u64 is {lo, hi}.
"""
import numpy as np

NUM_BODIES, NUM_DYN, NUM_CARS, NUM_WHEELS, NUM_PADS = 23, 3, 2, 4, 34
NUM_HANDLES, NUM_MESHES, NUM_PLANES = 24, 16, 4
NUM_EDGES = 2 * NUM_HANDLES
MAX_MANIFOLDS, MAX_POINTS, MAX_ACTIVE, MAX_RECORDS, MAX_PAIRS, MAX_DYN_CELLS = 16, 4, 16, 16, 16, 64
OVF_COUNT = 6
MAX_TRIS, MAX_NODES, MAX_LEAVES = 8020, 3754, 8020
TABLE_SIN_LEN, NUM_GRID_CELLS, MAX_CELL_ENTRIES, NUM_MUT = 16384, 242, 1210, 20

f4, u4, u2 = np.dtype("<f4"), np.dtype("<u4"), np.dtype("<u2")


def _s(*fields):
    return np.dtype(list(fields))


def _opt(t):
    return _s(("some", u4), ("v", t))


vec3 = _s(("x", f4), ("y", f4), ("z", f4))
quat = _s(("x", f4), ("y", f4), ("z", f4), ("w", f4))
mat3 = _s(("x_axis", vec3), ("y_axis", vec3), ("z_axis", vec3))
affine3 = _s(("matrix3", mat3), ("translation", vec3))
aabb = _s(("min", vec3), ("max", vec3))
u64 = _s(("lo", u4), ("hi", u4))

# ---- ArenaState (layout.rs:178-315) ------------------------------------------------------------------------------------------
controls = _s(("throttle", f4), ("steer", f4), ("pitch", f4), ("yaw", f4), ("roll", f4), ("jump", u4), ("boost", u4),
              ("handbrake", u4))
hit_info = _s(("hit_point", vec3), ("hit_normal", vec3), ("hit_fraction", f4), ("user_info", u4))
raycast_info = _s(("contact_normal", vec3), ("contact_point", vec3), ("ground_body_idx", u4), ("suspension_length", f4),
                  ("impulse", vec3), ("ground_stick", _opt(vec3)), ("is_in_contact_with_world", u4),
                  ("clipped_inv_contact_dot_suspension", f4), ("suspension_relative_vel", f4))
wheel = _s(("raycast_info", _opt(raycast_info)), ("hard_point", vec3), ("axle_dir", vec3), ("engine_force", f4),
           ("brake", f4), ("steer_angle", f4), ("vel_at_contact_point", vec3), ("lat_friction", f4), ("long_friction", f4),
           ("extra_pushback", f4))
car = _s(("dodge_deadzone", f4), ("vel_impulse_cache", vec3), ("sticky_gate_prev", u4), ("pos", vec3), ("rot_mat", mat3),
         ("vel", vec3), ("ang_vel", vec3), ("controls", controls), ("has_jumped", u4), ("has_double_jumped", u4),
         ("has_flipped", u4), ("flip_rel_torque", vec3), ("jump_ticks", u4), ("flip_time", f4), ("is_flipping", u4),
         ("is_jumping", u4), ("jump_hold_broken", u4), ("air_time", f4), ("air_time_since_jump", f4), ("boost", f4),
         ("handbrake_val", f4), ("is_supersonic", u4), ("supersonic_grace_timer", f4), ("time_since_boosted", f4),
         ("boosting_time", f4), ("is_boosting", u4), ("is_auto_flipping", u4), ("auto_flip_timer", f4),
         ("auto_flip_torque_scale", f4), ("bump_cooldown_timer", f4), ("bump_cooldown_victim", _opt(u4)),
         ("last_extra_hit_tick", _opt(u64)), ("is_demoed", u4), ("demo_respawn_timer", f4), ("is_on_ground", u4),
         ("wheels_with_contact", _opt(hit_info), (NUM_WHEELS,)), ("world_contact_normal", _opt(vec3)),
         ("prev_controls", controls), ("wheels", wheel, (NUM_WHEELS,)))
ball_dyn = _s(("pos", vec3), ("rot_mat", mat3), ("vel", vec3), ("ang_vel", vec3), ("tick_count_since_kickoff", u64))
body = _s(("collision_flags", u4), ("activation", u4), ("deactivation_time", f4), ("world_trans", affine3),
          ("world_quat", quat), ("interp_world_trans", affine3), ("lin_vel", vec3), ("ang_vel", vec3),
          ("inv_inertia_tensor_world", mat3), ("accum_lin_vel", vec3), ("accum_ang_vel", vec3),
          ("cached_linear_damping", f4), ("cached_damping_time_step", f4), ("cached_linear_damping_value", f4))
manifold_point = _s(("local_point_a", vec3), ("local_point_b", vec3), ("pos_world_on_b", vec3), ("pos_world_on_a", vec3),
                    ("normal_world_on_b", vec3), ("distance_1", f4), ("combined_friction", f4),
                    ("combined_restitution", f4), ("is_special", u4))
manifold = _s(("body0_idx", u4), ("body1_idx", u4), ("point_count", u4), ("points", manifold_point, (MAX_POINTS,)))
record_point = _s(("pos_world_on_b", vec3), ("pos_world_on_a", vec3), ("normal_world_on_b", vec3))
contact_record = _s(("is_swap", u4), ("rb_idx_a", u4), ("rb_idx_b", u4), ("point", record_point))
dyn_proxy = _s(("aabb", aabb), ("cell_idx", u4), ("indices", u4, (3,)))
dyn_cell = _s(("cell", u4), ("handle", u4))
pair = _s(("proxy0", u4), ("proxy1", u4))
pad_touching = _s(("n", u4), ("car_idx", u4, (NUM_CARS,)))
handle_edges = _s(("min", u2, (3,)), ("max", u2, (3,)))
edge = _s(("pos", u2), ("handle", u2))

ARENA = _s(
    ("tick_count", u64), ("rng", u64), ("ball", ball_dyn), ("cars", car, (NUM_CARS,)),
    ("pad_gave_boost", _opt(u64), (NUM_PADS,)), ("pad_touching", pad_touching, (NUM_PADS,)), ("bodies", body, (NUM_DYN,)),
    ("n_manifolds", u4), ("manifolds", manifold, (MAX_MANIFOLDS,)), ("n_dyn_cells", u4),
    ("dyn_cells", dyn_cell, (MAX_DYN_CELLS,)), ("dyn_proxy", dyn_proxy, (NUM_DYN,)),
    ("handle_edges", handle_edges, (NUM_HANDLES,)), ("edges", edge, (3, NUM_EDGES)), ("n_pairs", u4),
    ("pairs", pair, (MAX_PAIRS,)), ("overflow", u4, (OVF_COUNT,)), ("events", u4),
    # the mid-tick part: zero at every tick boundary
    ("n_active", u4), ("active", u4, (MAX_ACTIVE,)), ("fold_events", u4), ("fold_car_normal", vec3, (NUM_CARS,)),
    ("n_records", u4), ("records", contact_record, (MAX_RECORDS,)))

# ---- StaticData (layout.rs:318-461) -----------------------------------------------------------------------------------------
mesh_info = _s(("body", u4), ("tri_first", u4), ("tri_count", u4), ("node_first", u4), ("node_count", u4),
               ("leaf_first", u4), ("leaf_count", u4), ("aabb_ident_cache", aabb), ("root_aabb", aabb))
tri = _s(("p0", vec3), ("p1", vec3), ("p2", vec3), ("normal", vec3))
wide_node = _s(("min_x", f4, (4,)), ("min_y", f4, (4,)), ("min_z", f4, (4,)), ("max_x", f4, (4,)), ("max_y", f4, (4,)),
               ("max_z", f4, (4,)), ("children", u4, (4,)), ("child_count", u4), ("valid_mask", u4))
leaf = _s(("bounds", aabb), ("leaf_idx", u4))
tri_info = _s(("flags", u4), ("edge_v0_v1_angle", f4), ("edge_v1_v2_angle", f4), ("edge_v2_v0_angle", f4))
plane = _s(("body", u4), ("plane_normal", vec3), ("is_single_axis", u4), ("single_axis_idx", u4),
           ("single_axis_backwards", u4), ("aabb_ident_cache", aabb), ("aabb_cache", aabb))
body_const = _s(("world_array_idx", u4), ("user_idx", u4), ("user_pointer", u4), ("broadphase_handle", u4),
                ("filter_group", u4), ("filter_mask", u4), ("mass", f4), ("inv_mass", f4), ("inv_inertia_local", vec3),
                ("inv_mass_splat", vec3), ("friction", f4), ("restitution", f4), ("contact_processing_threshold", f4),
                ("contact_breaking_threshold", f4), ("linear_damping", f4), ("linear_sleeping_threshold", f4),
                ("angular_sleeping_threshold", f4), ("pose", affine3), ("shape_kind", u4), ("sphere_implicit_dim", vec3),
                ("sphere_margin", f4), ("box_half_extents", vec3), ("box_margin", f4), ("child_trans", affine3),
                ("compound_local_aabb", aabb), ("mesh_tri_count", u4))
wheel_const = _s(("chassis_connection_point_cs", vec3), ("suspension_rest_length_1", f4), ("wheels_radius", f4),
                 ("suspension_force_scale", f4), ("real_ray_length", f4))
car_const = _s(("car_idx", u4), ("team", u4), ("rigid_body_idx", u4), ("three_wheels", u4), ("hitbox_size", vec3),
               ("hitbox_pos_offset", vec3), ("wheels", wheel_const, (NUM_WHEELS,)))
pad_const = _s(("pos", vec3), ("is_big", u4), ("box_radius", f4), ("cyl_radius", f4), ("max_cooldown", f4),
               ("boost_amount", f4), ("aabb", aabb))
proxy_const = _s(("client_obj_idx", u4), ("filter_group", u4), ("filter_mask", u4), ("unique_id", u4), ("aabb", aabb),
                 ("cell_idx", u4), ("indices", u4, (3,)))
cell_range = _s(("first", u4), ("count", u4))

STATIC = _s(
    ("game_mode", u4), ("no_ball_rot", u4), ("ball_body", u4), ("gravity_uu", vec3), ("mut_f", f4, (NUM_MUT,)),
    ("unlimited_flips", u4), ("unlimited_double_jumps", u4), ("recharge_boost_enabled", u4),
    ("recharge_boost_per_second", f4), ("recharge_boost_delay", f4), ("demo_mode", u4), ("enable_team_demos", u4),
    ("goal_base_threshold_y", f4), ("boost_pads_enabled", u4), ("gravity_bt", vec3), ("dyn_body", u4, (NUM_DYN,)),
    ("grid_min_pos", vec3), ("grid_max_pos", vec3), ("grid_cell_size", f4), ("grid_cell_size_sq", f4),
    ("grid_num_cells", u4, (3,)), ("sap_world_aabb_min", vec3), ("sap_quantize", vec3), ("sap_handle_mask", u4),
    ("sap_sentinel", u4), ("sap_max_handles", u4), ("sap_world_half_extent", f4), ("max_pad_z", f4),
    ("dyn_handle", u4, (NUM_DYN,)), ("bodies", body_const, (NUM_BODIES,)), ("planes", plane, (NUM_PLANES,)),
    ("meshes", mesh_info, (NUM_MESHES,)), ("cars", car_const, (NUM_CARS,)), ("pads", pad_const, (NUM_PADS,)),
    ("proxies", proxy_const, (NUM_HANDLES,)), ("cells", cell_range, (NUM_GRID_CELLS,)),
    ("cell_handles", u4, (MAX_CELL_ENTRIES,)), ("infos", tri_info, (MAX_TRIS,)), ("tris", tri, (MAX_TRIS,)),
    ("nodes", wide_node, (MAX_NODES,)), ("leaves", leaf, (MAX_LEAVES,)), ("table_sin", f4, (TABLE_SIN_LEN,)))

ARENA_WORDS = ARENA.itemsize // 4
TICK_WORDS = ARENA.fields["n_active"][1] // 4
STATIC_WORDS = STATIC.itemsize // 4
assert ARENA.itemsize == 11136 and STATIC.itemsize == 1271580 and TICK_WORDS == 2567


def _view(words, dt, n_words, what):
    a = np.asarray(words)
    if a.dtype != np.uint32 or not a.flags.c_contiguous or a.size % n_words:
        raise ValueError("%s: a C-contiguous uint32 array of a multiple of %d words expected" % (what, n_words))
    return a.reshape(-1, n_words).view(dt).reshape(a.shape[:-1] if a.ndim > 1 else ())


def arena_view(states):
    """Record view (no copy) of uint32[..., ARENA_WORDS] states: writes through it change the words."""
    return _view(states, ARENA, ARENA_WORDS, "arena_view")


def static_view(world):
    """Record view (no copy) of a StaticData's uint32[STATIC_WORDS] words."""
    return _view(world, STATIC, STATIC_WORDS, "static_view")


def _leaf_kind(dt):
    return "i" if dt in (u64, handle_edges, edge) else "f"


def _is_leaf_struct(dt):
    """Vectors, matrices, boxes, u64 pairs and the u16 edge words: every word belongs to the one named field
    (layout.rs:67-85, :148-176)."""
    return dt in (vec3, quat, mat3, affine3, aabb, u64, handle_edges, edge)


def _leaves(dt, path, out):
    if dt.subdtype is not None:  # an array: every element under the array's name
        base, shape = dt.subdtype
        for _ in range(int(np.prod(shape))):
            _leaves(base, path, out)
    elif dt.fields is None:  # one word
        out.append((path, "f" if dt.kind == "f" else "i"))
    elif _is_leaf_struct(dt):
        out.extend([(path, _leaf_kind(dt))] * (dt.itemsize // 4))
    else:
        for name in dt.names:
            _leaves(dt.fields[name][0], name if not path else path + "." + name, out)


_ARENA_LEAVES = None


def arena_leaves():
    """(field path, "f" or "i") of every word of an ArenaState, the word-by-word naming the comparator reports diffs in
    (layout.rs:490-496). A vector, matrix, box or u64 is one field; the two u16 of an edge word are one."""
    global _ARENA_LEAVES
    if _ARENA_LEAVES is None:
        out = []
        _leaves(ARENA, "", out)
        assert len(out) == ARENA_WORDS, len(out)
        _ARENA_LEAVES = out
    return _ARENA_LEAVES
