"""TEST INFRASTRUCTURE: an independent encoder of V3 arena and static dumps, written from V3's own export code
(v3/rocketsim-dump.patch: Arena::dump_io and export_state, io_ball, io_car, io_car_state, io_pad, WheelInfo::dump_io,
RigidBody::dump_io / dump_static, PersistentManifold::dump_io, CollisionDispatcher::dump_io, GridBroadphase::dump_io,
AxisSweep3::dump_io, OverlappingPairCache::dump_io, SimulationIslandManager::dump_io, ArenaContactTracker::dump_io,
static_sections, BvhTriangleMeshShape::dump_static, Tree::dump_static, StaticPlaneShape::dump_static) over the named
layout of the state (rsim.layout, the field names of crates/rocketsim-cuda/src/layout.rs).

The library's codec (csrc/host/v3wire.cpp) follows the reference's reader (crates/rocketsim-cuda/src/dump.rs); this one
follows the writer on V3's side, so a field order or width that one of them gets wrong shows up as a byte difference.
Where the state leaves something out, the choices are the reference's (dump.rs): the static bodies' records come from
the template, the island manager's lists are written empty, the manifold table has stride 32, and V3's contact records
are the unfolded ones (dump.rs:425-447).
"""
import struct

import numpy as np

from rsim import layout

UI_NONE = 0
EV_CAR_HIT_WORLD, EV_BALL_HIT_WORLD, EV_BALL_WALL = 10, 12, 14


class Enc:
    def __init__(self):
        self.b = bytearray()

    def raw(self, x):
        self.b += bytes(x)

    def u8(self, v):
        self.b += struct.pack("<B", int(v))

    def u16(self, v):
        self.b += struct.pack("<H", int(v))

    def u32(self, v):
        self.b += struct.pack("<I", int(v))

    def u64(self, v):
        self.b += struct.pack("<Q", int(v))

    def f32(self, v):
        self.b += np.asarray(v, np.float32).tobytes()

    def bool(self, v):
        self.u8(1 if int(v) else 0)

    def floats(self, rec):
        """A vector, quaternion, matrix, affine transform or box: its f32 in field order (glam's column order)."""
        self.b += rec.tobytes()

    def u64w(self, r):
        self.u64(int(r["lo"]) | int(r["hi"]) << 32)

    def opt(self, o, f):
        self.bool(o["some"])
        f(o["v"])


# ---- arena pieces ------------------------------------------------------------------------------------------------------------
def io_controls(e, c):
    for k in ("throttle", "steer", "pitch", "yaw", "roll"):
        e.f32(c[k])
    for k in ("jump", "boost", "handbrake"):
        e.bool(c[k])


def io_hit(e, h):
    e.floats(h["hit_point"])
    e.floats(h["hit_normal"])
    e.f32(h["hit_fraction"])
    e.u8(h["user_info"])


def io_manifold_point(e, p, applied=0.0):
    for k in ("local_point_a", "local_point_b", "pos_world_on_b", "pos_world_on_a", "normal_world_on_b"):
        e.floats(p[k])
    for k in ("distance_1", "combined_friction", "combined_restitution"):
        e.f32(p[k])
    e.f32(applied)
    e.bool(p["is_special"])


def io_wheel(e, w):
    def ray(r):
        e.floats(r["contact_normal"])
        e.floats(r["contact_point"])
        e.u32(r["ground_body_idx"])
        e.f32(r["suspension_length"])
        e.floats(r["impulse"])
        e.opt(r["ground_stick"], e.floats)
        e.bool(r["is_in_contact_with_world"])
        e.f32(r["clipped_inv_contact_dot_suspension"])
        e.f32(r["suspension_relative_vel"])

    e.opt(w["raycast_info"], ray)
    e.floats(w["hard_point"])
    e.floats(w["axle_dir"])
    for k in ("engine_force", "brake", "steer_angle"):
        e.f32(w[k])
    e.floats(w["vel_at_contact_point"])
    for k in ("lat_friction", "long_friction", "extra_pushback"):
        e.f32(w[k])


def io_car(e, c):
    e.f32(c["dodge_deadzone"])
    e.floats(c["vel_impulse_cache"])
    e.bool(c["sticky_gate_prev"])
    # io_car_state
    for k in ("pos", "rot_mat", "vel", "ang_vel"):
        e.floats(c[k])
    io_controls(e, c["controls"])
    for k in ("has_jumped", "has_double_jumped", "has_flipped"):
        e.bool(c[k])
    e.floats(c["flip_rel_torque"])
    e.u32(c["jump_ticks"])
    e.f32(c["flip_time"])
    for k in ("is_flipping", "is_jumping", "jump_hold_broken"):
        e.bool(c[k])
    for k in ("air_time", "air_time_since_jump", "boost", "handbrake_val"):
        e.f32(c[k])
    e.bool(c["is_supersonic"])
    for k in ("supersonic_grace_timer", "time_since_boosted", "boosting_time"):
        e.f32(c[k])
    e.bool(c["is_boosting"])
    e.bool(c["is_auto_flipping"])
    for k in ("auto_flip_timer", "auto_flip_torque_scale", "bump_cooldown_timer"):
        e.f32(c[k])
    e.opt(c["bump_cooldown_victim"], e.u32)
    e.opt(c["last_extra_hit_tick"], e.u64w)
    e.bool(c["is_demoed"])
    e.f32(c["demo_respawn_timer"])
    e.bool(c["is_on_ground"])
    for h in c["wheels_with_contact"]:
        e.opt(h, lambda x: io_hit(e, x))
    e.opt(c["world_contact_normal"], e.floats)
    io_controls(e, c["prev_controls"])
    for w in c["wheels"]:
        io_wheel(e, w)


def io_rigid_body(e, b):
    e.u8(b["collision_flags"])
    e.bool(0)  # companion_id: none
    e.u32(0)
    e.u8(b["activation"])
    e.floats(b["world_trans"])
    e.floats(b["world_quat"])
    e.floats(b["interp_world_trans"])
    e.f32(b["deactivation_time"])
    for k in ("lin_vel", "ang_vel", "inv_inertia_tensor_world", "accum_lin_vel", "accum_ang_vel"):
        e.floats(b[k])
    for k in ("cached_linear_damping", "cached_damping_time_step", "cached_linear_damping_value"):
        e.f32(b[k])


def unfolded_records(a, sd):
    """(is_swap, a, b, pos_on_b, pos_on_a, normal) of the records that fold to the state (dump.rs:425-447)."""
    world = next(i for i in range(layout.NUM_BODIES) if sd["bodies"][i]["user_idx"] == UI_NONE)
    zero = np.zeros(3, np.float32)
    out = []
    ev = int(a["fold_events"])
    if ev >> EV_BALL_HIT_WORLD & 1:
        n = np.array([1, 0, 0] if ev >> EV_BALL_WALL & 1 else [0, 0, 1], np.float32)
        out.append((0, int(sd["dyn_body"][0]), world, zero, zero, n))
    for c in range(layout.NUM_CARS):
        if ev >> (EV_CAR_HIT_WORLD + c) & 1:
            out.append((0, int(sd["dyn_body"][1 + c]), world, zero, zero,
                        np.frombuffer(a["fold_car_normal"][c].tobytes(), np.float32)))
    for r in a["records"][:int(a["n_records"])]:
        p = r["point"]
        out.append((int(r["is_swap"]), int(r["rb_idx_a"]), int(r["rb_idx_b"]),
                    np.frombuffer(p["pos_world_on_b"].tobytes(), np.float32),
                    np.frombuffer(p["pos_world_on_a"].tobytes(), np.float32),
                    np.frombuffer(p["normal_world_on_b"].tobytes(), np.float32)))
    return out


def table_index(p0, p1):
    lo, hi = min(p0, p1), max(p0, p1)
    return hi * (hi + 1) // 2 + lo


def arena_dump(state, world, tpl, island=None, mutate=None):
    """The V3 arena dump of one ArenaState (uint32 words) in a world (StaticData words), with the template `tpl`
    (rsim.v3.Template). `island`: the island manager's six lists (V3 writes its own; the reference writes them empty).
    `mutate(name, value) -> value` may alter named scalars on the way out (for tests of the reader's refusals)."""
    a = layout.arena_view(np.ascontiguousarray(state, np.uint32))
    sd = layout.static_view(np.ascontiguousarray(world, np.uint32))
    m = mutate or (lambda name, v: v)
    e = Enc()
    e.raw(b"ARNA")
    e.u64w(a["tick_count"])
    e.u64w(a["rng"])
    e.bool(m("ball_only", 0))
    e.raw(b"BALL")
    ball = a["ball"]
    for k in ("pos", "rot_mat", "vel", "ang_vel"):
        e.floats(ball[k])
    e.raw(bytes(tpl.ball_other))
    e.u64w(ball["tick_count_since_kickoff"])
    e.bool(m("ground_stick_applied", 0))
    e.raw(b"CARS")
    e.u32(layout.NUM_CARS)
    for c in a["cars"]:
        io_car(e, c)
    e.raw(b"PADS")
    e.u32(layout.NUM_PADS)
    for gave, touching in zip(a["pad_gave_boost"], a["pad_touching"]):
        e.opt(gave, e.u64w)  # an i64 option: the same eight bytes
        n = int(touching["n"])
        e.u32(n)
        for k in range(n):
            e.u32(touching["car_idx"][k])
    # CONT: ArenaContactTracker::dump_io
    e.raw(b"CONT")
    recs = unfolded_records(a, sd)
    e.u32(len(recs))
    for is_swap, ra, rb, on_b, on_a, normal in recs:
        e.bool(is_swap)
        e.u32(ra)
        e.u32(rb)
        for v in (np.zeros(3, np.float32), np.zeros(3, np.float32), on_b, on_a, normal):
            e.raw(np.asarray(v, np.float32).tobytes())
        for _ in range(4):
            e.f32(0.0)  # distance, friction, restitution, applied impulse
        e.bool(0)
    e.u32(m("ball_world_records", 0))
    e.bool(0)
    # the Bullet world: DiscreteDynamicsWorld::dump_io
    e.raw(b"BODY")
    e.u32(layout.NUM_BODIES)
    dyn = [int(x) for x in sd["dyn_body"]]
    for i in range(layout.NUM_BODIES):
        if i in dyn:
            io_rigid_body(e, a["bodies"][dyn.index(i)])
        else:
            e.raw(bytes(tpl.static_body[i]))
    e.raw(b"DISP")
    nm = int(a["n_manifolds"])
    e.u32(nm)
    for mf in a["manifolds"][:nm]:
        b0, b1 = int(mf["body0_idx"]), int(mf["body1_idx"])
        e.u32(b0)
        e.u32(b1)
        e.u64(m("pair_key", min(b0, b1) << 32 | max(b0, b1)))
        bb = sd["bodies"]
        e.f32(m("cbt", min(bb[b0]["contact_breaking_threshold"], bb[b1]["contact_breaking_threshold"])))
        e.f32(min(bb[b0]["contact_processing_threshold"], bb[b1]["contact_processing_threshold"]))
        e.u32(mf["point_count"])
        for p in mf["points"]:
            io_manifold_point(e, p, m("applied_impulse", 0.0))
    na = int(a["n_active"])
    e.u32(na)
    for x in a["active"][:na]:
        e.u32(x)
    stride = 32 if nm else tpl.manifold_stride
    e.u32(stride)
    table = np.zeros(stride * stride, np.uint32)
    for i, mf in enumerate(a["manifolds"][:nm]):
        b0, b1 = int(mf["body0_idx"]), int(mf["body1_idx"])
        table[min(b0, b1) * stride + max(b0, b1)] = i + 1
    table = m("manifold_table", table)
    e.u32(table.size)
    e.raw(table.astype("<u4").tobytes())
    # BROA: GridBroadphase::dump_io, then AxisSweep3::dump_io, then OverlappingPairCache::dump_io
    e.raw(b"BROA")
    cells = []
    for dc in a["dyn_cells"][:int(a["n_dyn_cells"])]:
        if cells and cells[-1][0] == int(dc["cell"]):
            cells[-1][1].append(int(dc["handle"]))
        else:
            cells.append((int(dc["cell"]), [int(dc["handle"])]))
    cells = m("dyn_cells", cells)
    e.u32(len(cells))
    for c, hs in cells:
        e.u32(c)
        e.u32(len(hs))
        for h in hs:
            e.u32(h)
    e.u32(layout.NUM_HANDLES)
    dyn_h = [int(x) for x in sd["dyn_handle"]]
    for h in range(layout.NUM_HANDLES):
        p = sd["proxies"][h]
        e.u32(p["client_obj_idx"])
        e.u8(p["filter_group"])
        e.u8(p["filter_mask"])
        e.u32(p["unique_id"])
        src = a["dyn_proxy"][dyn_h.index(h)] if h in dyn_h else p
        e.floats(src["aabb"])
        e.u32(src["cell_idx"])
        for i in src["indices"]:
            e.u32(m("static_proxy_index", int(i)) if h not in dyn_h else i)
    e.u32(layout.NUM_HANDLES)
    for he in a["handle_edges"]:
        for v in list(he["min"]) + list(he["max"]):
            e.u16(v)
    for axis in a["edges"]:
        e.u32(layout.NUM_EDGES)
        for ed in axis:
            e.u16(ed["pos"])
            e.u16(ed["handle"])
    e.u16(tpl.num_handles)
    free = m("free_list", [])
    e.u32(len(free))
    for f in free:
        e.u16(f)
    e.u16(tpl.next_new_handle)
    np_ = int(a["n_pairs"])
    e.u32(np_)
    for pr in a["pairs"][:np_]:
        e.u32(pr["proxy0"])
        e.u32(pr["proxy1"])
    want = {table_index(int(pr["proxy0"]), int(pr["proxy1"])): i + 1 for i, pr in enumerate(a["pairs"][:np_])}
    length = max(want) + 1 if want else 0
    pt = np.zeros(length, np.uint32)
    for i, v in want.items():
        pt[i] = v
    pt = m("pair_indices", pt)
    e.u32(pt.size)
    e.raw(pt.astype("<u4").tobytes())
    # ISLD: SimulationIslandManager::dump_io (body_to_dynamic, parents: idx; island_tags, sort_keys: i64; the two caches:
    # idx), then the world's ball_only
    e.raw(b"ISLD")
    lists = island or ([], [], [], [], [], [])
    for vals, fmt in zip(lists, ("<I", "<I", "<q", "<q", "<I", "<I")):
        e.u32(len(vals))
        for v in vals:
            e.raw(struct.pack(fmt, v))
    e.bool(0)

    n_points = int(sum(int(mf["point_count"]) for mf in a["manifolds"][:nm]))
    counts = [layout.NUM_BODIES, nm, na, n_points, stride * stride, layout.NUM_HANDLES, np_, int(a["n_dyn_cells"]),
              layout.NUM_CARS, layout.NUM_PADS, len(recs), 0]
    head = b"V3ARENA\0" + struct.pack("<IQ", m("version", 2), tpl.config_hash) + struct.pack("<12I", *counts)
    return head + bytes(e.b)


# ---- static -----------------------------------------------------------------------------------------------------------------------
def static_dump(world):
    """The V3 static dump of a StaticData, in the section order of V3's static_sections / export_static."""
    sd = layout.static_view(np.ascontiguousarray(world, np.uint32))
    wr, bodies, meshes, planes, cells, cars, pads, prox = (Enc() for _ in range(8))
    # WRLD: the arena config, then DiscreteDynamicsWorld::dump_static (gravity, dynamic bodies), then the grid and
    # sweep-and-prune constants
    wr.u8(sd["game_mode"])
    wr.bool(sd["no_ball_rot"])
    wr.u32(sd["ball_body"])
    wr.floats(sd["gravity_uu"])
    for f in sd["mut_f"]:
        wr.f32(f)
    for k in ("unlimited_flips", "unlimited_double_jumps", "recharge_boost_enabled"):
        wr.bool(sd[k])
    wr.f32(sd["recharge_boost_per_second"])
    wr.f32(sd["recharge_boost_delay"])
    wr.u8(sd["demo_mode"])
    wr.bool(sd["enable_team_demos"])
    wr.f32(sd["goal_base_threshold_y"])
    wr.bool(sd["boost_pads_enabled"])
    wr.floats(sd["gravity_bt"])
    wr.u32(layout.NUM_DYN)
    for d in sd["dyn_body"]:
        wr.u32(d)
    wr.floats(sd["grid_min_pos"])
    wr.floats(sd["grid_max_pos"])
    wr.f32(sd["grid_cell_size"])
    wr.f32(sd["grid_cell_size_sq"])
    for v in sd["grid_num_cells"]:
        wr.u32(v)
    wr.floats(sd["sap_world_aabb_min"])
    wr.floats(sd["sap_quantize"])
    for k in ("sap_handle_mask", "sap_sentinel", "sap_max_handles"):
        wr.u16(sd[k])
    wr.f32(sd["sap_world_half_extent"])
    # BODY: RigidBody::dump_static per body, the mesh's triangle count after a mesh body
    bodies.u32(layout.NUM_BODIES)
    for b in sd["bodies"]:
        bodies.u32(b["world_array_idx"])
        bodies.u8(b["user_idx"])
        bodies.u32(b["user_pointer"])
        bodies.u32(b["broadphase_handle"])
        bodies.u8(b["filter_group"])
        bodies.u8(b["filter_mask"])
        bodies.f32(b["mass"])
        bodies.f32(b["inv_mass"])
        bodies.floats(b["inv_inertia_local"])
        bodies.floats(b["inv_mass_splat"])
        for k in ("friction", "restitution", "contact_processing_threshold", "contact_breaking_threshold",
                  "linear_damping", "linear_sleeping_threshold", "angular_sleeping_threshold"):
            bodies.f32(b[k])
        bodies.floats(b["pose"])
        kind = int(b["shape_kind"])
        bodies.u8(kind)
        if kind == 0:
            bodies.u32(b["mesh_tri_count"])
        elif kind == 2:
            bodies.floats(b["sphere_implicit_dim"])
            bodies.f32(b["sphere_margin"])
        elif kind == 3:
            bodies.floats(b["box_half_extents"])
            bodies.f32(b["box_margin"])
            bodies.floats(b["child_trans"])
            bodies.floats(b["compound_local_aabb"])
    # MESH: per mesh its body, then BvhTriangleMeshShape::dump_static (Tree::dump_static inside)
    for mi in sd["meshes"]:
        meshes.u32(mi["body"])
        meshes.floats(mi["aabb_ident_cache"])
        t0, nt = int(mi["tri_first"]), int(mi["tri_count"])
        meshes.u32(nt)
        meshes.raw(sd["tris"][t0:t0 + nt].tobytes())
        meshes.floats(mi["root_aabb"])
        n0, nn = int(mi["node_first"]), int(mi["node_count"])
        meshes.u32(nn)
        for node in sd["nodes"][n0:n0 + nn]:
            for k in ("min_x", "min_y", "min_z", "max_x", "max_y", "max_z", "children"):
                meshes.raw(node[k].tobytes())
            meshes.u8(node["child_count"])
            meshes.u32(node["valid_mask"])
        l0, nl = int(mi["leaf_first"]), int(mi["leaf_count"])
        meshes.u32(nl)
        meshes.raw(sd["leaves"][l0:l0 + nl].tobytes())
        meshes.u32(nt)
        for info in sd["infos"][t0:t0 + nt]:
            meshes.u8(info["flags"])
            for k in ("edge_v0_v1_angle", "edge_v1_v2_angle", "edge_v2_v0_angle"):
                meshes.f32(info[k])
    # PLAN: the plane's body, then StaticPlaneShape::dump_static
    for p in sd["planes"]:
        planes.u32(p["body"])
        planes.floats(p["plane_normal"])
        planes.bool(p["is_single_axis"])
        planes.u32(p["single_axis_idx"])
        planes.bool(p["single_axis_backwards"])
        planes.floats(p["aabb_ident_cache"])
        planes.floats(p["aabb_cache"])
    # CELL: the occupied static ray cells, ascending
    occupied = [(i, c) for i, c in enumerate(sd["cells"]) if int(c["count"])]
    cells.u32(len(occupied))
    for i, c in occupied:
        cells.u32(i)
        cells.u32(c["count"])
        f = int(c["first"])
        cells.raw(sd["cell_handles"][f:f + int(c["count"])].astype("<u4").tobytes())
    for c in sd["cars"]:
        cars.u32(c["car_idx"])
        cars.u8(c["team"])
        cars.u32(c["rigid_body_idx"])
        cars.bool(c["three_wheels"])
        cars.floats(c["hitbox_size"])
        cars.floats(c["hitbox_pos_offset"])
        for w in c["wheels"]:
            cars.floats(w["chassis_connection_point_cs"])
            for k in ("suspension_rest_length_1", "wheels_radius", "suspension_force_scale", "real_ray_length"):
                cars.f32(w[k])
    pads.f32(sd["max_pad_z"])
    for p in sd["pads"]:
        pads.floats(p["pos"])
        pads.bool(p["is_big"])
        for k in ("box_radius", "cyl_radius", "max_cooldown", "boost_amount"):
            pads.f32(p[k])
        pads.floats(p["aabb"])
    # PROX: GridBroadphase::dump_proxies
    prox.u32(layout.NUM_HANDLES)
    for p in sd["proxies"]:
        prox.u32(p["client_obj_idx"])
        prox.u8(p["filter_group"])
        prox.u8(p["filter_mask"])
        prox.u32(p["unique_id"])
        prox.floats(p["aabb"])
        prox.u32(p["cell_idx"])
        prox.raw(p["indices"].astype("<u4").tobytes())

    def section(tag, body, n=None):
        return tag + (struct.pack("<I", n) if n is not None else b"") + bytes(body.b)

    payload = (section(b"WRLD", wr) + section(b"BODY", bodies) + section(b"MESH", meshes, layout.NUM_MESHES) +
               section(b"PLAN", planes, layout.NUM_PLANES) + section(b"CELL", cells) +
               section(b"CARS", cars, layout.NUM_CARS) + section(b"PADS", pads, layout.NUM_PADS) + section(b"PROX", prox))
    tris = int(sd["meshes"]["tri_count"].sum())
    counts = [layout.NUM_BODIES, layout.NUM_DYN, layout.NUM_MESHES, tris, int(sd["meshes"]["node_count"].sum()),
              int(sd["meshes"]["leaf_count"].sum()), tris, layout.NUM_PLANES, layout.NUM_CARS, layout.NUM_PADS,
              layout.NUM_GRID_CELLS, int(sd["cells"]["count"].sum())]
    return (b"V3STATIC" + struct.pack("<I", 3) + struct.pack("<QQ", fnv1a64(payload), len(payload)) +
            struct.pack("<12I", *counts) + payload)


def fnv1a64(data):
    """FNV-1a 64 in plain Python (the published algorithm: offset basis 14695981039346656037, prime 1099511628211)."""
    h = 14695981039346656037
    for byte in data:
        h = ((h ^ byte) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h
