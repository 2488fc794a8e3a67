"""Start states and worlds crafted to reach the decision's rare outcomes, each checked against the reference's own build:

  overflow       the retained pair list filled to its capacity (16 entries, the free slots holding copies of the first
                 pair) so that the next new overlap of a contact run has no room: the pair cache's overflow counter
                 (cuda/sim/overlapping_pair_cache.cuh add_overlapping_pair; arena_state.cuh:52-61) and RSIM_SIM8_OVERFLOW;
  sticky         an overflow counter already nonzero in the start state: the result word reports it after any decision
                 (the flag reads the written-back state, arena_step.cuh:21-25);
  unimplemented  a world whose mesh table disagrees with its bodies (every MeshInfo.body past the mesh range), so each
                 ball-mesh or car-mesh pair reaches the dispatcher's no-algorithm branch (bvh.cuh:69-71 mesh_of_body;
                 collision_dispatcher.cuh:98) and the result word carries RSIM_SIM8_UNIMPLEMENTED;
  goal           the ball placed in a goal mouth, moving along y at 3000 uu/s (25 uu per tick) from half a tick's travel
                 short of tick k's crossing of the goal threshold, for k = 1..8 and both goals: the decision ends after
                 tick k (arena_step.cuh:16-34) and no later tick is written;
  autoflip       a kickoff car turned upside down (rolled half a turn about its forward axis) and lowered onto its roof:
                 one decision without jump gives it a world contact normal, the next presses jump, which takes the
                 auto-flip path with its f64 atan2 / hypot and the sine table (car_pre_tick.cuh:61-69).
"""
import numpy as np

from rsim import layout

TICK_WORDS = 2567
GOAL_SPEED = 3000.0  # uu/s: 25 uu per tick at 120 Hz
UU_TO_BT = np.float32(1.0 / 50.0)


def overflow_states(rsim, world, n, seed):
    s = rsim.sample_arenas(world, rsim.SCEN_CONTACT, seed, n)
    a = layout.arena_view(s)
    for i in range(n):
        k = int(a["n_pairs"][i])
        if k == 0:
            continue
        for j in range(k, layout.MAX_PAIRS):
            a["pairs"][i, j] = a["pairs"][i, 0]
        a["n_pairs"][i] = layout.MAX_PAIRS
    return s


def sticky_states(rsim, world, n, seed):
    s = rsim.sample_arenas(world, rsim.SCEN_MIXED, seed, n)
    a = layout.arena_view(s)
    for i in range(n):
        a["overflow"][i, i % layout.OVF_COUNT] = 1 + i % 3
    return s


def unimplemented_world(world):
    w = world.copy()
    sv = layout.static_view(w)
    sv["meshes"]["body"][:] = 99
    return w


def goal_states(rsim, world, seed):
    """16 arenas per tick k = 1..8 (half at each goal); returns (states, expected scored tick per arena)."""
    thr = ball_goal_threshold(world)
    states = rsim.sample_arenas(world, rsim.SCEN_KICKOFF, seed, 128)
    a = layout.arena_view(states)
    want = np.zeros(128, np.uint32)
    for i in range(128):
        k = 1 + i // 16
        sign = 1.0 if i % 2 == 0 else -1.0
        want[i] = k
        y0 = sign * (thr - 25.0 * (k - 0.5))
        place_ball(a, i, (float((i % 7) - 3) * 60.0, y0, 321.0), (0.0, sign * GOAL_SPEED, 0.0))
    return states, want


MF_BALL_RADIUS = 19  # the ball-radius entry of StaticData.mut_f (arena_state.cuh mutator order)


def ball_goal_threshold(world):
    """|ball.y| past which the ball is scored: goal_base_threshold_y + the ball radius mutator (arena_step.cuh:27-34)."""
    sv = layout.static_view(world)
    return float(np.float32(sv["goal_base_threshold_y"]) + np.float32(sv["mut_f"][MF_BALL_RADIUS]))


def place_ball(a, i, pos, vel):
    """Moves arena i's ball: the body's transforms (BT units) and the ball's position and velocity (uu)."""
    p = np.array(pos, np.float32)
    v = np.array(vel, np.float32)
    for f, x in zip("xyz", p * UU_TO_BT):
        a["bodies"][i, 0]["world_trans"]["translation"][f] = x
        a["bodies"][i, 0]["interp_world_trans"]["translation"][f] = x
    for f, x in zip("xyz", v * UU_TO_BT):
        a["bodies"][i, 0]["lin_vel"][f] = x
        a["bodies"][i, 0]["ang_vel"][f] = 0.0
    for f, x, y in zip("xyz", p, v):
        a["ball"][i]["pos"][f] = x
        a["ball"][i]["vel"][f] = y
    a["bodies"][i, 0]["activation"] = 0


def upside_down_states(rsim, world, seed, n=64):
    """Kickoff arenas with car 0 rolled half a turn about its forward axis, lying on its roof."""
    states = rsim.sample_arenas(world, rsim.SCEN_KICKOFF, seed, n)
    a = layout.arena_view(states)
    for i in range(n):
        for m in (a["cars"][i, 0]["rot_mat"], a["bodies"][i, 1]["world_trans"]["matrix3"],
                  a["bodies"][i, 1]["interp_world_trans"]["matrix3"]):
            for ax in ("y_axis", "z_axis"):
                for f in "xyz":
                    m[ax][f] = -m[ax][f]
        q = a["bodies"][i, 1]["world_quat"]
        x, y, z, w = float(q["x"]), float(q["y"]), float(q["z"]), float(q["w"])
        # q * (1, 0, 0, 0): a half turn about the local x (forward) axis, exact in float
        q["x"], q["y"], q["z"], q["w"] = np.float32(w), np.float32(z), np.float32(-y), np.float32(-x)
        zz = np.float32(30.0)
        a["cars"][i, 0]["pos"]["z"] = zz
        a["bodies"][i, 1]["world_trans"]["translation"]["z"] = zz * UU_TO_BT
        a["bodies"][i, 1]["interp_world_trans"]["translation"]["z"] = zz * UU_TO_BT
        for f in "xyz":
            a["bodies"][i, 1]["lin_vel"][f] = 0.0
            a["bodies"][i, 1]["ang_vel"][f] = 0.0
            a["cars"][i, 0]["vel"][f] = 0.0
            a["cars"][i, 0]["ang_vel"][f] = 0.0
    return states


def autoflip_controls(rsim, n, decision):
    """Decision 0: every control zero. Decision 1: car 0 presses jump (nothing else), car 1 idle."""
    c = np.zeros((n, 16), np.uint32)
    if decision == 1:
        c[:, 5] = 1
    return c
