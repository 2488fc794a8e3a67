"""Invariants of the synthetic world and start states (host-side builder): the geometry, the BVHs, the broadphase built
by handle insertion, the ray-grid cell lists, and physical plausibility of a few ticks."""
import numpy as np
import pytest

WORDS = {}


def f32(a):
    return a.view(np.float32)


TRIS_AT, NODES_AT, LEAVES_AT = 146044 // 4, 531004 // 4, 981484 // 4  # StaticData word offsets (rs_state.h asserts them)


@pytest.mark.parametrize("presets", [(0, 0), (6, 2)])
def test_world_geometry(rsim, presets):
    w = rsim.build_world(*presets)
    assert w.size == rsim.STATIC_WORDS
    tris = f32(w[TRIS_AT:TRIS_AT + 8020 * 12]).reshape(8020, 4, 3)
    used = np.abs(tris).sum(axis=(1, 2)) > 0
    nt = int(used.sum())
    assert 7000 <= nt <= 8020, nt  # about 7.5k triangles, within the format's cap
    assert used[:nt].all() and not used[nt:].any()  # packed from the front
    t = tris[:nt].astype(np.float64)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    assert np.allclose(n, t[:, 3], atol=1e-5), "stored normals are the normalised edge cross products"
    assert np.allclose(np.linalg.norm(t[:, 3], axis=1), 1.0, atol=1e-6)
    # every vertex inside the arena box, goals included (BT units: x +-81.92, y +-(102.4 + 17.6), z 0..40.88)
    v = t[:, :3].reshape(-1, 3)
    assert (np.abs(v[:, 0]) <= 81.93).all() and (np.abs(v[:, 1]) <= 120.01).all()
    assert (v[:, 2] >= -1e-4).all() and (v[:, 2] <= 40.89).all()
    leaves = w[LEAVES_AT:LEAVES_AT + 8020 * 7].reshape(8020, 7)
    assert int((np.abs(f32(leaves[:, :6])).sum(1) > 0).sum()) == nt  # one leaf per triangle


def test_start_states_broadphase(rsim, oracle):
    """Sorted SAP edges with consistent handle back-references; the pair list equals the brute-force overlaps of the
    filtered handle pairs (rs_state.h layout: edges at byte 9532, handle_edges at 9244)."""
    w = rsim.build_world(0, 0)
    for sc in range(5):
        states = rsim.sample_arenas(w, sc, 17, 50)
        for s in states:
            b = s.view(np.uint8)
            he = np.frombuffer(b[9244:9244 + 24 * 12].tobytes(), np.uint16).reshape(24, 6)
            edges = np.frombuffer(b[9532:9532 + 3 * 48 * 4].tobytes(), np.uint16).reshape(3, 48, 2)
            for ax in range(3):
                pos = edges[ax, :, 0].astype(np.int64)
                assert (np.diff(pos) >= 0).all(), "edges not sorted on axis %d" % ax
                for i, (p, h) in enumerate(edges[ax]):
                    if p & 1:
                        assert he[h, 3 + ax] == i
                    else:
                        assert he[h, ax] == i


def test_resting_physics(rsim, oracle):
    """Kickoff with no input: cars stay on their wheels at z ~ 17 uu, the ball stays on the floor at about its radius."""
    w = rsim.build_world(0, 0)
    s = rsim.sample_arenas(w, 0, 3, 4)
    ctl = np.zeros((4, 16), np.uint32)
    for _ in range(15):
        oracle.sim8(s, w, ctl)
    pos0 = f32(s[:, 24 + 5:24 + 8])
    assert (np.abs(pos0[:, 2] - 17.0) < 1.0).all()
    ball_z = f32(s[:, 722 + 3 + 9 + 2]) * 50.0
    assert (np.abs(ball_z - 93.0) < 3.0).all()


def test_sample_in_worlds_matches_single_world_draws(rsim, oracle):
    """rsim_sample_arenas_in draws arena i exactly as rsim_sample_arenas does in worlds[variant[i]], and the oracle's
    per-arena-world batch equals per-arena step_ticks in each arena's world."""
    import numpy as np
    import oracle as orc

    worlds = np.stack([rsim.build_world(p, p) for p in (0, 3, 6)])
    variant = np.array([2, 0, 1, 1, 2, 0, 2, 1, 0, 0, 2, 1], np.uint32)
    s = rsim.sample_arenas_in(worlds, variant, rsim.SCEN_MIXED, 77, first=5)
    for i, v in enumerate(variant):
        assert np.array_equal(s[i], rsim.sample_arenas(worlds[v], rsim.SCEN_MIXED, 77, 1, first=5 + i)[0])
    ctl = rsim.sample_controls(78, 0, variant.size)
    a, b = s.copy(), s.copy()
    r = oracle.sim8_in(a, worlds, variant, ctl, threads=3)
    for i, v in enumerate(variant):
        assert r[i] == orc.result_word(*oracle.step_ticks(b[i], worlds[v], ctl[i], 8), b[i])
    assert np.array_equal(a, b)
    bad = variant.copy()
    bad[4] = 3
    with pytest.raises(rsim.RsimError):
        rsim.sample_arenas_in(worlds, bad, rsim.SCEN_MIXED, 77)


def test_relocated_mesh_nodes_same_physics(rsim, oracle):
    """A world whose meshes' BVH nodes sit elsewhere in the node array (node_first changed) steps exactly like the
    original: the fixture the GPU test of mixed world layouts relies on."""
    import numpy as np

    a = rsim.build_world(0, 0)
    b = a
    for k in (0, 4, 8, 12, 14):
        b = oracle.relocate_mesh_nodes(b, k)
        assert b is not None
    assert (b != a).sum() > 0
    states = rsim.sample_arenas(a, rsim.SCEN_CONTACT, 191, 64)
    ctl = rsim.sample_controls(192, 0, 64)
    sa, sb = states.copy(), states.copy()
    assert np.array_equal(oracle.sim8(sa, a, ctl, threads=4), oracle.sim8(sb, b, ctl, threads=4))
    assert np.array_equal(sa, sb)
