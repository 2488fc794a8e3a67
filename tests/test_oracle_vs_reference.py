"""The CPU oracle against the reference's own C++ build (oracle/_ref/libref.so, compiled in place from the reference
sources by `make -C oracle ref`) on fresh seeded inputs: whole states after every tick, every phase, every decision.
Skipped where the reference tree is absent (the GPU host); the golden vectors carry the same check there."""
import os

import numpy as np
import pytest

import oracle as orc

pytestmark = pytest.mark.skipif(not os.path.exists(orc.REF_LIB), reason="reference build (oracle/_ref) not available")


@pytest.fixture(scope="module")
def ref():
    return orc.Reference()


@pytest.mark.parametrize("sc", [0, 1, 2, 3, 4])
def test_decisions_vs_reference(oracle, ref, sc):
    world = oracle.build_world(0, 0)
    n = 48
    states = oracle.sample_arenas(world, sc, 9000 + sc, n)
    mine, theirs = states.copy(), states.copy()
    for d in range(8):
        ctl = oracle.sample_controls(9100 + sc, d, n)
        for a in range(n):
            r1 = oracle.step_ticks(mine[a], world, ctl[a], 8)
            r2 = ref.step_ticks(theirs[a], world, ctl[a], 8)
            assert r1 == r2, "scenario %d decision %d arena %d" % (sc, d, a)
            if not np.array_equal(mine[a], theirs[a]):
                w = int(np.argmax(mine[a] != theirs[a]))
                pytest.fail("scenario %d decision %d arena %d: word %d differs" % (sc, d, a, w))


@pytest.mark.parametrize("presets", [(1, 3), (2, 5), (4, 6)])
def test_presets_and_mutators_vs_reference(oracle, ref, presets):
    rng = np.random.default_rng(presets[0])
    world = oracle.build_world(presets[0], presets[1], float(rng.uniform(-800, -500)), float(rng.uniform(0.8, 1.2)),
                               float(rng.uniform(91.25, 100)))
    states = oracle.sample_arenas(world, 6, 77 + presets[1], 40)
    mine, theirs = states.copy(), states.copy()
    for d in range(5):
        ctl = oracle.sample_controls(78, d, 40)
        for a in range(40):
            assert oracle.step_ticks(mine[a], world, ctl[a], 8) == ref.step_ticks(theirs[a], world, ctl[a], 8)
        assert np.array_equal(mine, theirs), "presets %s decision %d" % (presets, d)


def test_phases_vs_reference(oracle, ref):
    world = oracle.build_world(0, 0)
    states = oracle.sample_arenas(world, 6, 4321, 32)
    ctl = oracle.sample_controls(4322, 0, 32)
    for a in range(32):
        mine, theirs = states[a].copy(), states[a].copy()
        for t in range(6):  # several ticks, phase by phase
            oracle.set_controls(mine, ctl[a])
            oracle.set_controls(theirs, ctl[a])
            for k in range(10):
                assert oracle.phase(k, mine, world) == ref.phase(k, theirs, world)
                assert np.array_equal(mine, theirs), "arena %d tick %d phase %d" % (a, t, k)
    s = states[0].copy()
    assert oracle.phase(10, s, world) == ref.phase(10, s.copy(), world) == -2
    assert oracle.phase(-1, s, world) == -2


def test_threaded_reference_batch(rsim, oracle, ref):
    """The reference's decision over host threads (oracle/ref_batch.cpp, the bench's CPU baseline) gives the oracle's
    result words and states, and the same as the reference called arena by arena."""
    w = rsim.build_world(0, 0)
    st = rsim.sample_arenas(w, rsim.SCEN_CONTACT, 515, 96)
    a, b, c = st.copy(), st.copy(), st.copy()
    for d in range(4):
        ctl = rsim.sample_controls(516, d, 96)
        ra = ref.sim8(a, w, ctl, threads=5)
        rb = oracle.sim8(b, w, ctl, threads=3)
        rc = []
        for i in range(96):
            status, scored = ref.step_ticks(c[i], w, ctl[i], 8)
            rc.append(orc.result_word(status, scored, c[i]))
        assert np.array_equal(ra, rb) and np.array_equal(ra, np.array(rc, np.uint32)), d
        assert np.array_equal(a, b) and np.array_equal(a, c), d


def test_respawn_beside_wheel_rays(rsim, oracle, ref):
    """A car respawning in the same tick as the other car casts its wheel rays: one car of each arena is demoed with its
    respawn timer a few ticks from zero (a different tick per arena), so the respawn (which moves that car's body) falls
    inside P1 of one of the first decisions, and the whole states and result words must equal the reference's."""
    from rsim import layout
    w = rsim.build_world(0, 0)
    n = 256
    st = rsim.sample_arenas(w, rsim.SCEN_CONTACT, 4321, n)
    cars = layout.arena_view(st)["cars"]
    for a in range(n):
        ci = a % 2
        cars["is_demoed"][a, ci] = 1
        cars["demo_respawn_timer"][a, ci] = np.float32((1 + a % 12) / 120.0)
    assert int(layout.arena_view(st)["cars"]["is_demoed"].sum()) == n  # the view writes through to the states
    a_, b_ = st.copy(), st.copy()
    for d in range(3):
        ctl = rsim.sample_controls(4322, d, n)
        ra = ref.sim8(a_, w, ctl, threads=4)
        rb = oracle.sim8(b_, w, ctl, threads=4)
        assert np.array_equal(ra, rb), d
        assert np.array_equal(a_, b_), d
    after = layout.arena_view(a_)["cars"]["is_demoed"]
    assert not after.any(), "every demoed car respawned"
