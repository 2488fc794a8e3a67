"""Parity of the gfx950 kernels with the CPU oracle, through librsim.so's C ABI: every TICK_WORDS word of every arena and
every result word must be identical (the oracle is itself pinned to the reference's C++ build by the golden vectors)."""
import os
import sys

import numpy as np
import pytest

from conftest import first_diff
from helpers import TICK_WORDS, digest

pytestmark = pytest.mark.gpu


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def gpu_decisions(rsim, device, dev_world, states, controls, block=64, trace=False):
    n = states.shape[0]
    d = device.upload_states(states)
    r = device.alloc_words(n)
    results, traces = [], []
    for ctl in controls:
        c = device.upload_controls(ctl)
        t = device.alloc_words(8 * rsim.aosoa_words(n)) if trace else None
        device.sim8(d, dev_world, c, r, block, t)
        device.sync()
        results.append(device.read_words(r))
        if trace:
            w = device.read_words(t)
            slabs = w.reshape(8, -1)
            traces.append(np.stack([rsim.aosoa_unpack(slabs[k], n) for k in range(8)], 1))
    return device.download_states(d), results, traces


def cpu_decisions(oracle, world, states, controls):
    s = states.copy()
    res = [oracle.sim8(s, world, ctl, threads=8) for ctl in controls]
    return s, res


def assert_same(gs, gr, cs, cr, what):
    for k, (a, b) in enumerate(zip(gr, cr)):
        if not np.array_equal(a, b):
            i = int(np.argmax(a != b))
            pytest.fail("%s: decision %d result word of arena %d: gpu %#x cpu %#x" % (what, k, i, a[i], b[i]))
    fd = first_diff(gs[:, :TICK_WORDS], cs[:, :TICK_WORDS])
    assert fd is None, "%s: arena %d word %d differs (gpu %#x cpu %#x)" % (what, fd[0], fd[1], gs[fd], cs[fd])
    assert not gs[:, TICK_WORDS:].any(), "%s: the mid-tick tail is not zero on the device" % what


def test_identity_round_trip(rsim, device):
    note("identity round trip, 300 random states")
    rng = np.random.default_rng(1)
    states = rng.integers(0, 2**32, (300, rsim.ARENA_WORDS), dtype=np.uint64).astype(np.uint32)
    a = device.upload_states(states)
    b = device.alloc_states(300)
    device.identity(a, b)
    assert np.array_equal(device.download_states(b), states)
    device.identity_tick(a, b)
    got = device.download_states(b)
    assert np.array_equal(got[:, :TICK_WORDS], states[:, :TICK_WORDS])
    # AoSoA placement: word w of arena i at (i / 64) * 2784 * 64 + w * 64 + i % 64
    words = device.download_words(a)
    for i, w in [(0, 0), (63, 5), (64, 0), (299, 2783), (130, 1000)]:
        assert words[rsim.aosoa_index(i, w)] == states[i, w]
        assert rsim.aosoa_index(i, w) == (i // 64) * 2784 * 64 + w * 64 + i % 64


def test_kickoff_1024(rsim, oracle, world, device, dev_world):
    note("kickoff 1024 arenas x 4 decisions")
    states = rsim.sample_arenas(world, rsim.SCEN_KICKOFF, 11, 1024)
    ctl = [rsim.sample_controls(12, d, 1024) for d in range(4)]
    gs, gr, _ = gpu_decisions(rsim, device, dev_world, states, ctl)
    cs, cr = cpu_decisions(oracle, world, states, ctl)
    assert_same(gs, gr, cs, cr, "kickoff")


@pytest.mark.parametrize("scenario", [1, 2, 3, 4, 5])
def test_scenarios(rsim, oracle, world, device, dev_world, scenario):
    note("scenario %d, 2048 arenas x 3 decisions" % scenario)
    states = rsim.sample_arenas(world, scenario, 100 + scenario, 2048)
    ctl = [rsim.sample_controls(200 + scenario, d, 2048) for d in range(3)]
    gs, gr, _ = gpu_decisions(rsim, device, dev_world, states, ctl)
    cs, cr = cpu_decisions(oracle, world, states, ctl)
    assert_same(gs, gr, cs, cr, "scenario %d" % scenario)


def test_respawn_beside_wheel_rays(rsim, oracle, world, device, dev_world):
    """One car of each arena demoed with its respawn timer 1-12 ticks from zero, so it respawns (its body moves) in P1
    of a tick in which the other car casts its wheel rays from the walks pooled at the start of that P1."""
    from rsim import layout
    note("respawn beside wheel rays, 2048 contact arenas x 3 decisions")
    n = 2048
    states = rsim.sample_arenas(world, rsim.SCEN_CONTACT, 4321, n)
    cars = layout.arena_view(states)["cars"]
    for a in range(n):
        cars["is_demoed"][a, a % 2] = 1
        cars["demo_respawn_timer"][a, a % 2] = np.float32((1 + a % 12) / 120.0)
    assert int(layout.arena_view(states)["cars"]["is_demoed"].sum()) == n
    ctl = [rsim.sample_controls(4322, d, n) for d in range(3)]
    gs, gr, _ = gpu_decisions(rsim, device, dev_world, states, ctl)
    cs, cr = cpu_decisions(oracle, world, states, ctl)
    assert_same(gs, gr, cs, cr, "respawn")
    assert not layout.arena_view(gs)["cars"]["is_demoed"].any(), "every demoed car respawned"


@pytest.mark.parametrize("presets", [(1, 2), (3, 4), (5, 6), (6, 0)])
def test_presets(rsim, oracle, device, presets):
    note("presets %s, contact mix, 512 arenas x 3 decisions" % (presets,))
    w = rsim.build_world(*presets)
    dw = device.upload_static(w)
    states = rsim.sample_arenas(w, rsim.SCEN_MIXED, 300 + presets[0], 512)
    ctl = [rsim.sample_controls(301, d, 512) for d in range(3)]
    gs, gr, _ = gpu_decisions(rsim, device, dw, states, ctl)
    cs, cr = cpu_decisions(oracle, w, states, ctl)
    assert_same(gs, gr, cs, cr, "presets %s" % (presets,))


def test_mutators(rsim, oracle, device):
    note("mutator worlds (gravity, boost, ball radius)")
    rng = np.random.default_rng(5)
    for k in range(3):
        g, b, r = float(rng.uniform(-800, -500)), float(rng.uniform(0.8, 1.2)), float(rng.uniform(91.25, 100))
        w = rsim.build_world(k, 6 - k, g, b, r)
        dw = device.upload_static(w)
        states = rsim.sample_arenas(w, rsim.SCEN_MIXED, 400 + k, 256)
        ctl = [rsim.sample_controls(401, d, 256) for d in range(2)]
        gs, gr, _ = gpu_decisions(rsim, device, dw, states, ctl)
        cs, cr = cpu_decisions(oracle, w, states, ctl)
        assert_same(gs, gr, cs, cr, "mutators %d" % k)


def test_trace(rsim, oracle, world, device, dev_world):
    note("sim8 trace, 256 arenas of the contact mix")
    n = 256
    states = rsim.sample_arenas(world, rsim.SCEN_CONTACT, 21, n)
    ctl = rsim.sample_controls(22, 0, n)
    gs, gr, gt = gpu_decisions(rsim, device, dev_world, states, [ctl], trace=True)
    for a in range(n):
        s = states[a].copy()
        tr = np.zeros((8, rsim.ARENA_WORDS), np.uint32)
        st, scored = oracle.step_ticks(s, world, ctl[a], 8, tr)
        last = scored if scored else 8
        for t in range(last):
            assert np.array_equal(gt[0][a, t, :TICK_WORDS], tr[t, :TICK_WORDS]), "arena %d tick %d" % (a, t + 1)
        for t in range(last, 8):  # ticks after a goal are not written
            assert not gt[0][a, t].any(), "arena %d: tick %d after the goal was written" % (a, t + 1)


def test_step_tick_and_phases(rsim, oracle, world, device, dev_world):
    note("step_tick and the ten phases, 256 arenas")
    n = 256
    states = rsim.sample_arenas(world, rsim.SCEN_MIXED, 31, n)
    ctl = rsim.sample_controls(32, 0, n)
    for a in range(n):
        oracle.set_controls(states[a], ctl[a])
    # whole tick
    d = device.upload_states(states)
    status = device.step_tick(d, dev_world)
    got = device.download_states(d)
    for a in range(n):
        s = states[a].copy()
        assert oracle.step_tick(s, world) == status[a]
        assert np.array_equal(got[a, :TICK_WORDS], s[:TICK_WORDS]), "step_tick arena %d" % a
    # phase by phase on whole states (the mid-tick tail moves between phases)
    d = device.upload_states(states)
    cpu = states.copy()
    for k in range(10):
        st = device.phase(k, d, dev_world)
        got = device.download_states(d)
        for a in range(n):
            assert oracle.phase(k, cpu[a], world) == st[a]
        fd = first_diff(got, cpu)
        assert fd is None, "phase %d: arena %d word %d" % (k, fd[0], fd[1])
    assert (device.phase(10, d, dev_world) == -2).all()  # PHASE_BAD_INDEX


@pytest.mark.parametrize("block", [1, 7, 32, 64, 128, 256])
def test_block_size_invariance(rsim, world, device, dev_world, block):
    note("block size %d" % block)
    n = 1000
    states = rsim.sample_arenas(world, rsim.SCEN_MIXED, 41, n)
    ctl = [rsim.sample_controls(42, d, n) for d in range(2)]
    base = gpu_decisions(rsim, device, dev_world, states, ctl, block=64)
    got = gpu_decisions(rsim, device, dev_world, states, ctl, block=block)
    assert np.array_equal(base[0][:, :TICK_WORDS], got[0][:, :TICK_WORDS])
    for a, b in zip(base[1], got[1]):
        assert np.array_equal(a, b)


def test_permutation_invariance(rsim, world, device, dev_world):
    note("permutation invariance, 1000 arenas")
    n = 1000
    states = rsim.sample_arenas(world, rsim.SCEN_MIXED, 51, n)
    ctl = [rsim.sample_controls(52, d, n) for d in range(2)]
    perm = np.random.default_rng(53).permutation(n)
    base = gpu_decisions(rsim, device, dev_world, states, ctl)
    got = gpu_decisions(rsim, device, dev_world, states[perm].copy(), [c[perm].copy() for c in ctl])
    assert np.array_equal(base[0][perm, :TICK_WORDS], got[0][:, :TICK_WORDS])
    for a, b in zip(base[1], got[1]):
        assert np.array_equal(a[perm], b)


def test_contact_heavy_full_size(rsim, oracle, world, device, dev_world):
    note("contact-heavy 65,536 arenas x 1 decision, every word")
    n = 65536
    states = rsim.sample_arenas(world, rsim.SCEN_CONTACT, 61, n)
    ctl = [rsim.sample_controls(62, 0, n)]
    gs, gr, _ = gpu_decisions(rsim, device, dev_world, states, ctl)
    cs, cr = cpu_decisions(oracle, world, states, ctl)
    assert_same(gs, gr, cs, cr, "contact 65536")
    # a size-independent check: the digest of digests is stable across block sizes
    gs2, gr2, _ = gpu_decisions(rsim, device, dev_world, states, ctl, block=128)
    assert digest(gs2[:, :TICK_WORDS]) == digest(gs[:, :TICK_WORDS]) and np.array_equal(gr2[0], gr[0])


def test_edge_sizes(rsim, oracle, world, device, dev_world):
    note("ragged sizes 1, 63, 65 and an empty batch")
    for n in (1, 63, 65):
        states = rsim.sample_arenas(world, rsim.SCEN_MIXED, 70 + n, n)
        ctl = [rsim.sample_controls(71, 0, n)]
        gs, gr, _ = gpu_decisions(rsim, device, dev_world, states, ctl)
        cs, cr = cpu_decisions(oracle, world, states, ctl)
        assert_same(gs, gr, cs, cr, "n=%d" % n)
    d = device.alloc_states(0)
    c = device.alloc_words(16)
    r = device.alloc_words(1)
    device.sim8(d, dev_world, c, r)  # no arenas: nothing to launch
    device.sync()


def test_argument_checks(rsim, world, device, dev_world):
    note("argument checks")
    n = 10
    d = device.upload_states(rsim.sample_arenas(world, 0, 1, n))
    c = device.alloc_words(16 * n)
    r = device.alloc_words(n)
    with pytest.raises(rsim.RsimError):
        device.sim8(d, dev_world, c, r, block=0)
    with pytest.raises(rsim.RsimError):
        device.sim8(d, dev_world, c, r, block=257)
    with pytest.raises(rsim.RsimError):
        device.sim8(d, dev_world, device.alloc_words(16 * n - 1), r)
    with pytest.raises(rsim.RsimError):
        device.sim8(d, dev_world, c, device.alloc_words(n - 1))
    with pytest.raises(rsim.RsimError):
        device.sim8(d, dev_world, c, r, trace=device.alloc_words(8))


def test_world_variants_seven_pairings(rsim, oracle, device):
    """Open play over all seven preset pairings in one launch (a table of seven worlds, one index per arena), every
    arena word-identical to the oracle run in its own world; an index past the table leaves the arena untouched."""
    note("seven pairings in one launch, 4096 arenas x 3 decisions")
    n = 4096
    worlds = np.stack([rsim.build_world(p, p) for p in range(7)])
    variant = np.random.default_rng(81).integers(0, 7, n).astype(np.uint32)
    states = np.concatenate([rsim.sample_arenas(worlds[variant[a]], rsim.SCEN_OPEN, 82, 1, first=a) for a in range(n)])
    dw = device.upload_worlds(worlds)
    dv = device.upload_words(variant)
    d = device.upload_states(states)
    r = device.alloc_words(n)
    cpu = states.copy()
    for k in range(3):
        ctl = rsim.sample_controls(83, k, n)
        device.sim8_variants(d, dw, dv, device.upload_controls(ctl), r)
        device.sync()
        got_r = device.read_words(r)
        for p in range(7):
            idx = np.nonzero(variant == p)[0]
            sub = np.ascontiguousarray(cpu[idx])
            res = oracle.sim8(sub, worlds[p], np.ascontiguousarray(ctl[idx]), threads=8)
            cpu[idx] = sub
            assert np.array_equal(got_r[idx], res), "pairing %d decision %d" % (p, k)
    got = device.download_states(d)
    fd = first_diff(got[:, :TICK_WORDS], cpu[:, :TICK_WORDS])
    assert fd is None, "arena %d word %d (pairing %d)" % (fd[0], fd[1], variant[fd[0]])
    # an out-of-table index: untouched state, flagged result
    bad = variant.copy()
    bad[5] = 7
    before = device.download_states(d)
    device.sim8_variants(d, dw, device.upload_words(bad), device.upload_controls(rsim.sample_controls(84, 0, n)), r)
    device.sync()
    after = device.download_states(d)
    assert device.read_words(r)[5] == rsim.SIM8_BAD_VARIANT
    assert np.array_equal(after[5], before[5])


def test_world_variants_mutators(rsim, oracle, device):
    """Per-arena mutators (gravity, boost acceleration, ball radius) and presets from a table of 64 worlds, the scale
    workload's arena mix (kickoff, open play, contact 1:1:1)."""
    note("64 mutator worlds, 2048 arenas x 2 decisions")
    rng = np.random.default_rng(91)
    nw, n = 64, 2048
    worlds = np.stack([rsim.build_world(int(rng.integers(0, 7)), int(rng.integers(0, 7)), float(rng.uniform(-800, -500)),
                                        float(rng.uniform(0.8, 1.2)), float(rng.uniform(91.25, 100))) for _ in range(nw)])
    variant = rng.integers(0, nw, n).astype(np.uint32)
    states = np.concatenate([rsim.sample_arenas(worlds[variant[a]], rsim.SCEN_MIXED, 92, 1, first=a) for a in range(n)])
    dw = device.upload_worlds(worlds)
    dv = device.upload_words(variant)
    d = device.upload_states(states)
    r = device.alloc_words(n)
    cpu = states.copy()
    for k in range(2):
        ctl = rsim.sample_controls(93, k, n)
        device.sim8_variants(d, dw, dv, device.upload_controls(ctl), r)
        device.sync()
        got_r = device.read_words(r)
        for p in np.unique(variant):
            idx = np.nonzero(variant == p)[0]
            sub = np.ascontiguousarray(cpu[idx])
            res = oracle.sim8(sub, worlds[p], np.ascontiguousarray(ctl[idx]))
            cpu[idx] = sub
            assert np.array_equal(got_r[idx], res), "world %d decision %d" % (p, k)
    got = device.download_states(d)
    fd = first_diff(got[:, :TICK_WORDS], cpu[:, :TICK_WORDS])
    assert fd is None, "arena %d word %d" % fd


def test_run_equals_consecutive_sim8(rsim, oracle, world, device, dev_world):
    """rsim_run over five decisions in one launch: the same states and result words as five sim8 launches, and as the
    oracle."""
    note("rsim_run 5 decisions, 1500 arenas of the 1:1:1 mix")
    n, D = 1500, 5
    states = rsim.sample_arenas(world, rsim.SCEN_MIXED, 101, n)
    ctl = [rsim.sample_controls(102, d, n) for d in range(D)]
    d = device.upload_states(states)
    c = device.upload_words(np.concatenate(ctl))
    r = device.alloc_words(D * n)
    device.run(d, dev_world, c, r, D)
    device.sync()
    got, got_r = device.download_states(d), device.read_words(r).reshape(D, n)
    gs, gr, _ = gpu_decisions(rsim, device, dev_world, states, ctl)
    assert np.array_equal(got[:, :TICK_WORDS], gs[:, :TICK_WORDS])
    for k in range(D):
        assert np.array_equal(got_r[k], gr[k]), "decision %d" % k
    cs, cr = cpu_decisions(oracle, world, states, ctl)
    assert_same(got, list(got_r), cs, cr, "rsim_run")


def test_run_per_arena_worlds_and_sample_gather(rsim, oracle, device):
    """rsim_run with a per-arena world table (presets and mutators) equals the oracle in each arena's world, and a gather
    of a sample of arenas into a smaller buffer returns exactly those arenas."""
    note("rsim_run 4 decisions, 2000 arenas in 12 worlds; gather of a 37-arena sample")
    rng = np.random.default_rng(111)
    worlds = np.stack([rsim.build_world(p % 7, (p * 3) % 7, float(rng.uniform(-800, -500)), float(rng.uniform(0.8, 1.2)),
                                        float(rng.uniform(91.25, 100))) for p in range(12)])
    n, D = 2000, 4
    variant = rng.integers(0, 12, n).astype(np.uint32)
    states = rsim.sample_arenas_in(worlds, variant, rsim.SCEN_MIXED, 112)
    ctl = [rsim.sample_controls(113, d, n) for d in range(D)]
    dw, dv = device.upload_worlds(worlds), device.upload_words(variant)
    d = device.upload_states(states)
    r = device.alloc_words(D * n)
    device.run(d, dw, device.upload_words(np.concatenate(ctl)), r, D, variant=dv)
    sample = np.sort(rng.choice(n, 37, replace=False)).astype(np.uint32)
    g = device.alloc_states(sample.size)
    d_sample = device.upload_words(sample)  # held until the gather has run
    device.gather_states(d, d_sample, g)
    device.sync()
    got, got_r = device.download_states(d), device.read_words(r).reshape(D, n)
    cpu = states.copy()
    for k in range(D):
        assert np.array_equal(oracle.sim8_in(cpu, worlds, variant, ctl[k], threads=4), got_r[k]), "decision %d" % k
    fd = first_diff(got[:, :TICK_WORDS], cpu[:, :TICK_WORDS])
    assert fd is None, "arena %d word %d" % fd
    assert np.array_equal(device.download_states(g), got[sample])


def test_pooled_walks_with_mixed_world_layouts(rsim, oracle, device):
    """Arenas of one wave in worlds with the same physics but different BVH node layouts: each pooled walk must read its
    own arena's world. Results equal the oracle's and an all-original run's."""
    note("rsim_run 3 decisions, 4096 contact arenas over two node layouts")
    a = rsim.build_world(0, 0)
    b = a
    for k in (0, 4, 8, 12, 14):
        b = oracle.relocate_mesh_nodes(b, k)
    worlds = np.stack([a, b])
    n, D = 4096, 3
    rng = np.random.default_rng(193)
    variant = rng.integers(0, 2, n).astype(np.uint32)
    states = rsim.sample_arenas(a, rsim.SCEN_CONTACT, 194, n)
    ctl = [rsim.sample_controls(195, d, n) for d in range(D)]
    dw, dv = device.upload_worlds(worlds), device.upload_words(variant)
    d_mixed, d_plain = device.upload_states(states), device.upload_states(states)
    c = device.upload_words(np.concatenate(ctl))
    r_mixed, r_plain = device.alloc_words(D * n), device.alloc_words(D * n)
    device.run(d_mixed, dw, c, r_mixed, D, variant=dv)
    device.run(d_plain, dw, c, r_plain, D)  # every arena in world 0
    device.sync()
    got, got_r = device.download_states(d_mixed), device.read_words(r_mixed).reshape(D, n)
    assert np.array_equal(got_r, device.read_words(r_plain).reshape(D, n))
    assert np.array_equal(got[:, :TICK_WORDS], device.download_states(d_plain)[:, :TICK_WORDS])
    cpu = states.copy()
    for k in range(D):
        assert np.array_equal(oracle.sim8_in(cpu, worlds, variant, ctl[k], threads=4), got_r[k]), "decision %d" % k
    fd = first_diff(got[:, :TICK_WORDS], cpu[:, :TICK_WORDS])
    assert fd is None, "arena %d word %d" % fd


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                                                    "libref.so")), reason="reference build (oracle/_ref) not available")
def test_against_the_reference_build(rsim, world, device, dev_world):
    """The GPU against the reference's own C++ build directly (phase.cpp compiled in place, threaded by
    oracle/ref_batch.cpp): 8,192 arenas of the contact and mixed workloads, 10 decisions, every result word and
    persistent state word."""
    import oracle as orc
    ref = orc.Reference()
    for scen in (rsim.SCEN_CONTACT, rsim.SCEN_MIXED):
        note("GPU vs the reference build, scenario %d, 8192 arenas x 10 decisions" % scen)
        states = rsim.sample_arenas(world, scen, 6060 + scen, 8192)
        controls = [rsim.sample_controls(6070 + scen, d, 8192) for d in range(10)]
        gs, gr, _ = gpu_decisions(rsim, device, dev_world, states, controls)
        cs = states.copy()
        cr = [ref.sim8(cs, world, c, threads=8) for c in controls]
        assert_same(gs, gr, cs, cr, "reference build, scenario %d" % scen)


def _reference_in(ref, cpu, worlds, variant, controls, threads=8):
    """One decision of the reference's own build over arenas in per-arena worlds: each world's arenas as one batch."""
    res = np.zeros(cpu.shape[0], np.uint32)
    for p in np.unique(variant):
        idx = np.nonzero(variant == p)[0]
        sub = np.ascontiguousarray(cpu[idx])
        res[idx] = ref.sim8(sub, worlds[p], np.ascontiguousarray(controls[idx]), threads=threads)
        cpu[idx] = sub
    return res


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                                                    "libref.so")), reason="reference build (oracle/_ref) not available")
def test_world_table_and_run_against_the_reference_build(rsim, device):
    """rsim_sim8_variants over a world table of all seven preset pairings with per-world mutators (gravity, boost
    acceleration, ball radius), and rsim_run (k_run, several decisions in one launch) over the same table, against the
    reference's own C++ build run world by world: every result word of every decision, every persistent word."""
    import oracle as orc
    ref = orc.Reference()
    rng = np.random.default_rng(6080)
    worlds = np.stack([rsim.build_world(p, p, float(rng.uniform(-800, -500)), float(rng.uniform(0.8, 1.2)),
                                        float(rng.uniform(91.25, 100))) for p in range(7) for _ in range(2)])
    n, D = 4096, 4
    variant = rng.integers(0, worlds.shape[0], n).astype(np.uint32)
    states = rsim.sample_arenas_in(worlds, variant, rsim.SCEN_MIXED, 6081)
    controls = [rsim.sample_controls(6082, d, n) for d in range(D)]
    dw, dv = device.upload_worlds(worlds), device.upload_words(variant)
    cs = states.copy()
    cr = [_reference_in(ref, cs, worlds, variant, c) for c in controls]
    note("sim8_variants vs the reference build, 14 worlds (7 pairings x mutators), %d arenas x %d decisions" % (n, D))
    d = device.upload_states(states)
    r = device.alloc_words(n)
    gr = []
    for c in controls:
        device.sim8_variants(d, dw, dv, device.upload_controls(c), r)
        device.sync()
        gr.append(device.read_words(r))
    assert_same(device.download_states(d), gr, cs, cr, "sim8_variants vs reference build")
    note("rsim_run (k_run) vs the reference build, the same table, %d decisions in one launch" % D)
    d = device.upload_states(states)
    rr = device.alloc_words(D * n)
    device.run(d, dw, device.upload_words(np.concatenate(controls)), rr, D, variant=dv)
    device.sync()
    assert_same(device.download_states(d), list(device.read_words(rr).reshape(D, n)), cs, cr, "k_run vs reference build")
