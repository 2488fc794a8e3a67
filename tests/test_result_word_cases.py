"""The result word's rare outcomes on crafted inputs (tests/forced_cases.py): the overflow flag raised by a pair list at
capacity and carried by a nonzero counter, the unimplemented flag of a world whose mesh table disagrees with its bodies, a
goal on each chosen tick 1..8 at both goals (with nothing written after it), and the auto-flip path of an upside-down car.
On the CPU the oracle is checked against the reference's own build and against the outcome each case is built to reach;
on the GPU the kernels are checked against both, word for word (reference: cuda/sim/sim.cu:78-93 and
arena_step.cuh:16-34 for the result word)."""
import os
import sys

import numpy as np
import pytest

import forced_cases as fc
from conftest import first_diff
from helpers import TICK_WORDS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HAVE_REF = os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref.so"))
OVERFLOW, UNIMPLEMENTED = 0x40000000, 0x80000000


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


def cases(rsim, world):
    """(name, states, world for the run, controls per decision)."""
    n = 1024
    out = [("overflow", fc.overflow_states(rsim, world, n, 17), world, [rsim.sample_controls(18, d, n) for d in range(3)]),
           ("sticky", fc.sticky_states(rsim, world, 256, 19), world, [rsim.sample_controls(20, 0, 256)]),
           ("unimplemented", rsim.sample_arenas(world, rsim.SCEN_CONTACT, 21, n), fc.unimplemented_world(world),
            [rsim.sample_controls(22, d, n) for d in range(2)])]
    st, _ = fc.goal_states(rsim, world, 7)
    out.append(("goal", st, world, [np.zeros((st.shape[0], 16), np.uint32)]))
    out.append(("autoflip", fc.upside_down_states(rsim, world, 23, 64), world,
                [fc.autoflip_controls(rsim, 64, 0), fc.autoflip_controls(rsim, 64, 1)]))
    return out


def cpu_run(cpu, states, world, controls):
    s = states.copy()
    return s, [cpu.sim8(s, world, c, threads=8) for c in controls]


def check_outcome(rsim, name, states, final, results):
    """The outcome each crafted case is built to reach."""
    from rsim import layout
    if name == "overflow":
        assert sum(int(((r & OVERFLOW) != 0).sum()) for r in results) > 100
        assert (layout.arena_view(final)["overflow"][:, 3] > 0).any()  # the pair cache's counter
    elif name == "sticky":
        assert ((results[0] & OVERFLOW) != 0).all()
    elif name == "unimplemented":
        assert all(((r & UNIMPLEMENTED) != 0).sum() > 100 for r in results)
    elif name == "goal":
        assert np.array_equal(results[0] & 0xFF, np.repeat(np.arange(1, 9, dtype=np.uint32), 16))
    elif name == "autoflip":
        assert (layout.arena_view(final)["cars"]["is_auto_flipping"][:, 0] == 1).all()


@pytest.mark.parametrize("name", ["overflow", "sticky", "unimplemented", "goal", "autoflip"])
def test_oracle_reaches_the_outcome(rsim, oracle, world, name):
    c = {x[0]: x for x in cases(rsim, world)}[name]
    final, res = cpu_run(oracle, c[1], c[2], c[3])
    check_outcome(rsim, name, c[1], final, res)


@pytest.mark.skipif(not HAVE_REF, reason="reference build (oracle/_ref) not available")
@pytest.mark.parametrize("name", ["overflow", "sticky", "unimplemented", "goal", "autoflip"])
def test_oracle_equals_the_reference_build(rsim, oracle, world, name):
    import oracle as orc
    c = {x[0]: x for x in cases(rsim, world)}[name]
    a, ra = cpu_run(oracle, c[1], c[2], c[3])
    b, rb = cpu_run(orc.Reference(), c[1], c[2], c[3])
    for k, (x, y) in enumerate(zip(ra, rb)):
        assert np.array_equal(x, y), "%s decision %d" % (name, k)
    fd = first_diff(a[:, :TICK_WORDS], b[:, :TICK_WORDS])
    assert fd is None, "%s: arena %d word %d" % ((name,) + fd)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["overflow", "sticky", "unimplemented", "goal", "autoflip"])
def test_gpu_equals_the_oracle_and_the_reference_build(rsim, oracle, world, device, name):
    c = {x[0]: x for x in cases(rsim, world)}[name]
    states, w, controls = c[1], c[2], c[3]
    n = states.shape[0]
    note("forced case %s: %d arenas x %d decisions" % (name, n, len(controls)))
    dw = device.upload_static(w)
    d = device.upload_states(states)
    r = device.alloc_words(n)
    gr = []
    for ctl in controls:
        device.sim8(d, dw, device.upload_controls(ctl), r)
        device.sync()
        gr.append(device.read_words(r))
    got = device.download_states(d)
    check_outcome(rsim, name, states, got, gr)
    refs = [("oracle", oracle)]
    if HAVE_REF:
        import oracle as orc
        refs.append(("reference build", orc.Reference()))
    for what, cpu in refs:
        s, res = cpu_run(cpu, states, w, controls)
        for k, (x, y) in enumerate(zip(gr, res)):
            assert np.array_equal(x, y), "%s vs %s: decision %d, arena %d" % (name, what, k, int(np.argmax(x != y)))
        fd = first_diff(got[:, :TICK_WORDS], s[:, :TICK_WORDS])
        assert fd is None, "%s vs %s: arena %d word %d" % ((name, what) + fd)


@pytest.mark.gpu
def test_goal_trace_stops_at_the_scoring_tick(rsim, oracle, world, device):
    """The goal case with a trace: slabs 1..k hold the oracle's states after each tick, later slabs stay zero."""
    states, want = fc.goal_states(rsim, world, 7)
    n = states.shape[0]
    note("goal on ticks 1..8 with a trace, %d arenas" % n)
    dw = device.upload_static(world)
    d = device.upload_states(states)
    r = device.alloc_words(n)
    t = device.alloc_words(8 * rsim.aosoa_words(n))
    ctl = np.zeros((n, 16), np.uint32)
    device.sim8(d, dw, device.upload_controls(ctl), r, 64, t)
    device.sync()
    assert np.array_equal(device.read_words(r) & 0xFF, want)
    slabs = device.read_words(t).reshape(8, -1)
    tr = np.stack([rsim.aosoa_unpack(slabs[k], n) for k in range(8)], 1)
    for a in range(n):
        s = states[a].copy()
        ot = np.zeros((8, rsim.ARENA_WORDS), np.uint32)
        _, scored = oracle.step_ticks(s, world, ctl[a], 8, ot)
        assert scored == want[a]
        for k in range(scored):
            assert np.array_equal(tr[a, k, :TICK_WORDS], ot[k, :TICK_WORDS]), "arena %d tick %d" % (a, k + 1)
        assert not tr[a, scored:].any(), "arena %d: a tick after the goal was written" % a
