"""The CPU oracle against the golden vectors written from the reference's own C++ build (tests/golden/make_golden.py):
decision results and state digests per scenario and preset pairing, per-tick traces, per-phase states, and the correctly
rounded functions. These pin the oracle; the GPU tests then compare the kernels with it."""
import os

import numpy as np
import pytest

import make_golden as mg
import oracle as orc
from helpers import TICK_WORDS, digest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def gd():
    return np.load(os.path.join(GOLD, "golden_decisions.npz"))


@pytest.fixture(scope="module")
def gt():
    return np.load(os.path.join(GOLD, "golden_ticks.npz"))


@pytest.fixture(scope="module")
def gl():
    return np.load(os.path.join(GOLD, "golden_libm.npz"))


def test_world_digests(oracle, gd):
    for p0, p1 in mg.WORLD_PRESETS:
        assert digest(oracle.build_world(p0, p1)) == gd["world_digest_%d_%d" % (p0, p1)], "builder changed"


def cases():
    out = []
    for p0, p1 in mg.WORLD_PRESETS:
        for sc in mg.SCENARIOS:
            if (p0, p1) != (0, 0) and sc not in (1, 2):
                continue
            out.append((p0, p1, sc))
    return out


@pytest.mark.parametrize("p0,p1,sc", cases())
def test_decisions(oracle, gd, p0, p1, sc):
    tag = "p%d%d_s%d" % (p0, p1, sc)
    world = oracle.build_world(p0, p1)
    states = oracle.sample_arenas(world, sc, mg.SEED_ARENAS + sc, mg.N_ARENAS)
    assert np.array_equal(np.array([digest(s) for s in states], np.uint64), gd[tag + "_input_digest"])
    for d in range(mg.N_DECISIONS):
        ctl = oracle.sample_controls(mg.SEED_CONTROLS, d, mg.N_ARENAS)
        assert digest(ctl) == gd[tag + "_controls_digest"][d]
        for a in range(mg.N_ARENAS):
            st, sc_tick = oracle.step_ticks(states[a], world, ctl[a], 8)
            assert orc.result_word(st, sc_tick, states[a]) == gd[tag + "_result"][d, a], "%s decision %d arena %d" % (tag, d, a)
            assert digest(states[a][:TICK_WORDS]) == gd[tag + "_digest"][d, a], "%s decision %d arena %d" % (tag, d, a)
            if sc_tick:
                states[a] = oracle.sample_arenas(world, sc, mg.SEED_ARENAS + sc, 1, first=a)[0]
    fin = gd[tag + "_final"]
    assert np.array_equal(states[: fin.shape[0]], fin)


@pytest.mark.parametrize("sc", mg.SCENARIOS)
def test_trace_and_phases(oracle, gt, sc):
    world = oracle.build_world(0, 0)
    states = oracle.sample_arenas(world, sc, mg.SEED_ARENAS + 100 + sc, 8)
    ctl = oracle.sample_controls(mg.SEED_CONTROLS + 100, 0, 8)
    for a in range(8):
        tr = np.zeros((8, 2784), np.uint32)
        st, sc_tick = oracle.step_ticks(states[a], world, ctl[a], 8, tr)
        assert sc_tick == gt["s%d_trace_scored" % sc][a]
        for t in range(8):
            want = gt["s%d_trace_digest" % sc][a, t]
            got = digest(tr[t][:TICK_WORDS]) if (sc_tick == 0 or t < sc_tick) else 0
            assert got == want, "scenario %d arena %d tick %d" % (sc, a, t + 1)
    states = oracle.sample_arenas(world, sc, mg.SEED_ARENAS + 200 + sc, 8)
    for a in range(8):
        s = states[a]
        oracle.set_controls(s, ctl[a])
        for k in range(10):
            assert oracle.phase(k, s, world) == gt["s%d_phase_status" % sc][a, k]
            assert digest(s) == gt["s%d_phase_digest" % sc][a, k], "scenario %d arena %d phase %d" % (sc, a, k)


def test_libm_digests(oracle, gl):
    inp = mg.libm_inputs()
    for fn, name in enumerate(("sin", "cos", "tan")):
        assert digest(oracle.libm_unary_list(fn, inp["unary"])) == gl[name + "_digest"], name
    assert np.array_equal(oracle.libm_unary_list(0, gl["sample_unary"]), gl["sample_sin"])
    assert digest(oracle.libm_binary(0, inp["atan2_y"], inp["atan2_x"])) == gl["atan2f_digest"]
    assert digest(oracle.libm_binary(1, inp["pow_x"], inp["pow_y"])) == gl["powf_digest"]
    assert np.array_equal(oracle.libm_f64(0, inp["f64_a"][:512], inp["f64_b"][:512]), gl["sample_atan2_64"])
    assert digest(oracle.libm_f64(0, inp["f64_a"], inp["f64_b"])) == gl["atan2_64_digest"]
    assert digest(oracle.libm_f64(1, inp["f64_a"], inp["f64_b"])) == gl["hypot64_digest"]
