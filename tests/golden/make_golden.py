"""Writes the golden vectors of the decision path from the reference's own C++ build (oracle/_ref/libref.so, compiled from
the reference sources by `make -C oracle ref`). Run in the development container, where the reference tree exists:

    python tests/golden/make_golden.py

Inputs come from the synthetic-workload builder (world + seeded scenarios + seeded controls); their digests are stored so a
test can tell a changed builder from a changed tick. Outputs are the reference build's:
  golden_decisions.npz   per scenario: result words and TICK_WORDS digests after each decision, the final full states of
                         the first arenas, and the input digests
  golden_ticks.npz       per-tick trace digests (rscuda_step_ticks with a trace) and per-phase digests of one tick
  golden_libm.npz        digests of the correctly rounded sin/cos/tan over strided bit patterns and of atan2f/powf/
                         atan2_64/hypot64 on seeded inputs (the reference's lf_libm_* test shim)
"""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import oracle as orc  # noqa: E402

SEED_ARENAS = 20240601
SEED_CONTROLS = 7771
SCENARIOS = (0, 1, 2, 3, 4)
N_ARENAS = 24
N_DECISIONS = 6
N_FULL = 3  # full final states kept per scenario
WORLD_PRESETS = ((0, 0), (6, 2))  # Octane vs Octane; Psyclops (three wheels) vs Plank


def digest(words):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(words).tobytes()).digest()[:8], np.uint64)[0]


def decisions(ref, builder):
    out = {}
    for p0, p1 in WORLD_PRESETS:
        world = builder.build_world(p0, p1)
        out["world_digest_%d_%d" % (p0, p1)] = np.uint64(digest(world))
        for sc in SCENARIOS:
            if (p0, p1) != (0, 0) and sc not in (1, 2):
                continue
            tag = "p%d%d_s%d" % (p0, p1, sc)
            states = builder.sample_arenas(world, sc, SEED_ARENAS + sc, N_ARENAS)
            out[tag + "_input_digest"] = np.array([digest(s) for s in states], np.uint64)
            res = np.zeros((N_DECISIONS, N_ARENAS), np.uint32)
            dig = np.zeros((N_DECISIONS, N_ARENAS), np.uint64)
            ctl_dig = np.zeros(N_DECISIONS, np.uint64)
            for d in range(N_DECISIONS):
                ctl = builder.sample_controls(SEED_CONTROLS, d, N_ARENAS)
                ctl_dig[d] = digest(ctl)
                for a in range(N_ARENAS):
                    st, sc_tick = ref.step_ticks(states[a], world, ctl[a], 8)
                    res[d, a] = orc.result_word(st, sc_tick, states[a])
                    dig[d, a] = digest(states[a][: orc.TICK_WORDS])
                    if sc_tick:  # the arena restarts from its start state, as an RL loop resets it
                        states[a] = builder.sample_arenas(world, sc, SEED_ARENAS + sc, 1, first=a)[0]
            out[tag + "_result"] = res
            out[tag + "_digest"] = dig
            out[tag + "_controls_digest"] = ctl_dig
            out[tag + "_final"] = states[:N_FULL].copy()
    return out


def ticks(ref, builder):
    out = {}
    world = builder.build_world(0, 0)
    for sc in SCENARIOS:
        states = builder.sample_arenas(world, sc, SEED_ARENAS + 100 + sc, 8)
        ctl = builder.sample_controls(SEED_CONTROLS + 100, 0, 8)
        trace_dig = np.zeros((8, 8), np.uint64)
        scored = np.zeros(8, np.uint32)
        for a in range(8):
            trace = np.zeros((8, orc.ARENA_WORDS), np.uint32)
            st, sc_tick = ref.step_ticks(states[a], world, ctl[a], 8, trace)
            scored[a] = sc_tick
            for t in range(8):
                trace_dig[a, t] = digest(trace[t][: orc.TICK_WORDS]) if (sc_tick == 0 or t < sc_tick) else 0
        out["s%d_trace_digest" % sc] = trace_dig
        out["s%d_trace_scored" % sc] = scored
        # per phase: one tick of each arena, phase by phase, full-state digests (controls set as the decision loop does)
        states = builder.sample_arenas(world, sc, SEED_ARENAS + 200 + sc, 8)
        ph = np.zeros((8, 10), np.uint64)
        status = np.zeros((8, 10), np.int32)
        for a in range(8):
            s = states[a]
            builder.set_controls(s, ctl[a])
            for k in range(10):
                status[a, k] = ref.phase(k, s, world)
                ph[a, k] = digest(s)
        out["s%d_phase_digest" % sc] = ph
        out["s%d_phase_status" % sc] = status
    return out


def libm_inputs():
    """The libm check inputs: strided bit patterns for the unary functions, seeded pairs for the binary ones."""
    rng = np.random.default_rng(99)
    unary = np.arange(0, 1 << 32, 4093, dtype=np.uint64).astype(np.uint32)  # ~1.05 M patterns over the whole range
    near = (np.float32(np.pi / 4) * np.arange(-2000, 2000, dtype=np.float32)).view(np.uint32)  # near multiples of pi/4
    unary = np.concatenate([unary, near, near + 1, near - 1])
    n = 1 << 20
    ya = rng.standard_normal(n).astype(np.float32) * np.float32(10.0) ** rng.integers(-8, 8, n).astype(np.float32)
    xa = rng.standard_normal(n).astype(np.float32) * np.float32(10.0) ** rng.integers(-8, 8, n).astype(np.float32)
    px = np.abs(rng.standard_normal(n).astype(np.float32)) * np.float32(2.0) ** rng.integers(-20, 20, n).astype(np.float32)
    py = (rng.standard_normal(n) * 8).astype(np.float32)
    py[: n // 8] = np.round(py[: n // 8])  # integer exponents
    px[: n // 16] = np.float32(0.97)  # the ball's damping base
    py[: n // 16] = np.float32(1.0 / 120.0)
    specials = np.array([0.0, -0.0, 1.0, -1.0, 0.5, 2.0, np.inf, -np.inf, np.nan, 1e-45, -1e-45, 3.4028235e38, 1.1754944e-38,
                         0.97, 3.0, -3.0, 1.0 / 120.0], np.float32)
    sa, sb = np.meshgrid(specials, specials)
    a64 = rng.standard_normal(n) * 10.0 ** rng.integers(-5, 5, n)
    b64 = rng.standard_normal(n) * 10.0 ** rng.integers(-5, 5, n)
    a64[: n // 4] = a64[: n // 4].astype(np.float32)  # the auto-flip path passes floats widened to double
    b64[: n // 4] = b64[: n // 4].astype(np.float32)
    return dict(unary=unary, atan2_y=np.concatenate([ya, sa.ravel()]), atan2_x=np.concatenate([xa, sb.ravel()]),
                pow_x=np.concatenate([px, sa.ravel()]), pow_y=np.concatenate([py, sb.ravel()]), f64_a=a64, f64_b=b64)


def libm(reflm):
    inp = libm_inputs()
    out = {}
    for fn, name in enumerate(("sin", "cos", "tan")):
        out[name + "_digest"] = np.uint64(digest(reflm.unary(fn, inp["unary"])))
    out["atan2f_digest"] = np.uint64(digest(reflm.binary(0, inp["atan2_y"], inp["atan2_x"])))
    out["powf_digest"] = np.uint64(digest(reflm.binary(1, inp["pow_x"], inp["pow_y"])))
    out["atan2_64_digest"] = np.uint64(digest(reflm.f64(0, inp["f64_a"], inp["f64_b"])))
    out["hypot64_digest"] = np.uint64(digest(reflm.f64(1, inp["f64_a"], inp["f64_b"])))
    # a small explicit sample, for reports when a digest differs
    out["sample_unary"] = inp["unary"][::4096].copy()
    out["sample_sin"] = reflm.unary(0, out["sample_unary"])
    out["sample_atan2_64"] = reflm.f64(0, inp["f64_a"][:512], inp["f64_b"][:512])
    return out


def main():
    orc.build(ref=True)
    ref = orc.Reference()
    builder = orc.Oracle()  # only its builder is used here: the expected outputs all come from the reference build
    d = decisions(ref, builder)
    np.savez_compressed(os.path.join(HERE, "golden_decisions.npz"), **d)
    t = ticks(ref, builder)
    np.savez_compressed(os.path.join(HERE, "golden_ticks.npz"), **t)
    m = libm(orc.RefLibm())
    np.savez_compressed(os.path.join(HERE, "golden_libm.npz"), **m)
    for f in ("golden_decisions.npz", "golden_ticks.npz", "golden_libm.npz"):
        print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
