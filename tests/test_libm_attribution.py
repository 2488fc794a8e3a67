"""The libm attribution builds (the reference's phase_mixed.cpp / phase_platform.cpp and the attribution rule of its
tests/common/mod.rs:458-503), on the CPU: the oracle compiled with a per-function choice between the correctly rounded
functions and the platform C library.

Pinned by: mask 0 is the oracle itself; each mask bit routes exactly its function to the C library (checked call by
call against libm.so.6); where the reference tree is present, the reference's own mixed and platform builds, compiled
in place, give the same states as ours for every mask, tick and phase. Without V3 the attribution is exercised with the
platform build as the target (V3 on this platform computes with that library): every difference is charged to a libm
function, and a planted non-libm difference is charged to `other`.
"""
import ctypes
import ctypes.util
import os

import numpy as np
import pytest

import oracle as orc
from attribution import Attribution, attribute_step

MASKS = [0] + [1 << f for f in range(len(orc.LIBM_FUNCS))] + [orc.PLATFORM_ALL]


@pytest.fixture(scope="module")
def mixed():
    if not os.path.exists(os.path.join(orc.HERE, "liboracle_libm.so")):
        orc.build(ref=False)
    return orc.MixedOracle()


@pytest.fixture(scope="module")
def open_states(rsim, world):
    return rsim.sample_arenas(world, rsim.SCEN_OPEN, 3, 256), rsim.sample_controls(3, 0, 256)


def _tick(oracle, ctl):
    def step(m, s, w, mask):
        oracle.set_controls(s, ctl)
        m.step_tick(s, w, mask)
    return step


def test_mask_zero_is_the_oracle(oracle, mixed, world, open_states):
    st, ctl = open_states
    for i in range(0, 256, 8):
        a, b = st[i].copy(), st[i].copy()
        for _ in range(8):
            oracle.set_controls(a, ctl[i])
            oracle.step_tick(a, world)
            oracle.set_controls(b, ctl[i])
            mixed.step_tick(b, world, 0)
        assert np.array_equal(a, b), i


def test_each_bit_routes_one_function(mixed):
    libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
    for name in ("sinf", "cosf", "tanf"):
        getattr(libm, name).argtypes, getattr(libm, name).restype = [ctypes.c_float], ctypes.c_float
    for name in ("atan2f", "powf"):
        getattr(libm, name).argtypes, getattr(libm, name).restype = [ctypes.c_float, ctypes.c_float], ctypes.c_float
    for name in ("atan2", "hypot"):
        getattr(libm, name).argtypes, getattr(libm, name).restype = [ctypes.c_double, ctypes.c_double], ctypes.c_double
    plat = [lambda x, y: libm.sinf(x), lambda x, y: libm.cosf(x), lambda x, y: libm.tanf(x), libm.atan2f, libm.powf,
            libm.atan2, libm.hypot]
    rng = np.random.default_rng(1)
    xs = rng.uniform(-40, 40, 400).astype(np.float32)
    ys = rng.uniform(0.1, 3, 400).astype(np.float32)
    differs = [0] * 7
    for f in range(7):
        for x, y in zip(xs.tolist(), ys.tolist()):
            if f == 4:
                x = abs(x) / 10
            cr = mixed.libm(f, 0, x, y)
            assert mixed.libm(f, orc.PLATFORM_ALL & ~(1 << f), x, y) == cr  # the other bits leave f alone
            p = mixed.libm(f, 1 << f, x, y)
            assert p == plat[f](x, y), (orc.LIBM_FUNCS[f], x, y)
            differs[f] += p != cr
    # the C library is not correctly rounded everywhere (so the switch is observable), but it is close
    assert sum(differs) > 0, differs


def test_matches_the_reference_mixed_build(rsim, oracle, mixed, world, open_states):
    try:
        ref = orc.RefMixed()
    except RuntimeError as e:
        pytest.skip(str(e))
    st, ctl = open_states
    contact = rsim.sample_arenas(world, rsim.SCEN_CONTACT, 9, 32)
    cctl = rsim.sample_controls(9, 0, 32)
    cases = [(st[i], ctl[i]) for i in range(0, 256, 4)] + [(contact[i], cctl[i]) for i in range(32)]
    for mask in MASKS:
        for j, (s0, c) in enumerate(cases):
            a, b = s0.copy(), s0.copy()
            for _ in range(6):
                oracle.set_controls(a, c)
                mixed.step_tick(a, world, mask)
                oracle.set_controls(b, c)
                ref.step_tick(b, world, mask)
            assert np.array_equal(a, b), (mask, j)
            # per phase, one more tick
            oracle.set_controls(a, c)
            oracle.set_controls(b, c)
            for k in range(10):
                mixed.phase(k, a, world, mask)
                ref.phase(k, b, world, mask)
                assert np.array_equal(a, b), (mask, j, k)
    # the platform build is the all-bits mask
    for j, (s0, c) in enumerate(cases[:16]):
        a, b = s0.copy(), s0.copy()
        for _ in range(4):
            oracle.set_controls(a, c)
            mixed.step_tick(a, world, orc.PLATFORM_ALL)
            oracle.set_controls(b, c)
            ref.step_tick_platform(b, world)
        assert np.array_equal(a, b), j


def test_attribution_with_the_platform_build_as_target(oracle, mixed, world, open_states):
    st, ctl = open_states
    acc = Attribution()
    diffs = 0
    for i in range(256):
        s = st[i].copy()
        for t in range(8):
            dev, target = s.copy(), s.copy()
            oracle.set_controls(dev, ctl[i])
            oracle.step_tick(dev, world)          # the device computes with the correctly rounded functions
            oracle.set_controls(target, ctl[i])
            mixed.step_tick(target, world, orc.PLATFORM_ALL)  # V3 on this platform: the C library
            attribute_step(mixed, acc, "arena %d tick %d" % (i, t), s, world, dev, target, _tick(oracle, ctl[i]))
            diffs += int(not np.array_equal(dev, target))
            s = target  # follow the target, as the phase driver follows V3's states
    assert acc.ran == 256 * 8 and acc.states_diff == diffs > 0
    assert acc.other == 0, acc.examples
    assert acc.words > 0 and sum(acc.by_func) + acc.combo >= acc.words
    assert acc.by_func[orc.LIBM_FUNCS.index("sin")] > 0, acc.row()
    # a difference that is not the libm's is charged to `other`
    s = st[0].copy()
    dev, target = s.copy(), s.copy()
    for x in (dev, target):
        oracle.set_controls(x, ctl[0])
        oracle.step_tick(x, world)
    dev[100] ^= 1
    planted = Attribution()
    attribute_step(mixed, planted, "planted", s, world, dev, target, _tick(oracle, ctl[0]))
    assert (planted.words, planted.other, sum(planted.by_func), planted.combo) == (1, 1, 0, 0)


def _phase_step(k):
    def step(m, s, w, mask):
        m.phase(k, s, w, mask)
    return step


@pytest.mark.gpu
def test_device_attribution_against_the_platform_build(rsim, oracle, mixed, world, device, dev_world):
    """The attribution's device leg (crates/rocketsim-cuda/tests/common/mod.rs:458-541): the GPU's phases against the
    platform build (the stand-in for V3, which computes with this host's C library), phase by phase from the same input,
    following the platform build's states as the reference's phase driver follows V3's. Every word in which the GPU differs
    is charged to a libm function (or two together), none to `other`; a planted non-libm difference is."""
    print("[case] device phases vs the platform build, 192 arenas x 3 ticks x 10 phases", flush=True)
    st = np.concatenate([rsim.sample_arenas(world, rsim.SCEN_OPEN, 21, 128), rsim.sample_arenas(world, rsim.SCEN_CONTACT, 22, 64)])
    ctl = np.concatenate([rsim.sample_controls(21, 0, 128), rsim.sample_controls(22, 0, 64)])
    n = st.shape[0]
    acc = Attribution()
    diffs = 0
    cur = st.copy()
    for t in range(3):
        for a in range(n):
            oracle.set_controls(cur[a], ctl[a])
        for k in range(10):
            d = device.upload_states(cur)
            device.phase(k, d, dev_world)
            dev = device.download_states(d)
            target = cur.copy()
            for a in range(n):
                mixed.phase(k, target[a], world, orc.PLATFORM_ALL)
            for a in range(n):
                attribute_step(mixed, acc, "arena %d tick %d phase %d" % (a, t, k), cur[a], world, dev[a], target[a],
                               _phase_step(k))
                diffs += int(not np.array_equal(dev[a], target[a]))
            cur = target
    print("[attribution] " + str(acc.row()), flush=True)
    assert acc.ran == n * 3 * 10 and acc.states_diff == diffs > 0
    assert acc.other == 0, acc.examples
    assert acc.words > 0 and sum(acc.by_func) + acc.combo >= acc.words
    # a planted difference in the device's output is not the libm's
    d = device.upload_states(st[:1])
    device.phase(0, d, dev_world)
    dev = device.download_states(d)[0]
    target = st[0].copy()
    mixed.phase(0, target, world, orc.PLATFORM_ALL)
    dev[100] ^= 1
    planted = Attribution()
    attribute_step(mixed, planted, "planted", st[0], world, dev, target, _phase_step(0))
    assert planted.other >= 1
