"""The device API for user kernels (include/rsim_device.h) through its example, rocketsim-cuda_amd/examples/observe.hip:
its fp16 conversion against numpy's (round to nearest even; NaN -> 0x7FFF by the reference's convention), and on the GPU
a user kernel that steps and observes against the oracle, rsim_sim8 and a host computation of the observation."""
import ctypes
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "..", "rocketsim-cuda_amd", "rsim", "libobserve.so")


@pytest.fixture(scope="module")
def obs_lib(rsim):
    if not os.path.exists(LIB):
        pytest.skip("libobserve.so not built")
    L = ctypes.CDLL(LIB)
    L.observe_f2h_host.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p]
    L.observe_step.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_uint32] + [ctypes.c_void_p] * 3
    L.observe_step.restype = ctypes.c_int
    L.observe_layout.argtypes = [ctypes.c_void_p]
    return L


def f2h(lib, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.size, np.uint16)
    lib.observe_f2h_host(x.ctypes.data_as(ctypes.c_void_p), x.size, out.ctypes.data_as(ctypes.c_void_p))
    return out


def test_f2h_matches_round_to_nearest_even(obs_lib):
    rng = np.random.default_rng(1)
    bits = np.concatenate([rng.integers(0, 2**32, 2_000_000, dtype=np.uint64).astype(np.uint32),
                           # every float near the half-precision boundaries: subnormal ties, the normal range edges, overflow
                           np.arange(0x33000000 - 64, 0x33000000 + 64, dtype=np.uint32),
                           np.arange(0x387FC000, 0x38802000, 7, dtype=np.uint32),
                           np.arange(0x477FE000, 0x47800100, dtype=np.uint32),
                           np.array([0, 0x80000000, 0x7F800000, 0xFF800000, 0x00000001, 0x3F800000, 0x33800000], np.uint32)])
    bits = np.concatenate([bits, bits | np.uint32(0x80000000)])
    x = bits.view(np.float32)
    got = f2h(obs_lib, x)
    nan = np.isnan(x)
    with np.errstate(over="ignore"):
        want = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got[~nan], want[~nan])
    assert (got[nan] == 0x7FFF).all()


@pytest.mark.gpu
def test_user_kernel_steps_like_sim8_and_observes(rsim, oracle, obs_lib, world, device, dev_world):
    print("[case] example user kernel (step + observe) vs the oracle and rsim_sim8, 2048 arenas", flush=True)
    n = 2048
    states = rsim.sample_arenas(world, rsim.SCEN_MIXED, 181, n)
    host_ctl = rsim.sample_controls(182, 0, n)
    ctl = device.upload_controls(host_ctl)
    a, b = device.upload_states(states), device.upload_states(states)
    ra, rb = device.alloc_words(n), device.alloc_words(n)
    nobs = obs_lib.observe_words()
    obs = device.alloc_words((n * nobs + 1) // 2)
    rew = device.alloc_words(n)
    device.sim8(a, dev_world, ctl, ra)
    assert obs_lib.observe_step(device.stream, b.ptr, dev_world.ptr, ctl.ptr, n, rb.ptr, obs.ptr, rew.ptr) == 0
    device.sync()
    sa, sb = device.download_states(a), device.download_states(b)
    # the user kernel's stepping against the oracle (the CPU build of the tick), then against rsim_sim8
    cpu = states.copy()
    cpu_res = oracle.sim8(cpu, world, host_ctl, threads=8)
    assert np.array_equal(device.read_words(rb), cpu_res)
    assert np.array_equal(sb[:, :rsim.TICK_WORDS], cpu[:, :rsim.TICK_WORDS])
    assert np.array_equal(device.read_words(ra), device.read_words(rb))
    assert np.array_equal(sa[:, :rsim.TICK_WORDS], sb[:, :rsim.TICK_WORDS])
    got = device.read_words(obs).view(np.uint16)[:n * nobs].reshape(n, nobs)
    off = np.zeros(10, np.uint32)
    obs_lib.observe_layout(off.ctypes.data_as(ctypes.c_void_p))
    f = sb.view(np.float32)
    want = [f[:, off[0]:off[0] + 3] / np.float32(4096), f[:, off[1]:off[1] + 3] / np.float32(6000)]
    for c in range(2):
        p, v, bo, g = off[2 + 4 * c:6 + 4 * c]
        want += [f[:, p:p + 3] / np.float32(4096), f[:, v:v + 3] / np.float32(2300), f[:, bo:bo + 1] / np.float32(100),
                 (sb[:, g:g + 1] != 0).astype(np.float32)]
    want = np.concatenate(want, axis=1).astype(np.float32)
    assert np.array_equal(got, f2h(obs_lib, want.ravel()).reshape(n, nobs))
    y0 = states.view(np.float32)[:, off[0] + 1]
    r = device.read_words(rew).view(np.float32)
    assert np.array_equal(r, (f[:, off[0] + 1] - y0) / np.float32(100))
