"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle (liboracle.so) and, where built, the reference's own C++
build (_ref/libref.so). Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg use this module, and only as
the checker; the product (rocketsim-cuda_amd/rsim, librsim.so) never imports it.

Both libraries take states as uint32[n, 2784] arrays (the reference's ArenaState words), a StaticData as uint32[317895]
and controls as uint32[n, 16].
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libref.so")
REF_TREE = "/root/reference"

ARENA_WORDS = 2784
TICK_WORDS = 2567
STATIC_WORDS = 1271580 // 4
_vp = ctypes.c_void_p
_u32, _u64 = ctypes.c_uint32, ctypes.c_uint64


def build(ref=None):
    """Compiles liboracle.so (and _ref/libref.so when the reference tree is present, or when ref=True)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so", "liboracle_libm.so"], check=True)
    if ref or (ref is None and os.path.isdir(REF_TREE)):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _p(a):
    return a.ctypes.data_as(_vp)


class Oracle:
    """The CPU oracle: the single-source tick compiled for the host (cpu_oracle.cpp)."""

    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise RuntimeError("liboracle.so is not built: make -C oracle")
        L = ctypes.CDLL(path)
        L.rsim_cpu_phase.argtypes = [ctypes.c_int, _vp, _vp]
        L.rsim_cpu_step_tick.argtypes = [_vp, _vp]
        L.rsim_cpu_step_ticks.argtypes = [_vp, _vp, _vp, _u32, ctypes.POINTER(_u32), _vp]
        L.rsim_cpu_sim8_aos.argtypes = [_vp, _vp, _vp, _vp, _u32]
        L.rsim_cpu_sim8_batch.argtypes = [_vp, _vp, _vp, _vp, _u32, _u32]
        L.rsim_cpu_sim8_batch_in.argtypes = [_vp, _vp, _u32, _vp, _vp, _vp, _u32, _u32]
        L.rsim_cpu_static_layout.argtypes = [_vp]
        L.rsim_cpu_libm_unary.argtypes = [ctypes.c_int, _u64, _u64, _vp]
        L.rsim_cpu_libm_unary_list.argtypes = [ctypes.c_int, _vp, _u64, _vp]
        L.rsim_cpu_libm_binary.argtypes = [ctypes.c_int, _vp, _vp, _u64, _vp]
        L.rsim_cpu_libm_f64.argtypes = [ctypes.c_int, _vp, _vp, _u64, _vp]
        L.rsim_build_world.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _vp]
        L.rsim_sample_arenas.argtypes = [_vp, ctypes.c_int, _u64, _u32, _u32, _vp]
        L.rsim_sample_controls.argtypes = [_u64, _u32, _u32, _u32, _vp]
        L.rsim_cpu_check_move_cells.argtypes = [_vp, _u64, _u64]
        L.rsim_cpu_check_move_cells.restype = _u64
        for f in ("rsim_cpu_arena_words", "rsim_cpu_tick_words", "rsim_cpu_static_words", "rsim_cpu_controls_word"):
            getattr(L, f).restype = _u64
        self.L = L
        self.controls_word = [int(L.rsim_cpu_controls_word(c)) for c in (0, 1)]

    def set_controls(self, state, controls):
        """Writes one arena's controls (uint32[16]) into the state, as the decision loop does before each tick."""
        for c in (0, 1):
            state[self.controls_word[c]:self.controls_word[c] + 8] = controls[8 * c:8 * c + 8]

    # -- the reference's lower C ABI (phase.cpp:8-17) --
    def phase(self, k, state, world):
        return self.L.rsim_cpu_phase(k, _p(state), _p(world))

    def step_tick(self, state, world):
        return self.L.rsim_cpu_step_tick(_p(state), _p(world))

    def step_ticks(self, state, world, controls, n, trace=None):
        """In place on one state (uint32[2784]); returns (status, scored_tick)."""
        sc = _u32(0)
        st = self.L.rsim_cpu_step_ticks(_p(state), _p(world), _p(controls), n, ctypes.byref(sc),
                                         _p(trace) if trace is not None else None)
        return st, sc.value

    def sim8(self, states, world, controls, threads=1):
        """One decision of every arena of states (uint32[n, 2784], in place); returns the result words."""
        n = states.shape[0]
        res = np.zeros(max(n, 1), np.uint32)
        assert states.flags.c_contiguous and controls.flags.c_contiguous
        self.L.rsim_cpu_sim8_batch(_p(states), _p(world), _p(controls), _p(res), n, threads)
        return res[:n]

    def sim8_in(self, states, worlds, variant, controls, threads=1):
        """sim8 with arena a in worlds[variant[a]] (worlds: uint32[W, static words]); returns the result words."""
        n = states.shape[0]
        res = np.zeros(max(n, 1), np.uint32)
        worlds = np.ascontiguousarray(worlds, np.uint32)
        variant = np.ascontiguousarray(variant, np.uint32)
        assert states.flags.c_contiguous and controls.flags.c_contiguous and variant.size >= n
        self.L.rsim_cpu_sim8_batch_in(_p(states), _p(worlds), worlds.reshape(len(worlds), -1).shape[0], _p(variant), _p(controls),
                                      _p(res), n, threads)
        return res[:n]

    def relocate_mesh_nodes(self, world, mesh):
        """A copy of `world` (u32 words) with mesh `mesh`'s BVH nodes moved to the unused end of the node array and its
        node_first updated: the same physics, a different node layout. None if the array has no room."""
        lay = np.zeros(8, np.uint64)
        self.L.rsim_cpu_static_layout(_p(lay))
        meshes, msize, nf_off, nc_off, nodes, nsize, nmesh, cap = (int(x) for x in lay)
        b = np.array(world, np.uint32).view(np.uint8).copy()
        rec = lambda k, off: meshes + k * msize + off
        u32 = lambda at: int(b[at:at + 4].view(np.uint32)[0])
        used = max(u32(rec(k, nf_off)) + u32(rec(k, nc_off)) for k in range(nmesh))
        first, count = u32(rec(mesh, nf_off)), u32(rec(mesh, nc_off))
        if count == 0 or used + count > cap:
            return None
        b[nodes + used * nsize:nodes + (used + count) * nsize] = b[nodes + first * nsize:nodes + (first + count) * nsize]
        b[rec(mesh, nf_off):rec(mesh, nf_off) + 4] = np.array([used], np.uint32).view(np.uint8)
        return b.view(np.uint32)

    def check_move_cells(self, world, seed, iters):
        """Mismatches of the one-pass cell move against remove-then-push on random lists (brute force)."""
        return int(self.L.rsim_cpu_check_move_cells(_p(world), seed, iters))

    # -- workloads (the same builder source as the product's) --
    def build_world(self, p0=0, p1=0, gravity_z=-650.0, boost_accel_scale=1.0, ball_radius=91.25):
        w = np.zeros(STATIC_WORDS, np.uint32)
        assert self.L.rsim_build_world(p0, p1, gravity_z, boost_accel_scale, ball_radius, _p(w)) == 0
        return w

    def sample_arenas(self, world, scenario, seed, n, first=0):
        s = np.zeros((n, ARENA_WORDS), np.uint32)
        assert self.L.rsim_sample_arenas(_p(world), scenario, seed, first, n, _p(s)) == 0
        return s

    def sample_controls(self, seed, decision, n, first=0):
        c = np.zeros((n, 16), np.uint32)
        assert self.L.rsim_sample_controls(seed, decision, first, n, _p(c)) == 0
        return c

    # -- correctly rounded functions --
    def libm_unary(self, fn, base, n):
        out = np.zeros(n, np.uint32)
        self.L.rsim_cpu_libm_unary(fn, base, n, _p(out))
        return out

    def libm_unary_list(self, fn, x_bits):
        x = np.ascontiguousarray(x_bits, np.uint32)
        out = np.zeros(x.size, np.uint32)
        self.L.rsim_cpu_libm_unary_list(fn, _p(x), x.size, _p(out))
        return out

    def libm_binary(self, op, a, b):
        a = np.ascontiguousarray(a, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        out = np.zeros(a.size, np.uint32)
        self.L.rsim_cpu_libm_binary(op, _p(a), _p(b), a.size, _p(out))
        return out

    def libm_f64(self, op, a, b):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        out = np.zeros(a.size, np.float64)
        self.L.rsim_cpu_libm_f64(op, _p(a), _p(b), a.size, _p(out))
        return out


class RefLibm:
    """The reference's correctly rounded functions over caller-chosen inputs (_ref/libref_libm.so)."""

    def __init__(self, path=os.path.join(HERE, "_ref", "libref_libm.so")):
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libref_libm.so is not built (make -C oracle ref)")
        L = ctypes.CDLL(path)
        L.refm_unary.argtypes = [ctypes.c_int, _vp, _u64, _vp]
        L.refm_binary.argtypes = [ctypes.c_int, _vp, _vp, _u64, _vp]
        L.refm_f64.argtypes = [ctypes.c_int, _vp, _vp, _u64, _vp]
        self.L = L

    def unary(self, fn, x_bits):
        x = np.ascontiguousarray(x_bits, np.uint32)
        out = np.zeros(x.size, np.uint32)
        self.L.refm_unary(fn, _p(x), x.size, _p(out))
        return out

    def binary(self, op, a, b):
        a = np.ascontiguousarray(a, np.float32).view(np.uint32)
        b = np.ascontiguousarray(b, np.float32).view(np.uint32)
        out = np.zeros(a.size, np.uint32)
        self.L.refm_binary(op, _p(a), _p(b), a.size, _p(out))
        return out

    def f64(self, op, a, b):
        a = np.ascontiguousarray(a, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        out = np.zeros(a.size, np.float64)
        self.L.refm_f64(op, _p(a), _p(b), a.size, _p(out))
        return out


class Reference:
    """The reference's own host build (crates/rocketsim-cuda/csrc/phase.cpp + libm.cpp, compiled in place)."""

    def __init__(self, path=REF_LIB):
        if not os.path.exists(path):
            raise RuntimeError("oracle/_ref/libref.so is not built (make -C oracle ref, needs the reference tree)")
        L = ctypes.CDLL(path)
        L.rscuda_phase.argtypes = [ctypes.c_int, _vp, _vp]
        L.rscuda_step_tick.argtypes = [_vp, _vp]
        L.rscuda_step_ticks.argtypes = [_vp, _vp, _vp, _u32, ctypes.POINTER(_u32), _vp]
        L.refb_sim8_batch.argtypes = [_vp, _vp, _vp, _vp, _u32, _u32]
        self.L = L

    def phase(self, k, state, world):
        return self.L.rscuda_phase(k, _p(state), _p(world))

    def step_tick(self, state, world):
        return self.L.rscuda_step_tick(_p(state), _p(world))

    def step_ticks(self, state, world, controls, n, trace=None):
        sc = _u32(0)
        st = self.L.rscuda_step_ticks(_p(state), _p(world), _p(controls), n, ctypes.byref(sc),
                                      _p(trace) if trace is not None else None)
        return st, sc.value

    def sim8(self, states, world, controls, threads=1):
        """One decision of every arena of states (uint32[n, 2784], in place) by the reference's own CPU build, over
        `threads` host threads; returns the sim8 result words."""
        n = states.shape[0]
        res = np.zeros(max(n, 1), np.uint32)
        assert states.flags.c_contiguous and controls.flags.c_contiguous and controls.shape[0] >= n
        if self.L.refb_sim8_batch(_p(states), _p(world), _p(controls), _p(res), n, threads) != 0:
            raise RuntimeError("the reference's ArenaState is not 2,784 words")
        return res[:n]


LIBM_FUNCS = ("sin", "cos", "tan", "atan2", "pow", "atan2_64", "hypot64")  # bit f of a libm mask
PLATFORM_ALL = (1 << len(LIBM_FUNCS)) - 1


class MixedOracle:
    """The oracle's libm attribution build (liboracle_libm.so): the tick with each transcendental call choosing the
    correctly rounded function or the platform C library by a bit mask (bit f of LIBM_FUNCS set: platform)."""

    def __init__(self, path=os.path.join(HERE, "liboracle_libm.so")):
        if not os.path.exists(path):
            raise RuntimeError("liboracle_libm.so is not built: make -C oracle")
        L = ctypes.CDLL(path)
        L.rsim_cpu_phase_mixed.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_uint]
        L.rsim_cpu_step_tick_mixed.argtypes = [_vp, _vp, ctypes.c_uint]
        L.rsim_cpu_step_ticks_mixed.argtypes = [_vp, _vp, _vp, _u32, ctypes.POINTER(_u32), _vp, ctypes.c_uint]
        L.rsim_cpu_phase_platform.argtypes = [ctypes.c_int, _vp, _vp]
        L.rsim_cpu_step_tick_platform.argtypes = [_vp, _vp]
        L.rsim_cpu_libm_mixed.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_double, ctypes.c_double]
        L.rsim_cpu_libm_mixed.restype = ctypes.c_double
        self.L = L

    def phase(self, k, state, world, mask):
        return self.L.rsim_cpu_phase_mixed(k, _p(state), _p(world), mask)

    def step_tick(self, state, world, mask):
        return self.L.rsim_cpu_step_tick_mixed(_p(state), _p(world), mask)

    def step_ticks(self, state, world, controls, n, mask, trace=None):
        sc = _u32(0)
        st = self.L.rsim_cpu_step_ticks_mixed(_p(state), _p(world), _p(controls), n, ctypes.byref(sc),
                                              _p(trace) if trace is not None else None, mask)
        return st, sc.value

    def libm(self, fn, mask, x, y=0.0):
        return self.L.rsim_cpu_libm_mixed(fn, mask, x, y)


class RefMixed:
    """The reference's own mixed and platform builds (phase_mixed.cpp, phase_platform.cpp, compiled in place)."""

    def __init__(self, mixed=os.path.join(HERE, "_ref", "libref_mixed.so"),
                 platform=os.path.join(HERE, "_ref", "libref_platform.so")):
        for p in (mixed, platform):
            if not os.path.exists(p):
                raise RuntimeError("%s is not built (make -C oracle ref, needs the reference tree)" % p)
        M, P = ctypes.CDLL(mixed), ctypes.CDLL(platform)
        M.rscuda_phase_mixed.argtypes = [ctypes.c_int, _vp, _vp, ctypes.c_uint]
        M.rscuda_step_tick_mixed.argtypes = [_vp, _vp, ctypes.c_uint]
        P.rscuda_phase_platform.argtypes = [ctypes.c_int, _vp, _vp]
        P.rscuda_step_tick_platform.argtypes = [_vp, _vp]
        self.M, self.P = M, P

    def phase(self, k, state, world, mask):
        return self.M.rscuda_phase_mixed(k, _p(state), _p(world), mask)

    def step_tick(self, state, world, mask):
        return self.M.rscuda_step_tick_mixed(_p(state), _p(world), mask)

    def step_tick_platform(self, state, world):
        return self.P.rscuda_step_tick_platform(_p(state), _p(world))


def result_word(status, scored, state):
    """The sim8 result word from a CPU decision: scored tick | unimplemented | overflow (arena_step.cuh:16-25)."""
    ovf = state[_OVERFLOW_WORD:_OVERFLOW_WORD + 6].any()
    return (scored & 0xFF) | (0x80000000 if status != 0 else 0) | (0x40000000 if ovf else 0)


_OVERFLOW_WORD = (10264 - 6 * 4) // 4  # ArenaState.overflow[6] sits just before events (byte 10264)
