"""rsim: Python mirror of the reference crate's device API over librsim.so's C ABI.

`SimDevice` follows crates/rocketsim-cuda/src/device.rs:30-172 method by method (new, upload_static, upload_states,
alloc_states, download_words, download_states, identity, identity_tick, phase, step_tick, upload_controls, sim8): the
same arguments, the same in-place update of the state buffer, the same result words. States, StaticData and controls are
numpy arrays in the reference's word format (uint32; an ArenaState is 2,784 words, a StaticData 317,895, one arena's
controls 16 = two cars x (throttle, steer, pitch, yaw, roll as f32 bits, jump, boost, handbrake as 0/1)).

The GPU path is librsim.so (HIP kernels for gfx950). There is no CPU fallback: if the library is missing or no GPU is
present, SimDevice raises.
"""
import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librsim.so")

ARENA_WORDS = 2784
TICK_WORDS = 2567
STATIC_WORDS = 1271580 // 4
CONTROL_WORDS = 16  # per arena: 2 cars x 8 words
TICKS_PER_DECISION = 8
SIM8_SCORED_MASK = 0xFF
SIM8_UNIMPLEMENTED = 0x80000000
SIM8_OVERFLOW = 0x40000000
SIM8_BAD_VARIANT = 0x20000000
MAX_WORLDS = 1024
SIM8_BLOCK = 64
MAX_BLOCK = 256
AOSOA_WIDTH = 64
NUM_PHASES = 10
PHASE_OK, PHASE_UNIMPLEMENTED, PHASE_BAD_INDEX = 0, -1, -2

NUM_PRESETS = 7
PRESETS = ("octane", "dominus", "plank", "breakout", "hybrid", "merc", "psyclops")
SCEN_KICKOFF, SCEN_OPEN, SCEN_PINCH, SCEN_HEADON, SCEN_GOALFRAME, SCEN_CONTACT, SCEN_MIXED, SCEN_OPEN_CONTACT = range(8)

_lib = None
_vp = ctypes.c_void_p


class RsimError(RuntimeError):
    pass


def lib():
    """librsim.so, loaded once. Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RsimError("librsim.so is not built (%s): run __graft_entry__.build() or python -m rsim.build" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        u32, u64, i32, f32 = ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int32, ctypes.c_float
        sig = {
            "rsim_arena_words": ([], u64), "rsim_tick_words": ([], u64), "rsim_static_words": ([], u64),
            "rsim_aosoa_words": ([u32], u64), "rsim_aosoa_index": ([u32, u32], u64),
            "rsim_aosoa_pack": ([_vp, u32, _vp], ctypes.c_int), "rsim_aosoa_unpack": ([_vp, u32, _vp], ctypes.c_int),
            "rsim_open": ([ctypes.c_int, ctypes.POINTER(_vp)], ctypes.c_int), "rsim_close": ([_vp], ctypes.c_int),
            "rsim_stream": ([_vp], _vp),
            "rsim_upload_world": ([_vp, _vp, ctypes.POINTER(_vp)], ctypes.c_int), "rsim_world_free": ([_vp], ctypes.c_int),
            "rsim_upload_worlds": ([_vp, _vp, u32, ctypes.POINTER(_vp)], ctypes.c_int), "rsim_world_count": ([_vp], u32),
            "rsim_sim8_variants": ([_vp, _vp, _vp, _vp, _vp, _vp, u32, _vp], ctypes.c_int),
            "rsim_run": ([_vp, _vp, _vp, _vp, _vp, _vp, u32, u32], ctypes.c_int),
            "rsim_run_queued": ([_vp, _vp, _vp, _vp, _vp, _vp, u32, u32, _vp, u32], ctypes.c_int),
            "rsim_run_queue_capacity": ([_vp, u32], u32),
            "rsim_sim8_split": ([_vp, _vp, _vp, _vp, _vp, _vp, u32, u32], ctypes.c_int),
            "rsim_cost_block_order_work_words": ([u32], u64),
            "rsim_cost_block_order": ([_vp, _vp, u32, _vp, _vp], ctypes.c_int),
            "rsim_sim8_timed": ([_vp, _vp, _vp, _vp, _vp, _vp, u32, _vp, _vp], ctypes.c_int),
            "rsim_cost_features": ([_vp, _vp, _vp], ctypes.c_int),
            "rsim_cost_predict": ([_vp, _vp, _vp, u32, _vp], ctypes.c_int),
            "rsim_cost_gradients": ([_vp, _vp, _vp, _vp, _vp, u32, _vp, _vp, _vp], ctypes.c_int),
            "rsim_cost_adam": ([_vp, _vp, _vp, _vp, _vp, f32, u32], ctypes.c_int),
            "rsim_set_permutation": ([_vp, _vp, u32, _vp], ctypes.c_int),
            "rsim_cost_order_work_words": ([u32], u64),
            "rsim_cost_order": ([_vp, _vp, u32, ctypes.c_int, _vp, _vp], ctypes.c_int),
            "rsim_cost_order_noalloc": ([_vp, _vp, u32, ctypes.c_int, _vp], ctypes.c_int),
            "rsim_capture_begin": ([_vp], ctypes.c_int), "rsim_capture_end": ([_vp, ctypes.POINTER(_vp)], ctypes.c_int),
            "rsim_graph_launch": ([_vp, _vp], ctypes.c_int), "rsim_graph_free": ([_vp], ctypes.c_int),
            "rsim_gather_states": ([_vp, _vp, _vp, _vp], ctypes.c_int),
            "rsim_permute_words": ([_vp, _vp, _vp, u32, u32, ctypes.c_int, _vp], ctypes.c_int),
            "rsim_upload_states": ([_vp, _vp, u32, ctypes.POINTER(_vp)], ctypes.c_int),
            "rsim_alloc_states": ([_vp, u32, ctypes.POINTER(_vp)], ctypes.c_int),
            "rsim_write_states": ([_vp, _vp, _vp], ctypes.c_int), "rsim_copy_states": ([_vp, _vp, _vp], ctypes.c_int),
            "rsim_states_free": ([_vp], ctypes.c_int), "rsim_states_count": ([_vp], u32),
            "rsim_states_device_ptr": ([_vp], _vp), "rsim_world_device_ptr": ([_vp], _vp),
            "rsim_download_words": ([_vp, _vp, _vp], ctypes.c_int), "rsim_download_states": ([_vp, _vp, _vp], ctypes.c_int),
            "rsim_alloc_words": ([_vp, u64, ctypes.POINTER(_vp)], ctypes.c_int),
            "rsim_upload_words": ([_vp, _vp, u64, ctypes.POINTER(_vp)], ctypes.c_int),
            "rsim_write_words": ([_vp, _vp, _vp, u64], ctypes.c_int), "rsim_read_words": ([_vp, _vp, _vp, u64], ctypes.c_int),
            "rsim_upload_controls": ([_vp, _vp, u32, ctypes.POINTER(_vp)], ctypes.c_int),
            "rsim_words_free": ([_vp], ctypes.c_int), "rsim_words_count": ([_vp], u64), "rsim_words_device_ptr": ([_vp], _vp),
            "rsim_words_view": ([_vp, u64, u64, ctypes.POINTER(_vp)], ctypes.c_int),
            "rsim_sim8": ([_vp, _vp, _vp, _vp, _vp, u32, _vp], ctypes.c_int),
            "rsim_identity": ([_vp, _vp, _vp], ctypes.c_int), "rsim_identity_tick": ([_vp, _vp, _vp], ctypes.c_int),
            "rsim_phase": ([_vp, ctypes.c_int, _vp, _vp, _vp], ctypes.c_int),
            "rsim_step_tick": ([_vp, _vp, _vp, _vp], ctypes.c_int),
            "rsim_timer_begin": ([_vp], ctypes.c_int), "rsim_timer_end": ([_vp, ctypes.POINTER(f32)], ctypes.c_int),
            "rsim_sync": ([_vp], ctypes.c_int), "rsim_last_error": ([], ctypes.c_char_p),
            "rsim_build_world": ([ctypes.c_int, ctypes.c_int, f32, f32, f32, _vp], ctypes.c_int),
            "rsim_sample_arenas": ([_vp, ctypes.c_int, u64, u32, u32, _vp], ctypes.c_int),
            "rsim_sample_controls": ([u64, u32, u32, u32, _vp], ctypes.c_int),
            "rsim_sample_arenas_in": ([_vp, u32, _vp, ctypes.c_int, u64, u32, u32, _vp], ctypes.c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        del i32
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise RsimError("%s failed (%d): %s" % (what, rc, lib().rsim_last_error().decode()))


def _ptr(a):
    return a.ctypes.data_as(_vp)


def _words(a, n_words=None, what="buffer"):
    a = np.ascontiguousarray(a)
    if a.dtype != np.uint32:
        a = a.view(np.uint32)
    if n_words is not None and a.size != n_words:
        raise RsimError("%s: %d words, expected %d" % (what, a.size, n_words))
    return a


# ---- layout ---------------------------------------------------------------------------------------------------------------------
def aosoa_words(n):
    return int(lib().rsim_aosoa_words(n))


def aosoa_index(a, w):
    return int(lib().rsim_aosoa_index(a, w))


def aosoa_pack(states):
    s = _words(states).reshape(-1, ARENA_WORDS)
    out = np.empty(aosoa_words(s.shape[0]), np.uint32)
    _check(lib().rsim_aosoa_pack(_ptr(s), s.shape[0], _ptr(out)), "rsim_aosoa_pack")
    return out


def aosoa_unpack(words, n):
    w = _words(words)
    out = np.empty((n, ARENA_WORDS), np.uint32)
    _check(lib().rsim_aosoa_unpack(_ptr(w), n, _ptr(out)), "rsim_aosoa_unpack")
    return out


# ---- synthetic workloads (host) -----------------------------------------------------------------------------------------------
def build_world(preset0=0, preset1=0, gravity_z=-650.0, boost_accel_scale=1.0, ball_radius=91.25):
    """StaticData (uint32[STATIC_WORDS]) of the procedural Soccar arena with the given car presets and mutators."""
    w = np.zeros(STATIC_WORDS, np.uint32)
    _check(lib().rsim_build_world(preset0, preset1, gravity_z, boost_accel_scale, ball_radius, _ptr(w)), "rsim_build_world")
    return w


def sample_arenas(world, scenario, seed, n, first=0):
    """ArenaStates (uint32[n, ARENA_WORDS]) of arenas first..first+n-1 of a seeded scenario."""
    s = np.zeros((n, ARENA_WORDS), np.uint32)
    _check(lib().rsim_sample_arenas(_ptr(_words(world, STATIC_WORDS, "world")), scenario, seed, first, n, _ptr(s)),
           "rsim_sample_arenas")
    return s


def sample_arenas_in(worlds, variant, scenario, seed, first=0):
    """ArenaStates of arenas first..first+len(variant)-1, arena i drawn in worlds[variant[i]] (worlds: uint32[W, STATIC_WORDS])."""
    worlds = np.ascontiguousarray(worlds, np.uint32).reshape(-1, STATIC_WORDS)
    variant = np.ascontiguousarray(variant, np.uint32)
    s = np.zeros((variant.size, ARENA_WORDS), np.uint32)
    _check(lib().rsim_sample_arenas_in(_ptr(worlds), worlds.shape[0], _ptr(variant), scenario, seed, first, variant.size, _ptr(s)),
           "rsim_sample_arenas_in")
    return s


def sample_controls(seed, decision, n, first=0):
    """Controls (uint32[n, 16]) of one decision for arenas first..first+n-1."""
    c = np.zeros((n, CONTROL_WORDS), np.uint32)
    _check(lib().rsim_sample_controls(seed, decision, first, n, _ptr(c)), "rsim_sample_controls")
    return c


# ---- device ---------------------------------------------------------------------------------------------------------------------
class DeviceStatic:
    """StaticData on the device (device.rs:47-50)."""

    def __init__(self, dev, handle):
        self.dev, self.handle = dev, handle

    @property
    def ptr(self):
        """Device address of the first StaticData (for user kernels)."""
        return lib().rsim_world_device_ptr(self.handle)

    def free(self):
        if self.handle and self.dev.handle:  # once the device is closed, the memory is released at process exit
            if lib().rsim_world_free(self.handle) == 0:  # refused while the stream records a graph: the buffer stays
                self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceStates:
    """n arenas in one AoSoA-64 device buffer (device.rs:41-45)."""

    def __init__(self, dev, handle, n):
        self.dev, self.handle, self.n = dev, handle, n

    @property
    def ptr(self):
        return lib().rsim_states_device_ptr(self.handle)

    def free(self):
        if self.handle and self.dev.handle:  # once the device is closed, the memory is released at process exit
            if lib().rsim_states_free(self.handle) == 0:  # refused while the stream records a graph: the buffer stays
                self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Graph:
    """A recorded sequence of asynchronous calls (a HIP graph), replayed with `launch`."""

    def __init__(self, dev, handle):
        self.dev, self.handle = dev, handle

    def launch(self):
        _check(lib().rsim_graph_launch(self.dev.handle, self.handle), "rsim_graph_launch")

    def free(self):
        if self.handle and self.dev.handle:  # once the device is closed, the memory is released at process exit
            if lib().rsim_graph_free(self.handle) == 0:  # refused while the stream records a graph: the buffer stays
                self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class DeviceWords:
    """A u32 device buffer (CudaSlice<u32> in the reference)."""

    def __init__(self, dev, handle, n):
        self.dev, self.handle, self.n = dev, handle, n

    @property
    def ptr(self):
        return lib().rsim_words_device_ptr(self.handle)

    def view(self, offset, n):
        """Words offset..offset+n-1 as a DeviceWords of their own (no copy; keeps this buffer alive)."""
        h = _vp()
        _check(lib().rsim_words_view(self.handle, offset, n, ctypes.byref(h)), "rsim_words_view")
        v = DeviceWords(self.dev, h.value, n)
        v.parent = self
        return self.dev._own(v)

    def free(self):
        if self.handle and self.dev.handle:  # once the device is closed, the memory is released at process exit
            if lib().rsim_words_free(self.handle) == 0:  # refused while the stream records a graph: the buffer stays
                self.handle = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class SimDevice:
    """One GPU and its stream (device.rs:30-67). Buffers, worlds and graphs made on it register themselves; `close` frees
    the ones still open (graphs and views first), as the C ABI's rsim_close requires."""

    def __init__(self, ordinal=0):
        h = _vp()
        _check(lib().rsim_open(ordinal, ctypes.byref(h)), "rsim_open")
        self.handle = h
        self.ordinal = ordinal
        self._owned = weakref.WeakSet()

    def _own(self, obj):
        self._owned.add(obj)
        return obj

    def close(self):
        if self.handle:
            owned = list(self._owned)
            # graphs, then views (they only borrow their parent's memory), then buffers and worlds
            owned.sort(key=lambda o: 0 if isinstance(o, Graph) else 1 if getattr(o, "parent", None) is not None else 2)
            for o in owned:
                o.free()
            _check(lib().rsim_close(self.handle), "rsim_close")
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self):
        return lib().rsim_stream(self.handle)

    def upload_static(self, world):
        h = _vp()
        _check(lib().rsim_upload_world(self.handle, _ptr(_words(world, STATIC_WORDS, "world")), ctypes.byref(h)),
               "rsim_upload_world")
        return self._own(DeviceStatic(self, h))

    def upload_worlds(self, worlds):
        """A table of StaticData (uint32[count, STATIC_WORDS]) for rsim_sim8_variants."""
        w = _words(worlds).reshape(-1, STATIC_WORDS)
        h = _vp()
        _check(lib().rsim_upload_worlds(self.handle, _ptr(w), w.shape[0], ctypes.byref(h)), "rsim_upload_worlds")
        return self._own(DeviceStatic(self, h))

    def upload_states(self, states):
        s = _words(states).reshape(-1, ARENA_WORDS)
        h = _vp()
        _check(lib().rsim_upload_states(self.handle, _ptr(s), s.shape[0], ctypes.byref(h)), "rsim_upload_states")
        return self._own(DeviceStates(self, h, s.shape[0]))

    def alloc_states(self, n):
        h = _vp()
        _check(lib().rsim_alloc_states(self.handle, n, ctypes.byref(h)), "rsim_alloc_states")
        return self._own(DeviceStates(self, h, n))

    def write_states(self, d, states):
        s = _words(states, d.n * ARENA_WORDS, "states")
        _check(lib().rsim_write_states(self.handle, d.handle, _ptr(s)), "rsim_write_states")

    def copy_states(self, dst, src):
        _check(lib().rsim_copy_states(self.handle, dst.handle, src.handle), "rsim_copy_states")

    def download_words(self, d):
        out = np.empty(aosoa_words(d.n), np.uint32)
        _check(lib().rsim_download_words(self.handle, d.handle, _ptr(out)), "rsim_download_words")
        return out

    def download_states(self, d):
        out = np.empty((d.n, ARENA_WORDS), np.uint32)
        _check(lib().rsim_download_states(self.handle, d.handle, _ptr(out)), "rsim_download_states")
        return out

    def identity(self, inp, out):
        _check(lib().rsim_identity(self.handle, inp.handle, out.handle), "rsim_identity")

    def identity_tick(self, inp, out):
        _check(lib().rsim_identity_tick(self.handle, inp.handle, out.handle), "rsim_identity_tick")

    def phase(self, k, states, sd):
        st = np.zeros(max(states.n, 1), np.int32)
        _check(lib().rsim_phase(self.handle, k, states.handle, sd.handle, _ptr(st)), "rsim_phase")
        return st[: states.n]

    def step_tick(self, states, sd):
        st = np.zeros(max(states.n, 1), np.int32)
        _check(lib().rsim_step_tick(self.handle, states.handle, sd.handle, _ptr(st)), "rsim_step_tick")
        return st[: states.n]

    def alloc_words(self, n):
        h = _vp()
        _check(lib().rsim_alloc_words(self.handle, n, ctypes.byref(h)), "rsim_alloc_words")
        return self._own(DeviceWords(self, h, n))

    def upload_words(self, words):
        w = _words(words)
        h = _vp()
        _check(lib().rsim_upload_words(self.handle, _ptr(w), w.size, ctypes.byref(h)), "rsim_upload_words")
        return self._own(DeviceWords(self, h, w.size))

    def write_words(self, d, words):
        w = _words(words)
        _check(lib().rsim_write_words(self.handle, d.handle, _ptr(w), w.size), "rsim_write_words")

    def read_words(self, d, n=None):
        n = d.n if n is None else n
        out = np.empty(n, np.uint32)
        _check(lib().rsim_read_words(self.handle, d.handle, _ptr(out), n), "rsim_read_words")
        return out

    def upload_controls(self, controls):
        c = _words(controls).reshape(-1, CONTROL_WORDS)
        h = _vp()
        _check(lib().rsim_upload_controls(self.handle, _ptr(c), c.shape[0], ctypes.byref(h)), "rsim_upload_controls")
        return self._own(DeviceWords(self, h, c.size))

    def sim8(self, states, sd, controls, result, block=SIM8_BLOCK, trace=None):
        """One decision of every arena, in place, asynchronously (device.rs:151-168)."""
        _check(lib().rsim_sim8(self.handle, states.handle, sd.handle, controls.handle, result.handle, block,
                               trace.handle if trace is not None else None), "rsim_sim8")

    def sim8_variants(self, states, worlds, variant, controls, result, block=SIM8_BLOCK, trace=None):
        """sim8 with arena a stepped in worlds[variant[a]] (variant: a DeviceWords of n indices, or None for world 0)."""
        _check(lib().rsim_sim8_variants(self.handle, states.handle, worlds.handle, variant.handle if variant is not None else None,
                                        controls.handle, result.handle, block, trace.handle if trace is not None else None),
               "rsim_sim8_variants")

    def run(self, states, worlds, controls, results, decisions, variant=None, block=SIM8_BLOCK):
        """`decisions` decisions in one launch (controls: decisions x n x 16 words, results: decisions x n words)."""
        _check(lib().rsim_run(self.handle, states.handle, worlds.handle, variant.handle if variant is not None else None,
                              controls.handle, results.handle, decisions, block), "rsim_run")

    def run_queued(self, states, worlds, controls, results, decisions, variant=None, block=SIM8_BLOCK, order=None,
                   workgroups=0):
        """rsim_run with persistent work-groups taking AoSoA blocks from a queue (in `order`, a DeviceWords permutation of
        the block indices, or in index order); workgroups 0: as many as the GPU holds."""
        _check(lib().rsim_run_queued(self.handle, states.handle, worlds.handle, variant.handle if variant is not None else None,
                                     controls.handle, results.handle, decisions, block,
                                     order.handle if order is not None else None, workgroups), "rsim_run_queued")

    def run_queue_capacity(self, block=SIM8_BLOCK):
        return int(lib().rsim_run_queue_capacity(self.handle, block))

    def sim8_split(self, states, worlds, controls, result, ticks_per_launch, variant=None, block=SIM8_BLOCK):
        """One decision as launches of `ticks_per_launch` ticks (the state back in HBM between launches)."""
        _check(lib().rsim_sim8_split(self.handle, states.handle, worlds.handle, variant.handle if variant is not None else None,
                                     controls.handle, result.handle, block, ticks_per_launch), "rsim_sim8_split")

    def gather_states(self, src, perm, dst):
        """dst arena i = src arena perm[i] for the dst.n arenas of dst (asynchronous). perm is a DeviceWords; the caller
        keeps every buffer alive until the work is done (freeing a buffer waits for the stream first)."""
        _check(lib().rsim_gather_states(self.handle, src.handle, perm.handle, dst.handle), "rsim_gather_states")

    def sync(self):
        _check(lib().rsim_sync(self.handle), "rsim_sync")

    def capture_begin(self):
        """Start recording the asynchronous calls on this device's stream into a HIP graph."""
        _check(lib().rsim_capture_begin(self.handle), "rsim_capture_begin")

    def capture_end(self):
        """Stop recording; returns the instantiated Graph."""
        h = _vp()
        _check(lib().rsim_capture_end(self.handle, ctypes.byref(h)), "rsim_capture_end")
        return self._own(Graph(self, h))

    def timer_begin(self):
        _check(lib().rsim_timer_begin(self.handle), "rsim_timer_begin")

    def timer_end(self):
        ms = ctypes.c_float(0)
        _check(lib().rsim_timer_end(self.handle, ctypes.byref(ms)), "rsim_timer_end")
        return float(ms.value)
