"""The V3 wire format over librsim.so (include/rsim_v3.h): RocketSim V3's arena and static dumps to and from the
ArenaState / StaticData words the device API steps.

Mirrors the public functions of crates/rocketsim-cuda/src/dump.rs: `parse_static` (:976), `parse_arena` (:885, returns
the state and its `Template`), `write_arena` (:903), `arena_header` (:871), `static_header` (:947), `fnv1a64` (:30),
`fold_records` (:395) and `unfold_records` (:425). Where the reference returns `Err(String)` these raise `FormatError`
with the same message. `write_static` has no counterpart there (V3 writes static dumps); it runs the parser's traversal
the other way.

    world = v3.parse_static(v3_static_bytes)          # uint32[STATIC_WORDS] -> SimDevice.upload_static
    state, template = v3.parse_arena(v3_arena_bytes, world)
    ...                                               # step it on the GPU
    v3_bytes = v3.write_arena(state, world, template)  # V3 imports it back
"""
import ctypes
from collections import namedtuple

import numpy as np

from . import ARENA_WORDS, STATIC_WORDS, RsimError, _ptr, _words, lib

ARENA_VERSION = 2
STATIC_VERSION = 3
ARENA_HEADER_LEN = 68
STATIC_HEADER_LEN = 76
BODY_RECORD_LEN = 219
NUM_BODIES = 23
COUNTS = 12
RECORD_WORDS = 12  # one ContactRecord: is_swap, body a, body b, world point on b, on a, normal on b (f32 bits)
ERR_SHORT, ERR_FORMAT = -2, -5


class FormatError(ValueError):
    """A dump that does not parse, or a state the format cannot carry."""


class Template(ctypes.Structure):
    """What an ArenaState leaves out of a V3 arena dump (dump.rs:205-238); `rsim_v3_template` in include/rsim_v3.h."""
    _fields_ = [
        ("config_hash", ctypes.c_uint64),
        ("static_body", (ctypes.c_uint8 * BODY_RECORD_LEN) * NUM_BODIES),
        ("ball_other", ctypes.c_uint8 * 24),
        ("num_handles", ctypes.c_uint16),
        ("next_new_handle", ctypes.c_uint16),
        ("manifold_stride", ctypes.c_uint32),
        ("ball_other_set", ctypes.c_uint32),
    ]

    def static_record(self, body):
        return bytes(self.static_body[body])


ArenaHeader = namedtuple("ArenaHeader", "version config_hash counts")
StaticHeader = namedtuple("StaticHeader", "hash counts")

_ready = False


def _lib():
    global _ready
    L = lib()
    if not _ready:
        vp, u32, u64, c_int = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64, ctypes.c_int
        sig = {
            "rsim_v3_template_bytes": ([], u64),
            "rsim_fnv1a64": ([vp, u64], u64),
            "rsim_v3_static_header": ([vp, u64, ctypes.POINTER(u64), vp], c_int),
            "rsim_v3_parse_static": ([vp, u64, vp], c_int),
            "rsim_v3_write_static": ([vp, vp, u64, ctypes.POINTER(u64)], c_int),
            "rsim_v3_arena_header": ([vp, u64, ctypes.POINTER(u64), vp], c_int),
            "rsim_v3_parse_arena": ([vp, u64, vp, vp, ctypes.POINTER(Template)], c_int),
            "rsim_v3_write_arena": ([vp, vp, ctypes.POINTER(Template), vp, u64, ctypes.POINTER(u64)], c_int),
            "rsim_v3_fold_records": ([vp, vp, vp, u32], c_int),
            "rsim_v3_unfold_records": ([vp, vp, vp, u32, ctypes.POINTER(u32)], c_int),
        }
        for name, (args, res) in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
        if L.rsim_v3_template_bytes() != ctypes.sizeof(Template):
            raise RsimError("rsim_v3_template is %d bytes in the library, %d here" %
                            (L.rsim_v3_template_bytes(), ctypes.sizeof(Template)))
        _ready = True
    return L


def _raise(rc, what):
    msg = lib().rsim_last_error().decode()
    if rc == ERR_FORMAT:
        raise FormatError(msg)
    raise RsimError("%s failed (%d): %s" % (what, rc, msg))


def _buf(data):
    b = np.frombuffer(bytes(data), np.uint8) if not isinstance(data, np.ndarray) else np.ascontiguousarray(data, np.uint8)
    return b, b.size


def _sized_call(fn, *args):
    """Calls a writer twice: once for the size, once into a buffer of that size."""
    n = ctypes.c_uint64(0)
    rc = fn(*args, None, 0, ctypes.byref(n))
    if rc not in (0, ERR_SHORT):
        _raise(rc, fn.__name__)
    out = np.empty(max(int(n.value), 1), np.uint8)
    rc = fn(*args, _ptr(out), out.size, ctypes.byref(n))
    if rc:
        _raise(rc, fn.__name__)
    return out[:n.value].tobytes()


def fnv1a64(data):
    b, n = _buf(data)
    return int(_lib().rsim_fnv1a64(_ptr(b), n))


def static_header(data):
    b, n = _buf(data)
    h, c = ctypes.c_uint64(0), np.zeros(COUNTS, np.uint32)
    rc = _lib().rsim_v3_static_header(_ptr(b), n, ctypes.byref(h), _ptr(c))
    if rc:
        _raise(rc, "rsim_v3_static_header")
    return StaticHeader(int(h.value), [int(x) for x in c])


def arena_header(data):
    b, n = _buf(data)
    h, c = ctypes.c_uint64(0), np.zeros(COUNTS, np.uint32)
    rc = _lib().rsim_v3_arena_header(_ptr(b), n, ctypes.byref(h), _ptr(c))
    if rc:
        _raise(rc, "rsim_v3_arena_header")
    return ArenaHeader(ARENA_VERSION, int(h.value), [int(x) for x in c])


def parse_static(data):
    """A V3 static dump as StaticData (uint32[STATIC_WORDS])."""
    b, n = _buf(data)
    w = np.zeros(STATIC_WORDS, np.uint32)
    rc = _lib().rsim_v3_parse_static(_ptr(b), n, _ptr(w))
    if rc:
        _raise(rc, "rsim_v3_parse_static")
    return w


def write_static(world):
    """StaticData as a V3 static dump (bytes)."""
    return _sized_call(_lib().rsim_v3_write_static, _ptr(_words(world, STATIC_WORDS, "world")))


def parse_arena(data, world):
    """A V3 arena dump as (ArenaState uint32[ARENA_WORDS], Template)."""
    b, n = _buf(data)
    s = np.zeros(ARENA_WORDS, np.uint32)
    t = Template()
    rc = _lib().rsim_v3_parse_arena(_ptr(b), n, _ptr(_words(world, STATIC_WORDS, "world")), _ptr(s), ctypes.byref(t))
    if rc:
        _raise(rc, "rsim_v3_parse_arena")
    return s, t


def write_arena(state, world, template):
    """An ArenaState as a V3 arena dump (bytes) V3's importer takes."""
    s = _words(state, ARENA_WORDS, "state")
    return _sized_call(_lib().rsim_v3_write_arena, _ptr(s), _ptr(_words(world, STATIC_WORDS, "world")),
                       ctypes.byref(template))


def fold_records(state, world, records):
    """Folds V3's contact records of a tick (uint32[n, RECORD_WORDS]) into a copy of `state`; returns the copy."""
    s = np.array(_words(state, ARENA_WORDS, "state"), copy=True)
    r = np.ascontiguousarray(records, np.uint32).reshape(-1, RECORD_WORDS)
    rc = _lib().rsim_v3_fold_records(_ptr(s), _ptr(_words(world, STATIC_WORDS, "world")), _ptr(r), r.shape[0])
    if rc:
        _raise(rc, "rsim_v3_fold_records")
    return s


def unfold_records(state, world):
    """Records (uint32[n, RECORD_WORDS]) that fold to the state's fold fields."""
    L = _lib()
    s, w = _words(state, ARENA_WORDS, "state"), _words(world, STATIC_WORDS, "world")
    n = ctypes.c_uint32(0)
    rc = L.rsim_v3_unfold_records(_ptr(s), _ptr(w), None, 0, ctypes.byref(n))
    if rc not in (0, ERR_SHORT):
        _raise(rc, "rsim_v3_unfold_records")
    out = np.zeros((max(n.value, 1), RECORD_WORDS), np.uint32)
    rc = L.rsim_v3_unfold_records(_ptr(s), _ptr(w), _ptr(out), out.shape[0], ctypes.byref(n))
    if rc:
        _raise(rc, "rsim_v3_unfold_records")
    return out[:n.value]
