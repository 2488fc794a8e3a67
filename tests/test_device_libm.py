"""The correctly rounded functions as the GPU computes them (the core's rs::fsin / fcos / ftan / fatan2 / fpow / datan2 /
dhypot, built for gfx950 with the product's flags into tests/devlibm/libdevlibm.so, not into librsim.so).

Pinned against the reference's own correctly rounded functions (cuda/common/libm_cr.cuh compiled in place):
  - sin, cos, tan over ALL 2^32 float inputs: 256 chunk digests per function on the device against
    tests/golden/golden_libm_all.npz (written from the reference build by make_golden_libm_all.py); the reference's own
    device check is the same exhaustive sweep (crates/rocketsim-cuda/tests/libm_cr.rs:151, kernels
    cuda/simtest/simtest.cu:28-165);
  - atan2f, powf, atan2_64, hypot64 (and the unary functions) over the golden input streams of golden_libm.npz;
  - a special-value grid (signed zeros, subnormals, the normal range's ends, infinities, NaN, huge arguments, floats
    next to multiples of pi/2, exact and near powers) against the oracle, element by element.
The CPU test pins the oracle itself on six of the chunks (subnormals, the middle of the range, the infinities and NaNs,
the negative side)."""
import ctypes
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "devlibm"))
import devlibm_build  # noqa: E402
import make_golden as mg  # noqa: E402
import make_golden_libm_all as mga  # noqa: E402
from helpers import digest  # noqa: E402

GOLD = os.path.join(HERE, "golden")
FUNCS = mga.FUNCS


@pytest.fixture(scope="module")
def gall():
    return np.load(os.path.join(GOLD, "golden_libm_all.npz"))


@pytest.fixture(scope="module")
def gl():
    return np.load(os.path.join(GOLD, "golden_libm.npz"))


def note(msg):
    print("[case] " + msg, flush=True)
    sys.stdout.flush()


# ---- CPU: the oracle on a few whole chunks ----------------------------------------------------------------------------------

@pytest.mark.parametrize("chunk", [0, 63, 127, 128, 200, 255])
def test_oracle_chunks(oracle, gall, chunk):
    x = np.arange(chunk << 24, (chunk + 1) << 24, dtype=np.uint64).astype(np.uint32)
    for fn, name in enumerate(FUNCS):
        out = oracle.libm_unary(fn, chunk << 24, 1 << 24)
        assert mga.chunk_digest(x, out) == gall[name + "_chunks"][chunk], (name, chunk)


def test_chunk_digest_detects_one_flipped_bit(oracle, gall):
    """A planted one-ulp error in one output of a chunk changes its digest (the check can fail)."""
    x = np.arange(0, 1 << 24, dtype=np.uint64).astype(np.uint32)
    out = oracle.libm_unary(0, 0, 1 << 24)
    out[12345] ^= 1
    assert mga.chunk_digest(x, out) != gall["sin_chunks"][0]


# ---- GPU ----------------------------------------------------------------------------------------------------------------------

class DevLibm:
    def __init__(self, device):
        if not os.path.exists(devlibm_build.LIB):  # only when nothing is built yet
            devlibm_build.build()
        L = ctypes.CDLL(devlibm_build.LIB)
        vp, u32, u64 = ctypes.c_void_p, ctypes.c_uint32, ctypes.c_uint64
        L.devlibm_unary_digests.argtypes = [vp, ctypes.c_int, u32, u32, vp]
        L.devlibm_unary.argtypes = [vp, ctypes.c_int, vp, u64, vp]
        L.devlibm_binary.argtypes = [vp, ctypes.c_int, vp, vp, u64, vp]
        L.devlibm_f64.argtypes = [vp, ctypes.c_int, vp, vp, u64, vp]
        self.L, self.dev = L, device

    def _run(self, rc, what):
        assert rc == 0, "%s: hipError %d" % (what, rc)
        self.dev.sync()

    def digests(self, fn):
        out = self.dev.alloc_words(2 * 256)
        self._run(self.L.devlibm_unary_digests(self.dev.stream, fn, 0, 256, out.ptr), "unary digests")
        return self.dev.read_words(out).view(np.uint64)

    def unary(self, fn, x_bits):
        x = np.ascontiguousarray(x_bits, np.uint32)
        dx, dout = self.dev.upload_words(x), self.dev.alloc_words(x.size)
        self._run(self.L.devlibm_unary(self.dev.stream, fn, dx.ptr, x.size, dout.ptr), "unary")
        return self.dev.read_words(dout)

    def binary(self, op, a, b):
        a = np.ascontiguousarray(a, np.float32).view(np.uint32)
        b = np.ascontiguousarray(b, np.float32).view(np.uint32)
        da, db, dout = self.dev.upload_words(a), self.dev.upload_words(b), self.dev.alloc_words(a.size)
        self._run(self.L.devlibm_binary(self.dev.stream, op, da.ptr, db.ptr, a.size, dout.ptr), "binary")
        return self.dev.read_words(dout)

    def f64(self, op, a, b):
        a = np.ascontiguousarray(a, np.float64).view(np.uint32)
        b = np.ascontiguousarray(b, np.float64).view(np.uint32)
        da, db, dout = self.dev.upload_words(a), self.dev.upload_words(b), self.dev.alloc_words(a.size)
        self._run(self.L.devlibm_f64(self.dev.stream, op, da.ptr, db.ptr, a.size // 2, dout.ptr), "f64")
        return self.dev.read_words(dout).view(np.float64)


@pytest.fixture(scope="module")
def dl(device):
    return DevLibm(device)


@pytest.mark.gpu
@pytest.mark.parametrize("fn", [0, 1, 2])
def test_every_float_input(dl, oracle, gall, fn):
    name = FUNCS[fn]
    note("device %sf over all 2^32 inputs (256 chunk digests)" % name)
    got = dl.digests(fn)
    want = gall[name + "_chunks"]
    bad = np.flatnonzero(got != want)
    if bad.size:  # name the inputs of the first differing chunk
        c = int(bad[0])
        x = np.arange(c << 24, (c + 1) << 24, dtype=np.uint64).astype(np.uint32)
        d, o = dl.unary(fn, x), oracle.libm_unary(fn, c << 24, 1 << 24)
        i = np.flatnonzero(d != o)[:5]
        pytest.fail("%sf: %d of 256 chunks differ; chunk %d: inputs %s device %s oracle %s" %
                    (name, bad.size, c, [hex(v) for v in x[i]], [hex(v) for v in d[i]], [hex(v) for v in o[i]]))


@pytest.mark.gpu
def test_golden_streams(dl, gl):
    note("device libm over the golden input streams")
    inp = mg.libm_inputs()
    for fn, name in enumerate(FUNCS):
        assert digest(dl.unary(fn, inp["unary"])) == gl[name + "_digest"], name
    assert np.array_equal(dl.unary(0, gl["sample_unary"]), gl["sample_sin"])
    assert digest(dl.binary(0, inp["atan2_y"], inp["atan2_x"])) == gl["atan2f_digest"]
    assert digest(dl.binary(1, inp["pow_x"], inp["pow_y"])) == gl["powf_digest"]
    got = dl.f64(0, inp["f64_a"], inp["f64_b"])
    assert np.array_equal(got[:512].view(np.uint64), gl["sample_atan2_64"].view(np.uint64))
    assert digest(got) == gl["atan2_64_digest"]
    assert digest(dl.f64(1, inp["f64_a"], inp["f64_b"])) == gl["hypot64_digest"]


def _specials_f32():
    v = [0.0, 1.0, 0.5, 2.0, 3.0, 0.97, 1.0 / 120.0, 1e-45, 2e-45, 1.1754942e-38, 1.1754944e-38, 3.4028235e38, 1e30, 1e10,
         123456.79, 8388608.0, 16777216.0, 0.1, 10.0, 100.0, 1e-8, 1.0000001, 0.99999994, np.inf, np.nan]
    v = np.array(v, np.float32)
    halfpi = (np.float32(np.pi / 2) * np.arange(1, 40, dtype=np.float32))
    near = np.concatenate([halfpi, np.nextafter(halfpi, np.float32(np.inf)), np.nextafter(halfpi, np.float32(0))])
    big = np.array([2.0 ** e for e in range(20, 128, 7)], np.float32)
    v = np.concatenate([v, near, big])
    v = np.concatenate([v, -v])
    bits = v.view(np.uint32)
    return np.concatenate([bits, np.array([0x7FC00001, 0x7FA00000, 0xFFC00000, 0x00000001, 0x007FFFFF, 0x80000001,
                                           0x807FFFFF, 0x7F7FFFFF, 0xFF7FFFFF], np.uint32)])


def _specials_f64():
    d = np.array([0.0, 1.0, 2.0, 0.5, 3.0, 1e-300, 5e-324, 2.2250738585072014e-308, 1.7976931348623157e308, 1e300, 1e-20,
                  1e20, 0.1, 123.456, 65536.0, 1.0 + 2 ** -52, np.inf, np.nan, 2.0 ** 1023, 2.0 ** 1022, 1e308, 2.0 ** -1022,
                  2.0 ** -1000, 3e-310], np.float64)
    return np.concatenate([d, -d])


def test_special_value_grid_oracle_vs_reference(oracle):
    """The oracle on the special-value grid against the reference's own correctly rounded functions (compiled in place).
    Among the pairs: atan2_64 with |x| = DBL_MAX, where scaling the pair by 2^-1023 in one factor gives 0."""
    try:
        import oracle as orc
        ref = orc.RefLibm()
    except RuntimeError as e:
        pytest.skip(str(e))
    s = _specials_f32()
    for fn in range(3):
        assert np.array_equal(oracle.libm_unary_list(fn, s), ref.unary(fn, s)), FUNCS[fn]
    sf = s.view(np.float32)
    a, b = [g.ravel() for g in np.meshgrid(sf, sf)]
    for op in range(2):
        assert np.array_equal(oracle.libm_binary(op, a, b), ref.binary(op, a, b)), ("atan2f", "powf")[op]
    d = _specials_f64()
    a, b = [g.ravel() for g in np.meshgrid(d, d)]
    for op in range(2):
        got, want = oracle.libm_f64(op, a, b).view(np.uint64), ref.f64(op, a, b).view(np.uint64)
        assert np.array_equal(got, want), (("atan2_64", "hypot64")[op], a[got != want][:4], b[got != want][:4])


@pytest.mark.gpu
def test_special_value_grid(dl, oracle):
    note("device libm on the special-value grid")
    s = _specials_f32()
    for fn in range(3):
        got, want = dl.unary(fn, s), oracle.libm_unary_list(fn, s)
        assert np.array_equal(got, want), (FUNCS[fn], s[got != want][:5])
    sf = s.view(np.float32)
    a, b = [g.ravel() for g in np.meshgrid(sf, sf)]
    for op in range(2):
        got, want = dl.binary(op, a, b), oracle.libm_binary(op, a, b)
        bad = np.flatnonzero(got != want)
        assert bad.size == 0, (("atan2f", "powf")[op], [(hex(a.view(np.uint32)[i]), hex(b.view(np.uint32)[i])) for i in bad[:5]])
    d = _specials_f64()
    a, b = [g.ravel() for g in np.meshgrid(d, d)]
    for op in range(2):
        got, want = dl.f64(op, a, b), oracle.libm_f64(op, a, b)
        assert np.array_equal(got.view(np.uint64), want.view(np.uint64)), (("atan2_64", "hypot64")[op],
                                                                           a[got.view(np.uint64) != want.view(np.uint64)][:5])
