"""The correctly rounded functions against what the mathematics fixes: a correctly rounded result is the one float nearest
the exact value, which mpmath gives at high precision. This pins sinf/cosf/tanf/atan2f/powf and hypot64 independently of
the reference; atan2_64 is not correctly rounded in the reference and is pinned by the golden vectors instead."""
import numpy as np
import pytest

mpmath = pytest.importorskip("mpmath")
mpmath.mp.prec = 200


def round_f32(x):
    """mpmath value -> nearest float32 (ties to even), via the exact double then a careful float step."""
    if mpmath.isnan(x):
        return np.float32(np.nan)
    if mpmath.isinf(x):
        return np.float32(x)
    d = float(x)  # nearest double
    f = np.float32(d)
    # fix the double-rounding case: compare the two float neighbours exactly
    lo = np.nextafter(f, np.float32(-np.inf), dtype=np.float32)
    hi = np.nextafter(f, np.float32(np.inf), dtype=np.float32)
    best = min((lo, f, hi), key=lambda c: (abs(mpmath.mpf(float(c)) - x), int(np.float32(c).view(np.uint32)) & 1))
    return np.float32(best)


def sample_floats(rng, n, lo_exp=-30, hi_exp=30):
    m = rng.uniform(1, 2, n) * (2.0 ** rng.integers(lo_exp, hi_exp, n)) * rng.choice([-1, 1], n)
    return m.astype(np.float32)


@pytest.mark.parametrize("fn", [0, 1, 2])
def test_trig_correct_rounding(oracle, fn):
    rng = np.random.default_rng(10 + fn)
    xs = np.concatenate([sample_floats(rng, 400, -20, 40), (np.float32(np.pi / 2) * np.arange(-50, 50)).astype(np.float32),
                         np.array([1e-30, 3.0, 1e22, 1.5707964, 3.1415927, 6.2831855], np.float32)])
    got = oracle.libm_unary_list(fn, xs.view(np.uint32)).view(np.float32)
    f = (mpmath.sin, mpmath.cos, mpmath.tan)[fn]
    for x, g in zip(xs, got):
        want = round_f32(f(mpmath.mpf(float(x))))
        assert g.view(np.uint32) == want.view(np.uint32), "fn %d at %r: got %r want %r" % (fn, x, g, want)


def test_atan2_correct_rounding(oracle):
    rng = np.random.default_rng(3)
    y = sample_floats(rng, 400, -10, 10)
    x = sample_floats(rng, 400, -10, 10)
    got = oracle.libm_binary(0, y, x).view(np.float32)
    for a, b, g in zip(y, x, got):
        want = round_f32(mpmath.atan2(mpmath.mpf(float(a)), mpmath.mpf(float(b))))
        assert g == want, (a, b, g, want)


def test_powf_correct_rounding(oracle):
    rng = np.random.default_rng(4)
    x = np.abs(sample_floats(rng, 400, -6, 6))
    y = (rng.standard_normal(400) * 6).astype(np.float32)
    x = np.concatenate([x, np.float32([0.97, 2.0, 0.5, 1.5])])
    y = np.concatenate([y, np.float32([1.0 / 120.0, 10.0, -3.0, 2.0])])
    got = oracle.libm_binary(1, x, y).view(np.float32)
    for a, b, g in zip(x, y, got):
        want = round_f32(mpmath.power(mpmath.mpf(float(a)), mpmath.mpf(float(b))))
        assert g == want, (a, b, g, want)
    # exact cases
    assert oracle.libm_binary(1, np.float32([2.0, 4.0, 1.0]), np.float32([10.0, 0.5, 1e30])).view(np.float32).tolist() == \
        [1024.0, 2.0, 1.0]


def test_hypot64_correct_rounding(oracle):
    rng = np.random.default_rng(5)
    a = rng.standard_normal(300) * 10.0 ** rng.integers(-5, 5, 300)
    b = rng.standard_normal(300) * 10.0 ** rng.integers(-5, 5, 300)
    got = oracle.libm_f64(1, a, b)
    for x, y, g in zip(a, b, got):
        exact = mpmath.sqrt(mpmath.mpf(x) ** 2 + mpmath.mpf(y) ** 2)
        assert g == float(exact), (x, y, g, float(exact))  # float(mpf) rounds to nearest double


def test_atan2_64_accuracy(oracle):
    """atan2_64 is a double-double evaluation: within an ulp of the exact value (its exact bits are pinned by the golden
    vectors); a dropped term or wrong quadrant would be off by far more."""
    rng = np.random.default_rng(6)
    a = rng.standard_normal(300)
    b = rng.standard_normal(300)
    got = oracle.libm_f64(0, a, b)
    for x, y, g in zip(a, b, got):
        exact = mpmath.atan2(mpmath.mpf(x), mpmath.mpf(y))
        assert abs(mpmath.mpf(g) - exact) <= abs(exact) * 2.3e-16, (x, y)
