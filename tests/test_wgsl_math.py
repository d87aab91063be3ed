"""The implementation-defined WGSL builtins (include/rt_wgsl_math.h) against
double-precision libm and the WGSL accuracy requirements. CPU only."""
import numpy as np
import pytest

from .oracle_ffi import math_eval

RNG = np.random.default_rng(1234)


def test_sin_cos_absolute_error():
    x = np.concatenate([np.linspace(0, 2 * np.pi, 200001), RNG.uniform(-8192, 8192, 200000),
                        RNG.uniform(-10, 10, 100000)]).astype(np.float32)
    xd = x.astype(np.float64)
    for name, ref in (("sin", np.sin), ("cos", np.cos)):
        err = np.abs(math_eval(name, x).astype(np.float64) - ref(xd))
        assert err.max() < 2e-6, (name, err.max(), x[np.argmax(err)])
    # WGSL: absolute error <= 2^-11 inside [-pi, pi]; ours is far tighter
    assert 2e-6 < 2.0 ** -11


def test_sin_cos_special_values():
    x = np.array([0.0, -0.0, np.pi / 2, np.pi, 8192.5, np.inf, np.nan], np.float32)
    s, c = math_eval("sin", x), math_eval("cos", x)
    assert s[0] == 0.0 and c[0] == 1.0
    assert abs(s[2] - 1.0) < 1e-7 and abs(c[3] + 1.0) < 1e-7
    assert np.isnan(s[4:]).all() and np.isnan(c[4:]).all()  # outside the reduction range


def test_exp2_log2_relative_error():
    x = np.concatenate([RNG.uniform(-125, 127, 200000), RNG.uniform(-1, 1, 100000)]).astype(np.float32)
    e = math_eval("exp2", x).astype(np.float64)
    rel = np.abs(e / np.exp2(x.astype(np.float64)) - 1)
    assert rel.max() < 4e-7, rel.max()
    y = np.exp(RNG.uniform(-80, 80, 300000)).astype(np.float32)
    l = math_eval("log2", y).astype(np.float64)
    ref = np.log2(y.astype(np.float64))
    err = np.abs(l - ref)
    assert (err / np.maximum(np.abs(ref), 1.0)).max() < 4e-7


def test_exp_pow_accuracy_and_special_values():
    x = RNG.uniform(-80, 80, 100000).astype(np.float32)
    rel = np.abs(math_eval("exp", x).astype(np.float64) / np.exp(x.astype(np.float64)) - 1)
    # exp(x) = exp2(x*log2(e)): the rounded product adds |x|*2^-24 of relative error
    assert (rel <= 4e-7 + np.abs(x) * 1.2e-7).all()
    b = RNG.uniform(0, 1, 100000).astype(np.float32)
    for ex in (5.0, 0.5, 12.0, 1.0 / 2.2, 2.0):
        p = math_eval("pow", b, np.full_like(b, ex)).astype(np.float64)
        ref = np.power(b.astype(np.float64), np.float32(ex).astype(np.float64))
        # inherited from exp2(y * log2(x)): absolute error of the exponent times ln 2
        arg = np.abs(np.float64(ex) * np.log2(b.astype(np.float64)))
        ok = np.abs(p - ref) <= (5e-7 + arg * 2.5e-7) * ref + 1e-37
        assert ok.all(), (ex, (np.abs(p - ref) / ref).max())
    sp = math_eval("pow", np.array([0.0, 0.0, 2.0, 1.0, -1.0], np.float32), np.array([5.0, 0.5, 0.0, 7.0, 2.0], np.float32))
    assert sp[0] == 0.0 and sp[1] == 0.0 and sp[2] == 1.0 and sp[3] == 1.0 and np.isnan(sp[4])
    ex2 = math_eval("exp2", np.array([-np.inf, 128.0, -149.0, -150.5, np.nan], np.float32))
    assert ex2[0] == 0.0 and np.isinf(ex2[1]) and ex2[2] == np.float32(2.0 ** -149) and np.isnan(ex2[4])
    lg = math_eval("log2", np.array([0.0, -1.0, np.inf, 1.0, 1e-45, 8.0], np.float32))
    assert lg[0] == -np.inf and np.isnan(lg[1]) and lg[2] == np.inf and lg[3] == 0.0 and lg[5] == 3.0
    assert abs(lg[4] - np.log2(np.float64(np.float32(1e-45)))) < 1e-5


def test_sqrt_is_correctly_rounded():
    x = np.abs(RNG.normal(size=200000) * 1e3).astype(np.float32)
    assert np.array_equal(math_eval("sqrt", x), np.sqrt(x))
    assert np.array_equal(math_eval("inverse_sqrt", x[x > 0]), np.float32(1.0) / np.sqrt(x[x > 0]))


def test_float_to_int_saturates():
    x = np.array([np.nan, -5.0, -0.5, 0.0, 3.9, 4294967296.0, 1e30, 2147483648.0], np.float32)
    assert list(math_eval("u32", x)) == [0, 0, 0, 0, 3, 4294967296.0, 4294967296.0, 2147483648.0]
    i = math_eval("i32", np.array([np.nan, -3.9, 3.9, 3e9, -3e9], np.float32))
    assert list(i) == [0, -3, 3, 2147483648.0, -2147483648.0]


def test_smoothstep_closed_form():
    t = np.linspace(-0.5, 1.5, 2001).astype(np.float32)
    tc = np.clip(t.astype(np.float64), 0, 1)
    assert np.allclose(math_eval("smoothstep", t), tc * tc * (3 - 2 * tc), atol=1e-6)
