"""The shared f32 arithmetic contract (pike_amd/csrc/pk_math.h), pinned
against float64 libm (numpy), IEEE float32 division/sqrt, exact special
values, and the PCG hash's published definition."""
import math

import numpy as np
import pytest

from oracle import oracle

RNG = np.random.default_rng(1234)


def ulp32(x):
    x = np.abs(np.asarray(x, np.float32))
    return np.spacing(np.maximum(x, np.float32(np.finfo(np.float32).tiny)))


def err_ulps(got, ref):
    return np.abs(got.astype(np.float64) - ref) / ulp32(ref.astype(np.float32)).astype(np.float64)


@pytest.mark.parametrize("fn,ref,lo,hi", [("sin", np.sin, -20.0, 20.0), ("cos", np.cos, -20.0, 20.0)])
def test_sin_cos_accuracy(fn, ref, lo, hi):
    x = RNG.uniform(lo, hi, 200_000).astype(np.float32)
    got = oracle.math_eval(fn, x)
    r = ref(x.astype(np.float64))
    # absolute error near zeros of the function, relative elsewhere
    abs_err = np.abs(got - r)
    assert np.all((err_ulps(got, r) <= 3.0) | (abs_err <= 1.2e-7)), abs_err.max()
    assert abs_err.max() < 2.5e-7


def test_tan_atan2_acos_accuracy():
    x = RNG.uniform(0.0, 1.4, 100_000).astype(np.float32)
    t = oracle.math_eval("tan", x)
    assert np.all(err_ulps(t, np.tan(x.astype(np.float64))) <= 6.0)
    y = RNG.normal(size=200_000).astype(np.float32)
    xx = RNG.normal(size=200_000).astype(np.float32)
    a = oracle.math_eval("atan2", y, xx)
    r = np.arctan2(y.astype(np.float64), xx.astype(np.float64))
    assert np.abs(a - r).max() < 5e-7
    c = RNG.uniform(-1, 1, 200_000).astype(np.float32)
    ac = oracle.math_eval("acos", c)
    assert np.abs(ac - np.arccos(c.astype(np.float64))).max() < 5e-7


def test_log2_exp2_pow_accuracy():
    x = np.exp(RNG.uniform(-80, 80, 100_000)).astype(np.float32)
    x = x[np.isfinite(x) & (x > 0)]
    l2 = oracle.math_eval("log2", x)
    assert np.abs(l2 - np.log2(x.astype(np.float64))).max() < 2e-5
    assert np.all(err_ulps(l2, np.log2(x.astype(np.float64))) <= 3.0)
    y = RNG.uniform(-120, 120, 100_000).astype(np.float32)
    e2 = oracle.math_eval("exp2", y)
    assert np.all(err_ulps(e2, np.exp2(y.astype(np.float64))) <= 4.0)
    b = RNG.uniform(0, 1, 100_000).astype(np.float32)
    k = RNG.uniform(0, 4, 100_000).astype(np.float32)
    p = oracle.math_eval("pow", b, k)
    r = np.power(b.astype(np.float64), k.astype(np.float64))
    assert np.all((err_ulps(p, r) <= 64.0) | (np.abs(p - r) < 1e-30))


def test_special_values():
    f = np.float32
    assert oracle.math_eval("sin", np.array([0.0], f))[0] == 0.0
    assert oracle.math_eval("cos", np.array([0.0], f))[0] == 1.0
    pi32 = f(math.pi)
    assert oracle.math_eval("atan2", np.array([0.0], f), np.array([-1.0], f))[0] == pi32
    assert oracle.math_eval("atan2", np.array([1.0], f), np.array([0.0], f))[0] == f(math.pi / 2)
    assert oracle.math_eval("acos", np.array([1.0], f))[0] == 0.0
    assert oracle.math_eval("acos", np.array([-1.0], f))[0] == pi32
    ints = np.arange(-126, 128, dtype=f)
    np.testing.assert_array_equal(oracle.math_eval("exp2", ints), np.exp2(ints.astype(np.float64)).astype(f))
    assert oracle.math_eval("pow", np.array([0.0], f), np.array([0.5], f))[0] == 0.0
    assert oracle.math_eval("pow", np.array([0.3], f), np.array([0.0], f))[0] == 1.0


def test_divide_and_sqrt_are_correctly_rounded():
    a = RNG.normal(size=100_000).astype(np.float32)
    b = RNG.normal(size=100_000).astype(np.float32)
    np.testing.assert_array_equal(oracle.math_eval("div", a, b), a / b)
    s = np.abs(a)
    np.testing.assert_array_equal(oracle.math_eval("sqrt", s), np.sqrt(s))


def _pcg_hash_int(seed):
    """pcg_hash of Jarzynski & Olano 2020 in unbounded integers mod 2^32."""
    state = seed & 0xFFFFFFFF
    word = (((state >> ((state >> 28) + 4)) ^ state) * 277803737) & 0xFFFFFFFF
    return ((word >> 22) ^ word) & 0xFFFFFFFF


def test_pcg_hash_known_answers():
    seeds = np.array([0, 1, 2, 0xDEADBEEF, 0xFFFFFFFF, 1973, 9277 * 3 + 1, 26699 | 1], np.uint32)
    got = oracle.math_eval("pcg", seeds.view(np.float32)).view(np.uint32)
    want = np.array([_pcg_hash_int(int(s)) for s in seeds], np.uint32)
    np.testing.assert_array_equal(got, want)
    # hash of 0 is 0 (RXS-M-XS fixes 0); of 1: state>>((1>>28)+4)=0 -> word = 277803737
    assert want[0] == 0
    assert want[1] == ((277803737 >> 22) ^ 277803737)


def test_pcg_stream_and_rand_edge():
    # the state advances by the LCG x -> 747796405 x + 2891336453 (mod 2^32):
    # after n steps x_n = a^n x0 + c (a^n - 1)/(a - 1), checked in closed form
    a, c, m = 747796405, 2891336453, 1 << 32
    x0 = 12345
    x = x0
    for n in range(1, 50):
        x = (a * x + c) % m
        an = pow(a, n, m * (a - 1))
        closed = (an * x0 + c * ((an - 1) // (a - 1))) % m
        assert x == closed
    # rand() = f32(h) / 2^32 rounds to exactly 1.0 for h >= 2^32 - 128
    h = np.array([0xFFFFFF80, 0xFFFFFF7F, 0xFFFFFFFF], np.uint32)
    r = h.astype(np.float32) / np.float32(4294967296.0)
    assert r[0] == 1.0 and r[2] == 1.0 and r[1] < 1.0


def test_pcg_uniformity():
    seeds = np.arange(1, 200_001, dtype=np.uint32)
    h = oracle.math_eval("pcg", seeds.view(np.float32)).view(np.uint32)
    counts = np.bincount((h >> np.uint32(24)).astype(np.int64), minlength=256)
    expected = len(seeds) / 256
    chi2 = float(((counts - expected) ** 2 / expected).sum())
    assert chi2 < 340  # 255 dof, p ~ 0.0003


@pytest.mark.parametrize("fn,lo,hi,den", [("unorm16", 0, 65535, 65535), ("snorm16", -32768, 32767, 32767),
                                          ("snorm8", -128, 127, 127), ("unorm8", 0, 255, 255)])
def test_integer_unpacking_is_the_correctly_rounded_quotient(fn, lo, hi, den):
    """The kernels unpack quantized integers (bary masks, uv, tangent frames,
    RGBA8 texels) with a reciprocal product and one fma correction instead of
    a division; over every integer of each domain that equals the correctly
    rounded q / den (numpy's float32 division), clamped at -1 for snorm as
    WGSL's unpack does."""
    q = np.arange(lo, hi + 1).astype(np.float32)
    got = oracle.math_eval(fn, q)
    want = q / np.float32(den)
    if fn.startswith("snorm"):
        want = np.maximum(want, np.float32(-1.0))
    np.testing.assert_array_equal(got.view(np.uint32), want.astype(np.float32).view(np.uint32))
