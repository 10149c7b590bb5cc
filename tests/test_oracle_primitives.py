"""Oracle primitives against the reference's own known answers and bounds
(simd.rs:654-818) and against exact rational evaluation."""
import math

import numpy as np
import pytest

from helpers import f32c, xfma, xmul

pytestmark = []


def test_x86_min_max_semantics(oracle):
    L = oracle.lib()
    # minps/maxps return the second operand on unordered or equal inputs
    assert math.copysign(1, L.or_x86_min(0.0, -0.0)) == -1
    assert math.copysign(1, L.or_x86_min(-0.0, 0.0)) == 1
    assert L.or_x86_min(float("nan"), 1.0) == 1.0
    assert math.isnan(L.or_x86_min(1.0, float("nan")))
    assert L.or_x86_max(float("nan"), 2.0) == 2.0
    assert math.isnan(L.or_x86_max(2.0, float("nan")))
    assert L.or_x86_min(1.0, 2.0) == 1.0 and L.or_x86_max(1.0, 2.0) == 2.0


def test_geq_is_not_less_unordered(oracle):
    L = oracle.lib()
    assert L.or_geq_mask(1.0, 1.0) == 0xFFFFFFFF
    assert L.or_geq_mask(2.0, 1.0) == 0xFFFFFFFF
    assert L.or_geq_mask(1.0, 2.0) == 0
    assert L.or_geq_mask(float("nan"), 1.0) == 0xFFFFFFFF  # cmpps 21 true on NaN
    assert L.or_geq_mask(1.0, float("nan")) == 0xFFFFFFFF


def test_cvt_round_nearest_even(oracle):
    L = oracle.lib()
    for x, y in [(0.5, 0), (1.5, 2), (2.5, 2), (-0.5, 0), (-1.5, -2), (127.5, 128), (254.5, 254)]:
        assert L.or_cvt_f32_i32(x) == y
    assert L.or_cvt_f32_i32(float("nan")) == -2147483648
    assert L.or_cvt_f32_i32(3e9) == -2147483648


def test_portable_rcp_rsqrt_correctly_rounded(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(1)
    for x in rng.uniform(-100, 100, 200).astype(np.float32):
        x = float(x)
        assert L.or_rcp(x, oracle.APPROX_PORTABLE) == _div(1.0, x)
    for x in rng.uniform(1e-3, 100, 200).astype(np.float32):
        x = float(x)
        s = np.float32(np.sqrt(np.float64(x)))  # correctly rounded sqrt
        assert L.or_rsqrt(x, oracle.APPROX_PORTABLE) == _div(1.0, float(s))


def _div(a, b):
    from fractions import Fraction
    from helpers import f32r
    return f32r(Fraction(a) / Fraction(b))


def test_host_rcp_rsqrt_12_bit(oracle):
    # rcpps / rsqrtps: relative error at most 1.5 * 2^-12 (Intel and AMD)
    L = oracle.lib()
    xs = np.random.default_rng(2).uniform(0.01, 1000, 2000).astype(np.float32)
    for x in xs:
        r = L.or_rcp(float(x), oracle.APPROX_HOST_X86)
        assert abs(r * float(x) - 1) <= 1.5 * 2 ** -12
        s = L.or_rsqrt(float(x), oracle.APPROX_HOST_X86)
        assert abs(s * math.sqrt(float(x)) - 1) <= 1.5 * 2 ** -12


def test_recip_precise_error(oracle):
    # simd.rs:710-726: relative error of recip_precise below 1e-6
    L = oracle.lib()
    xs = np.random.default_rng(3).uniform(-1, 1, 4096).astype(np.float32)
    for x in xs:
        if x == 0:
            continue
        approx = L.or_recip_precise(float(x), oracle.APPROX_HOST_X86)
        assert abs(1 - approx / (1.0 / float(x))) <= 1e-6


@pytest.mark.parametrize("name,bound,scale", [("or_sin_poly", 0.0082, math.pi),
                                              ("or_cos_poly", 0.0021, math.pi)])
def test_sin_cos_error_bounds(oracle, name, bound, scale):
    # simd.rs:729-770
    fn = getattr(oracle.lib(), name)
    xs = (np.random.default_rng(4).uniform(-1, 1, 4096) * np.float32(scale)).astype(np.float32)
    ref = np.sin if "sin" in name else np.cos
    for x in xs:
        assert abs(fn(float(x)) - ref(float(x))) <= bound


def test_sin_cos_poly_exact_evaluation(oracle):
    """The polynomials with their FMA placement, evaluated exactly (simd.rs:190-233)."""
    L = oracle.lib()
    sa, sb, sc = f32c(0.982006669405), f32c(-0.152175915294), f32c(0.00533738081738135917)
    ca, cb, cc = f32c(-0.496000299455), f32c(0.0392596924214), f32c(-0.000966231179636657107)
    xs = (np.random.default_rng(5).uniform(-1, 1, 300) * np.float32(math.pi)).astype(np.float32)
    for x in xs:
        x = float(x)
        x2 = xmul(x, x)
        x3 = xmul(x, x2)
        x5 = xmul(x3, x2)
        s = xfma(x5, sc, xfma(x3, sb, xmul(x, sa)))
        assert L.or_sin_poly(x) == s
        x4 = xmul(x2, x2)
        x6 = xmul(x2, x4)
        c = xfma(x6, cc, xfma(x4, cb, xfma(x2, ca, 1.0)))
        assert L.or_cos_poly(x) == c


def test_acos_bounds(oracle):
    # simd.rs:772-810
    L = oracle.lib()
    xs = np.random.default_rng(6).uniform(-1, 1, 4096).astype(np.float32)
    for x in xs:
        assert abs(L.or_acos_fast_poly(float(x)) - math.acos(float(x))) <= 0.075
        y = float(np.float32(x * np.float32(0.98)))
        assert abs(L.or_acos_poly(y) - math.acos(y)) <= 0.017
