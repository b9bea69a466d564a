"""Pins of the oracle's scalar math against closed forms and exact arithmetic.

Wendland C2 kernel (WendlandKernel.cuh:39-84): W(0,h) = 21/(16 pi h^3),
W(2h) = 0, unit normalisation, W' = dW/dr. Tait EOS (Common.cuh:6-12):
c0, B and the interior pressure. The deterministic powf / cbrtf / sin / cos
replacements are checked against exactly rounded values.
"""
import math
import struct
from decimal import Decimal, getcontext
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle


def test_wendland_closed_forms():
    for h in (0.013, 0.00507, 1.0):
        w0 = oracle.wendland(0.0, h)
        assert w0 == pytest.approx(21.0 / (16.0 * math.pi * h ** 3), rel=1e-14)
        assert oracle.wendland(2.0 * h, h) == 0.0
        assert oracle.wendland(2.0 * h * (1 + 1e-9), h) == 0.0
        assert oracle.wendland_grad(0.0, h) == 0.0
        assert oracle.wendland_grad(2.0 * h, h) == 0.0


def test_wendland_normalisation():
    h = 0.013
    n = 20000
    r = (np.arange(n) + 0.5) * (2.0 * h / n)
    w = np.array([oracle.wendland(x, h) for x in r])
    integral = np.sum(4.0 * math.pi * r * r * w) * (2.0 * h / n)
    assert integral == pytest.approx(1.0, abs=1e-7)


def test_wendland_gradient_is_derivative():
    h = 0.02
    for q in (0.1, 0.5, 1.0, 1.5, 1.9):
        r = q * h
        eps = 1e-7 * h
        fd = (oracle.wendland(r + eps, h) - oracle.wendland(r - eps, h)) / (2 * eps)
        assert oracle.wendland_grad(r, h) == pytest.approx(fd, rel=1e-6)


def test_fp32_kernel_close_to_fp64():
    h = np.float32(0.013)
    for q in np.linspace(0.0, 2.0, 101):
        r = float(np.float32(q * h))
        a = oracle.wendland(r, float(h), precision=32)
        b = oracle.wendland(r, float(h), precision=64)
        assert a == pytest.approx(b, rel=2e-6, abs=1e-3)


def test_tait_closed_forms():
    c0 = 10.0 * math.sqrt(2 * 9.81 * 0.5)
    assert c0 == pytest.approx(31.320920, abs=1e-6)
    ratio = 1.009500783785
    for gamma, p_expected in ((5.0, 9499.06), (7.0, 9590.17)):
        B = 1000.0 * c0 * c0 / gamma
        assert B == pytest.approx({5.0: 196200.0, 7.0: 140142.857}[gamma], rel=1e-6)
        p = oracle.tait(float(np.float32(1000.0 * ratio)), 1000.0, float(np.float32(c0)), gamma)
        assert p == pytest.approx(p_expected, rel=2e-4)
    assert oracle.tait(1000.0, 1000.0, 30.0, 7.0) == 0.0


def _f32(x: float) -> float:
    return struct.unpack("f", struct.pack("f", x))[0]


def _ulp_diff(a: float, b: float) -> int:
    ia = struct.unpack("i", struct.pack("f", a))[0]
    ib = struct.unpack("i", struct.pack("f", b))[0]
    return abs(ia - ib)


def _round_f32(fr: Fraction) -> float:
    """Exactly rounded (nearest-even) FP32 value of a positive rational."""
    x = float(fr)  # nearest double
    f = _f32(x)
    # correct double rounding: compare the two FP32 neighbours exactly
    lo = np.nextafter(np.float32(f), np.float32(0)) if Fraction(f) > fr else np.float32(f)
    hi = np.nextafter(lo, np.float32(np.inf))
    dlo = abs(Fraction(float(lo)) - fr)
    dhi = abs(Fraction(float(hi)) - fr)
    if dlo < dhi:
        return float(lo)
    if dhi < dlo:
        return float(hi)
    return float(lo) if (struct.unpack("i", struct.pack("f", float(lo)))[0] & 1) == 0 else float(hi)


@pytest.mark.parametrize("gamma", [5, 7])
def test_powf_integer_exponent_is_exactly_rounded(gamma):
    rng = np.random.default_rng(gamma)
    xs = rng.uniform(0.5, 1.5, 4000).astype(np.float32)
    exact = 0
    for x in xs:
        got = oracle.powf(float(x), float(gamma))
        want = _round_f32(Fraction(float(x)) ** gamma)
        d = _ulp_diff(got, want)
        assert d <= 1
        exact += d == 0
    assert exact >= len(xs) - 2


def _cbrt_exact(x: float) -> float:
    getcontext().prec = 60
    d = Decimal(x)
    y = Decimal(float(x) ** (1.0 / 3.0))
    for _ in range(8):
        y = y - (y * y * y - d) / (3 * y * y)
    fr = Fraction(y)
    return _round_f32(fr)


def test_cbrtf_within_one_ulp_and_exact_cubes():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(1e-4, 10.0, 1500), rng.uniform(0.05, 1.1, 500)]).astype(np.float32)
    exact = 0
    for x in xs:
        got = oracle.cbrtf(float(x))
        want = _cbrt_exact(float(x))
        d = _ulp_diff(got, want)
        assert d <= 1, (x, got, want)
        exact += d == 0
    assert exact >= len(xs) - 2
    for v, r in ((8.0, 2.0), (27.0, 3.0), (0.125, 0.5), (1.0, 1.0), (0.0, 0.0), (-8.0, -2.0)):
        assert oracle.cbrtf(v) == r


def test_sincos_against_libm():
    for a in np.linspace(-50.0, 50.0, 2001):
        s, c = oracle.sincos(float(a))
        assert abs(s - math.sin(a)) <= 4e-16 * max(1.0, abs(a))
        assert abs(c - math.cos(a)) <= 4e-16 * max(1.0, abs(a))


def test_mix64_is_splitmix_finaliser():
    # SplitMix64 with state 0: first output 0xE220A8397B1DCDAF
    assert oracle.mix64(0) == 0xE220A8397B1DCDAF
