"""The logarithm of the analytical moments, I1_0 = log(P / Q) with
P = |2 + b + 2 sqrt(1 + b + l)| and Q = |-2 + b + 2 sqrt(1 - b + l)|
(fortran/compute_inner_integral_analytically.incl:12-20), is formed on the
device by rf::log_ratio (rigidfibers_amd/csrc/rf_device.h): exact exponent and
mantissa split, mantissa ratio brought into [sqrt(1/2), sqrt(2)), then
e ln 2 + 2 atanh(y) with y = (mP - mQ) / (mP + mQ) by the odd Taylor series to
y^23 / 23.  This test restates those steps (fma exact through fractions, as
the device's fused operations round once) and checks them against a 50-digit
decimal logarithm of the exact ratio P / Q over the (b, l) the assembly meets:
far pairs (the ratio near 1, where the relative precision matters), near and
touching pairs, nearly collinear fibers, and ratios across powers of two.
"""
import math
import random
from decimal import Decimal, getcontext
from fractions import Fraction

import pytest


def _fma(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def log_ratio(P, Q, top=23):
    """The steps of rf::log_ratio for normal positive P, Q (series to y^top / top)."""
    mP, eP = math.frexp(P)
    mQ, eQ = math.frexp(Q)
    adj = int(mP > 1.4142135623730951 * mQ) - int(mP < 0.70710678118654752 * mQ)
    mP = math.ldexp(mP, -adj)
    e = eP - eQ + adj
    num, den = mP - mQ, mP + mQ
    r = 1.0 / den  # the device starts from its reciprocal instruction; two refinements follow either way
    r = _fma(_fma(-den, r, 1.0), r, r)
    r = _fma(_fma(-den, r, 1.0), r, r)
    y = num * r
    y = _fma(_fma(-den, y, num), r, y)
    z = y * y
    p = 1.0 / top
    for k in range((top - 3) // 2, 0, -1):
        p = _fma(p, z, 1.0 / (2 * k + 1))
    t = 2.0 * y
    tail = _fma(t * z, p, t)
    de = float(e)
    return _fma(de, 0.69314718055994529, _fma(de, 2.3190468138462996e-17, tail))


def _exact(P, Q):
    getcontext().prec = 50
    return (Decimal(P) / Decimal(Q)).ln()


def _ulp_err(got, P, Q):
    ex = _exact(P, Q)
    if ex == 0:
        return 0.0 if got == 0.0 else math.inf
    return float(abs(Decimal(got) - ex) / Decimal(math.ulp(float(ex))))


def _pq(b, l):
    uu, ul = math.sqrt(1.0 + b + l), math.sqrt(1.0 - b + l)
    return abs(2.0 + b + 2.0 * uu), abs(-2.0 + b + 2.0 * ul)


def _cases():
    rng = random.Random(7)
    out = []
    for _ in range(300):  # far pairs: l in [10, 1e4], any relative direction
        l = 10.0 ** rng.uniform(1.0, 4.0)
        b = rng.uniform(-2.0, 2.0) * math.sqrt(l) * rng.random()
        out.append(_pq(b, l))
    for _ in range(300):  # near pairs and the forward-recurrence range l < 10
        l = rng.uniform(0.01, 10.0)
        b = rng.uniform(-1.0, 1.0) * 2.0 * math.sqrt(l)
        out.append(_pq(b, l))
    for _ in range(100):  # nearly collinear: l - b^2 / 4 small
        b = rng.choice([-1.0, 1.0]) * rng.uniform(2.5, 60.0)
        l = 0.25 * b * b + 10.0 ** rng.uniform(-7.0, -1.0)
        out.append(_pq(b, l))
    for _ in range(200):  # ratios near 1 and across powers of two
        Q = math.ldexp(rng.uniform(0.5, 1.0), rng.randint(-40, 40))
        P = Q * (1.0 + rng.choice([-1.0, 1.0]) * 10.0 ** rng.uniform(-15.0, -0.2))
        out.append((P, Q))
    for k in range(-3, 4):  # both mantissas at the ends of [1/2, 1)
        out.append((math.ldexp(1.0, k), 0.9999999999999999))
        out.append((math.ldexp(0.5, k), math.nextafter(0.5, 1.0)))
    return [(P, Q) for P, Q in out if P > 0 and Q > 0]


def test_log_ratio_within_four_ulp():
    worst = 0.0
    for P, Q in _cases():
        err = _ulp_err(log_ratio(P, Q), P, Q)
        worst = max(worst, err)
        assert err <= 4.0, (P, Q, log_ratio(P, Q), _exact(P, Q))
    print(f"worst error {worst:.2f} ulp over {len(_cases())} ratios")


def test_log_ratio_truncation_is_below_rounding():
    """The first omitted term y^25 / 25 at the largest |y| the range reduction
    allows (3 - 2 sqrt(2)) is below 2^-60 of 2 atanh(y)."""
    y = 3.0 - 2.0 * math.sqrt(2.0)
    assert 2.0 * y ** 25 / 25.0 < 2.0 ** -60 * 2.0 * math.atanh(y)


def test_log_ratio_catches_a_short_series():
    """A series stopped at y^17 / 17 (three terms fewer) misses the bound
    (~7 ulp) at the edge of the reduced range, so the cases above see a
    mistake of that size; the full series is within it there (the last two
    terms keep a margin: the series to y^19 already gives 0.8 ulp here)."""
    P, Q = 1.4142, 1.0  # mantissa ratio just below sqrt(2): |y| ~ 0.1716
    assert _ulp_err(log_ratio(P, Q, top=17), P, Q) > 4.0
    assert _ulp_err(log_ratio(P, Q), P, Q) <= 4.0


@pytest.mark.parametrize("P,Q", [(3.0, 3.0), (1.0, 2.0), (2.0, 1.0), (1e300, 1e-300), (5e-300, 7e300)])
def test_log_ratio_special_values(P, Q):
    assert _ulp_err(log_ratio(P, Q), P, Q) <= 4.0
