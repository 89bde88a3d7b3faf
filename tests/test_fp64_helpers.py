"""The FP64 logarithm and reciprocal square root of the CDM kernels'
near-singular sides (altar_amd/csrc/cdm_math.hpp, f_log and f_rsq): their
steps and constants are restated here, with one rounding per operation and
fused multiply-adds rounded once, and checked against exact values from
Python's decimal module.  The device's seed reciprocal and reciprocal square
root are replaced by FP32-accurate estimates with a planted relative error,
so the Newton steps are what makes the result accurate.  CPU only."""
import math
import os
import re
from decimal import Decimal, getcontext

import numpy as np

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "altar_amd", "csrc",
                   "cdm_math.hpp")
getcontext().prec = 60


def source_of(name):
    src = open(HDR).read()
    i = src.index(name)
    return src[i:src.index("#else", i)]


def fma(a, b, c):
    return float(Decimal(a) * Decimal(b) + Decimal(c))


def committed_ln2_split():
    body = source_of("inline double f_log(double x)")
    m = re.search(r"LN2_HI = (0x[0-9a-fp.+-]+), LN2_LO = (0x[0-9a-fp.+-]+);", body)
    return float.fromhex(m.group(1)), float.fromhex(m.group(2))


def committed_series():
    """The atanh series' coefficients as written (highest degree first)."""
    body = source_of("inline double f_log(double x)")
    first = re.search(r"double q = 2\.0 / (\d+)\.0;", body).group(1)
    rest = re.findall(r"q = fma\(q, z, 2\.0 / (\d+)\.0\);", body)
    return [2.0 / int(d) for d in [first] + rest]


def f_log(x, hi, lo, coeffs, seed_rel=3e-8):
    m, e = math.frexp(x)
    if m < 0.7071067811865476:
        m, e = m + m, e - 1
    f, d = m - 1.0, m + 1.0
    r = (1.0 / d) * (1.0 + seed_rel)
    r = fma(fma(-d, r, 1.0), r, r)
    r = fma(fma(-d, r, 1.0), r, r)
    s = f * r
    s = fma(fma(-s, d, f), r, s)
    z = s * s
    q = coeffs[0]
    for c in coeffs[1:]:
        q = fma(q, z, c)
    lm = fma(s * z, q, s + s)
    ed = float(e)
    return fma(ed, hi, fma(ed, lo, lm))


def f_rsq(x, seed_rel=5e-8):
    y = (1.0 / math.sqrt(x)) * (1.0 + seed_rel)
    h = 0.5 * y
    y = fma(h, fma(-x * y, y, 1.0), y)
    h = 0.5 * y
    return fma(h, fma(-x * y, y, 1.0), y)


def ulps(got, exact):
    return float(abs(Decimal(got) - exact) / Decimal(math.ulp(float(exact))))


def test_ln2_split():
    hi, lo = committed_ln2_split()
    ln2 = Decimal(2).ln()
    assert abs(Decimal(hi) + Decimal(lo) - ln2) < Decimal("1e-30")
    # at most 42 significant bits, so e * LN2_HI is exact for |e| < 2^11
    m, _ = math.frexp(hi)
    assert (m * 2 ** 42).is_integer()


def test_series_coefficients_and_truncation():
    c = committed_series()
    assert c == [2.0 / k for k in range(19, 2, -2)]
    # |s| <= (sqrt 2 - 1) / (sqrt 2 + 1): the first omitted term s^21 2/21
    # relative to the leading 2 s is below 2.4e-17
    s = (math.sqrt(2) - 1) / (math.sqrt(2) + 1)
    assert s ** 20 / 21 < 2.4e-17


def test_log_within_two_ulp():
    hi, lo = committed_ln2_split()
    c = committed_series()
    rng = np.random.default_rng(3)
    xs = np.concatenate([np.exp(rng.uniform(-40, 40, 1500)), 1 + rng.uniform(-1e-3, 1e-3, 300),
                         1 + rng.uniform(-0.35, 0.5, 300), [1.0, 2.0, 0.5, 1e-300, 1e300, 5e-324]])
    worst = 0.0
    for x in xs:
        exact = Decimal(float(x)).ln()
        for rel in (3e-8, -3e-8):
            got = f_log(float(x), hi, lo, c, rel)
            if exact == 0:
                assert got == 0.0
                continue
            worst = max(worst, ulps(got, exact))
    assert worst < 2.5, worst
    # a planted mistake: a truncated ln 2 (LN2_LO dropped) fails
    bad = max(ulps(f_log(x, hi, 0.0, c), Decimal(x).ln()) for x in (1e-200, 3e150))
    assert bad > 100


def test_rsq_within_one_ulp():
    rng = np.random.default_rng(4)
    worst = 0.0
    for x in np.exp(rng.uniform(-30, 50, 2000)):
        exact = 1 / Decimal(float(x)).sqrt()
        for rel in (5e-8, -5e-8):
            worst = max(worst, ulps(f_rsq(float(x), rel), exact))
    assert worst < 1.5, worst
    # a planted mistake: with one Newton step, 1.5 times the seed's squared
    # error (about 4e-15, some 17 ulp) is left
    def one_step(x, rel=5e-8):
        y = (1.0 / math.sqrt(x)) * (1.0 + rel)
        return fma(0.5 * y, fma(-x * y, y, 1.0), y)
    assert min(ulps(one_step(float(x)), 1 / Decimal(float(x)).sqrt()) for x in (3.0, 7e5)) > 5
