"""The FP64 logarithm, reciprocal square root and atan2 of the CDM kernels'
near-singular sides (altar_amd/csrc/cdm_math.hpp, f_log, f_rsq and
f_atan2_exact): their steps and constants are restated here, with one
rounding per operation and fused multiply-adds rounded once, and checked
against exact values from Python's decimal module (mpmath for atan2).  The
device's seed reciprocal and reciprocal square root are replaced by
FP32-accurate estimates with a planted relative error, so the Newton steps
are what makes the result accurate.  CPU only."""
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


def committed_atan2():
    """f_atan2_exact's polynomial (highest degree first) and constants."""
    body = source_of("inline double f_atan2_exact(double y, double x)")
    coeffs = [re.search(r"double p = ([-0-9a-fx.p+]+);", body).group(1)]
    coeffs += re.findall(r"p = fma\(p, s, ([-0-9a-fx.p+]+)\);", body)
    consts = dict(re.findall(r"(PIO4_HI|PIO4_LO|PIO2_HI|PIO2_LO|PI_HI|PI_LO|ATAN05_HI|ATAN05_LO|TAN_PIO8) = "
                             r"([-0-9a-fx.p+]+)", body))
    return [float.fromhex(c) for c in coeffs], {k: float.fromhex(v) for k, v in consts.items()}


def div_refined(a, b, seed_rel):
    r = (1.0 / b) * (1.0 + seed_rel)
    r = fma(r, fma(-b, r, 1.0), r)
    r = fma(r, fma(-b, r, 1.0), r)
    q = a * r
    return fma(fma(-q, b, a), r, q)


def f_atan2_exact(y, x, coeffs, k, seed_rel=3e-8):
    ax, ay = abs(x), abs(y)
    mx, mn = max(ax, ay), min(ax, ay)
    t = div_refined(mn, mx, seed_rel) if mx > 0.0 else 0.0
    red, one = t > k["TAN_PIO8"], t > 0.6875
    c = 1.0 if one else 0.5
    u = div_refined(t - c, fma(c, t, 1.0), -seed_rel) if red else t
    s = u * u
    p = coeffs[0]
    for c in coeffs[1:]:
        p = fma(p, s, c)
    a = fma(u * s, p, u)
    def add3(h, a, lo):
        s = h + a
        return s + (((h - s) + a) + lo)
    if red:
        a = add3(k["PIO4_HI"] if one else k["ATAN05_HI"], a, k["PIO4_LO"] if one else k["ATAN05_LO"])
    if ay > ax:
        a = add3(k["PIO2_HI"], -a, k["PIO2_LO"])
    if x < 0.0:
        a = add3(k["PI_HI"], -a, k["PI_LO"])
    return math.copysign(a, y)


def atan2_ulps(got, y, x):
    import mpmath
    mpmath.mp.dps = 40
    # atan2 is odd in y, -0 included (mpmath has no signed zero)
    exact = mpmath.atan2(mpmath.mpf(abs(y)), mpmath.mpf(x)) * (-1 if math.copysign(1.0, y) < 0 else 1)
    if exact == 0:
        return 0.0 if got == 0.0 else math.inf
    return float(abs(mpmath.mpf(got) - exact) / mpmath.mpf(math.ulp(float(exact))))


def test_atan2_constants_and_polynomial():
    import mpmath
    mpmath.mp.dps = 40
    coeffs, k = committed_atan2()
    assert len(coeffs) == 11
    pi = mpmath.pi
    for name, exact in (("PIO4", pi / 4), ("PIO2", pi / 2), ("PI", pi), ("ATAN05", mpmath.atan(mpmath.mpf(1) / 2))):
        assert abs(mpmath.mpf(k[name + "_HI"]) + mpmath.mpf(k[name + "_LO"]) - exact) < mpmath.mpf("1e-32")
    assert abs(k["TAN_PIO8"] - math.tan(math.pi / 8)) < 1e-16
    # the polynomial itself, evaluated exactly: atan u = u + u s P(s) to 1e-17 relative
    worst = 0
    for i in range(1, 801):
        u = mpmath.tan(pi / 8) * i / 800
        s = u * u
        p = mpmath.mpf(0)
        for c in coeffs:
            p = p * s + mpmath.mpf(c)
        worst = max(worst, abs(u + u * s * p - mpmath.atan(u)) / mpmath.atan(u))
    assert worst < 1e-17, worst
    # a planted mistake: the leading coefficient -1/3 perturbed in its 13th digit fails
    bad = list(coeffs)
    bad[-1] *= 1 + 1e-12
    u = mpmath.tan(pi / 8)
    s = u * u
    p = mpmath.mpf(0)
    for c in bad:
        p = p * s + mpmath.mpf(c)
    assert abs(u + u * s * p - mpmath.atan(u)) / mpmath.atan(u) > 1e-14


def test_atan2_within_two_ulp():
    coeffs, k = committed_atan2()
    rng = np.random.default_rng(5)
    ang = rng.uniform(-math.pi, math.pi, 1500)
    rad = np.exp(rng.uniform(-20, 20, 1500))
    pts = [(float(r * math.sin(a)), float(r * math.cos(a))) for r, a in zip(rad, ang)]
    # both sides of the reduction thresholds and of the octant swap, axes, signed zeros
    tp8 = math.tan(math.pi / 8)
    for f in (tp8 * (1 - 1e-15), tp8, tp8 * (1 + 1e-15), 0.6875 * (1 - 1e-16), 0.6875, 0.6875 * (1 + 1e-15),
              0.45, 0.5, 0.55, 1 - 1e-16, 1.0, 1e-300, 0.3, 0.7):
        for sy in (1, -1):
            for sx in (1, -1):
                pts += [(sy * f, sx * 1.0), (sy * 1.0, sx * f)]
    pts += [(0.0, 1.0), (-0.0, 1.0), (0.0, -1.0), (-0.0, -1.0), (1.0, 0.0), (-1.0, 0.0), (0.0, 0.0), (-0.0, 0.0),
            (3e-310, 1.0), (1.0, 3e-310)]
    worst = 0.0
    for y, x in pts:
        for rel in (3e-8, -3e-8):
            got = f_atan2_exact(y, x, coeffs, k, rel)
            worst = max(worst, atan2_ulps(got, y, x))
            assert math.copysign(1.0, got) == math.copysign(1.0, y) or got == 0.0
    assert worst < 1.15, worst      # measured 1.06
    assert f_atan2_exact(-0.0, 1.0, coeffs, k) == 0.0 and math.copysign(1.0, f_atan2_exact(-0.0, 1.0, coeffs, k)) < 0
    # planted mistakes: each constant's low part dropped (1.25-1.52 ulp on the
    # same points), or pi/2 in place of pi
    for name in ("PI_LO", "PIO2_LO", "PIO4_LO", "ATAN05_LO"):
        bad = dict(k, **{name: 0.0})
        assert max(atan2_ulps(f_atan2_exact(y, x, coeffs, bad), y, x) for y, x in pts) > 1.15, name
    bad = dict(k, PI_HI=k["PIO2_HI"])
    assert atan2_ulps(f_atan2_exact(0.3, -1.0, coeffs, bad), 0.3, -1.0) > 1e15
