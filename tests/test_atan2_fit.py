"""The FP32 atan2 of the CDM kernels (altar_amd/csrc/cdm_math.hpp, f_atan2):
its polynomial is re-derived here and its error checked against numpy.

f_atan2 reduces (y, x) to t = min(|x|, |y|) / max(|x|, |y|) in [0, 1] and
evaluates atan(t) = t + t^3 q(t^2) with q of degree 6.  q comes from a
minimax fit on [0, 1], approached by Lawson's iteratively reweighted least
squares.  CPU only."""
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "altar_amd", "csrc",
                   "cdm_math.hpp")
F32 = np.float32


def committed_coefficients():
    """q's coefficients as written in f_atan2, highest degree first."""
    src = open(HDR).read()
    body = src[src.index("inline float f_atan2(float y, float x)"):]
    body = body[:body.index("return copysignf")]
    first = re.search(r"float q = (-?[0-9.e-]+)f;", body).group(1)
    rest = re.findall(r"q = q \* s ([+-]) ([0-9.e-]+)f;", body)
    return np.array([float(first)] + [float(v) if sgn == "+" else -float(v) for sgn, v in rest])


def horner(c, t):
    s = t * t
    q = np.zeros_like(t)
    for k in c:
        q = q * s + k
    return t + t * s * q


def lawson_minimax(t, degree=6, iters=300):
    """Minimax fit of atan(t) - t by t^3 q(t^2) (Lawson 1961: weights
    multiplied by the absolute error each pass)."""
    s = t * t
    A = np.stack([t ** 3 * s ** k for k in range(degree, -1, -1)], 1)
    b = np.arctan(t) - t
    w = np.full_like(t, 1.0 / len(t))
    for _ in range(iters):
        sw = np.sqrt(w)
        c, *_ = np.linalg.lstsq(A * sw[:, None], b * sw, rcond=None)
        w = w * np.abs(A @ c - b)
        w /= w.sum()
    return c


def fma32(a, b, c):
    # a*b is exact in FP64 for FP32 operands; one rounding to FP32 after the add
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F32)


def atan2_f32(y, x, coeffs, rcp_rel=0.0):
    """f_atan2 in FP32 arithmetic, with contracted multiply-adds as the device
    compiles them; rcp_rel perturbs the reciprocal (the device's is within
    1 ulp)."""
    y, x = y.astype(F32), x.astype(F32)
    ax, ay = np.abs(x), np.abs(y)
    mx, mn = np.maximum(ax, ay), np.minimum(ax, ay)
    with np.errstate(divide="ignore", invalid="ignore"):
        r = ((F32(1.0) / mx) * F32(1.0 + rcp_rel)).astype(F32)
        t = np.where(mx > 0, (mn * r).astype(F32), F32(0.0))
    s = (t * t).astype(F32)
    q = np.full_like(t, F32(coeffs[0]))
    for k in coeffs[1:]:
        q = fma32(q, s, np.full_like(t, F32(k)))
    a = fma32((t * s).astype(F32), q, t)
    a = np.where(ay > ax, (F32(1.57079632679489662) - a).astype(F32), a)
    a = np.where(x < 0, (F32(3.14159265358979324) - a).astype(F32), a)
    return np.copysign(a, y)


def test_coefficients_rederive_from_a_minimax_fit():
    c = committed_coefficients()
    assert len(c) == 7
    t = np.linspace(0.0, 1.0, 20001)
    fit = lawson_minimax(t)
    # the committed q is this fit, to the digits that survive FP32 storage
    assert np.max(np.abs(fit - c) / np.abs(c)) < 2e-5
    # and within 25% of the minimax error in exact arithmetic
    e_fit = np.max(np.abs(horner(fit, t) - np.arctan(t)))
    e_com = np.max(np.abs(horner(c, t) - np.arctan(t)))
    assert e_fit < 5.5e-8 and e_com < 1.25 * e_fit
    # a planted mistake: one coefficient's sign flipped fails the bound
    bad = c.copy()
    bad[3] = -bad[3]
    assert np.max(np.abs(horner(bad, t) - np.arctan(t))) > 1e-2


def test_fp32_error_against_numpy():
    c = committed_coefficients()
    # first octant, t = y / x in [0, 1]: 1.2e-7 rad
    t = np.linspace(0.0, 1.0, 200001).astype(F32)
    got = atan2_f32(t, np.ones_like(t), c)
    assert np.max(np.abs(got - np.arctan(t.astype(np.float64)))) < 1.2e-7
    # the whole circle at three radii, with a reciprocal 1 ulp off either
    # way: 4e-7 rad (the result's own FP32 spacing near pi is 2.4e-7)
    ang = np.linspace(-np.pi, np.pi, 400001)
    for rad in (1e-3, 1.0, 7.3e4):
        y, x = (rad * np.sin(ang)).astype(F32), (rad * np.cos(ang)).astype(F32)
        want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
        for rel in (0.0, 2.0 ** -23, -(2.0 ** -23)):
            got = atan2_f32(y, x, c, rel)
            err = np.abs(got.astype(np.float64) - want)
            assert err.max() < 4e-7, (rad, rel)
            inner = np.abs(want) <= np.pi / 4
            assert err[inner].max() < (1.4e-7 if rel == 0.0 else 1.8e-7), (rad, rel)
    # signed zeros and the axes select the branch as atan2 does
    y = np.array([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 0.0], F32)
    x = np.array([1.0, 1.0, -1.0, -1.0, 0.0, 0.0, 0.0], F32)
    got = atan2_f32(y, x, c)
    want = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    assert np.all(np.abs(got - want) < 2.5e-7) and np.array_equal(np.signbit(got), np.signbit(want))
