// Compound dislocation model (Nikkhoo et al. 2017) arithmetic shared by the
// device kernels and host-side checks.  Free of HIP types so that g++ can
// build it for precision studies of the FP32 formulation.
//
// The displacement of one angular-dislocation pair (one side of one
// rectangular dislocation) is v(B) - v(A), where v is AngDisDispSurf
// (models/cdm/cdm/libcdm.py:375-430).  v is linear in three transcendental
// terms, Fi = 2 atan2(y2, (r+a) cot(beta/2) - y1), log(r+a) and log(r-z3),
// whose coefficients depend only on the side.  This file therefore splits v
// into a rational part evaluated per vertex and the three transcendental
// terms, which are differenced across the two vertices of the side:
//   log(rB+aB) - log(rA+aA) = log((rB+aB)/(rA+aA))
//   log(rB-z3B) - log(rA-z3A) = log((rB-z3B)/(rA-z3A))
// and r - z3 is formed as (z1^2 + y2^2)/(r + z3) when z3 > 0, an exact
// identity (r^2 - z3^2 = z1^2 + y2^2) that removes the cancellation of the
// direct difference.
#pragma once

#include <math.h>

#if defined(__HIPCC__)
#define ALTAR_HD __host__ __device__
#else
#define ALTAR_HD
#endif

namespace altar {
namespace cdm {

// Arithmetic helpers.  In FP32 on the device they map to the hardware
// reciprocal, square root, reciprocal square root and base-2 logarithm
// (about 1 ulp each); host builds use the correctly rounded operations.
ALTAR_HD inline float f_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
// FP64: the hardware reciprocal refined by two Newton steps (error within a
// few ulp; each step squares the relative error of the estimate)
ALTAR_HD inline double f_rcp(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / x;
#endif
}
ALTAR_HD inline float f_div(float a, float b) { return a * f_rcp(b); }
ALTAR_HD inline double f_div(double a, double b) { return a * f_rcp(b); }
ALTAR_HD inline float f_sqrt(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
ALTAR_HD inline double f_sqrt(double x) { return sqrt(x); }
// 1 / sqrt(x): the hardware estimate in FP32 on the device (about 1 ulp)
ALTAR_HD inline float f_rsq(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}
// FP64: the hardware estimate refined by two Newton steps
// y <- y + y (1 - x y^2) / 2, each of which squares the relative error
// (within about 2 ulp for positive normal x; the library's sqrt and a
// reciprocal cost 27 instructions, this and the product r = x y 8)
ALTAR_HD inline double f_rsq(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * y;
    y = fma(h, fma(-x * y, y, 1.0), y);
    h = 0.5 * y;
    return fma(h, fma(-x * y, y, 1.0), y);
#else
    return 1.0 / sqrt(x);
#endif
}
// FP64 natural logarithm.  On the device: x = m 2^e with m in [sqrt(1/2),
// sqrt(2)), log m = 2 atanh(s) = 2 (s + s^3/3 + ... + s^19/19) with
// s = (m - 1) / (m + 1), |s| <= 0.1716, so the omitted terms stay below
// 2.4e-17 of the sum; s comes from the hardware reciprocal with two Newton
// steps and one residual correction, and e ln 2 is added in two parts (ln 2
// rounded to 42 bits, so e * LN2_HI is exact, and the remainder).  About
// 2 ulp for positive finite x (40 instructions; the library's log spends 98
// on double-double arithmetic); 0 gives -inf, +inf gives +inf, negative
// arguments and NaN give NaN.  Host builds use log().
ALTAR_HD inline double f_log(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    constexpr double LN2_HI = 0x1.62e42fefa3800p-1, LN2_LO = 0x1.ef35793c76730p-45;
    double m = __builtin_amdgcn_frexp_mant(x);
    int e = __builtin_amdgcn_frexp_exp(x);
    const bool low = m < 0.7071067811865476;
    m = low ? m + m : m;
    e = low ? e - 1 : e;
    const double f = m - 1.0;   // exact
    const double d = m + 1.0;
    double r = __builtin_amdgcn_rcp(d);
    r = fma(fma(-d, r, 1.0), r, r);
    r = fma(fma(-d, r, 1.0), r, r);
    double s = f * r;
    s = fma(fma(-s, d, f), r, s);
    const double z = s * s;
    double q = 2.0 / 19.0;
    q = fma(q, z, 2.0 / 17.0);
    q = fma(q, z, 2.0 / 15.0);
    q = fma(q, z, 2.0 / 13.0);
    q = fma(q, z, 2.0 / 11.0);
    q = fma(q, z, 2.0 / 9.0);
    q = fma(q, z, 2.0 / 7.0);
    q = fma(q, z, 2.0 / 5.0);
    q = fma(q, z, 2.0 / 3.0);
    const double lm = fma(s * z, q, s + s);
    const double ed = (double)e;
    double res = fma(ed, LN2_HI, fma(ed, LN2_LO, lm));
    res = x > 0.0 ? res : (x == 0.0 ? -__builtin_inf() : __builtin_nan(""));
    return x == __builtin_inf() ? x : res;
#else
    return log(x);
#endif
}

// The logarithm of the side's log(r+a) and log(r-z3) differences, in the
// unit its cL and cZ coefficients assume: the FP32 geometry carries them
// multiplied by ln 2 (convert_side), so FP32 sides take the base-2 logarithm
// (one hardware instruction, no scaling); FP64 sides the natural one.
ALTAR_HD inline float side_log(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_logf(x);
#else
    return log2f(x);
#endif
}
ALTAR_HD inline double side_log(double x) { return f_log(x); }

// atan2 in FP32: octant reduction to t in [0, 1] and atan(t) = t + t^3 q(t^2)
// with a degree-6 q from a minimax fit on [0, 1] (Lawson's iteratively
// reweighted least squares; 4.9e-8 in exact arithmetic).  In FP32: 1.2e-7 rad
// for t in [0, 1] with an exact reciprocal, 1.8e-7 with the hardware one
// (1 ulp) where |atan2| <= pi/4, and 4e-7 rad over the whole circle, where
// the result's own spacing near pi is 2.4e-7.  The sign of y, including -0,
// selects the branch, as atan2 does.
ALTAR_HD inline float f_atan2(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = mx > 0.0f ? mn * f_rcp(mx) : 0.0f;
    const float s = t * t;
    float q = -0.004355341661721468f;
    q = q * s + 0.02303992584347725f;
    q = q * s - 0.057773325592279434f;
    q = q * s + 0.09794218093156815f;
    q = q * s - 0.13976576924324036f;
    q = q * s + 0.19962702691555023f;
    q = q * s - 0.3333165943622589f;
    float a = t + t * s * q;
    a = ay > ax ? 1.57079632679489662f - a : a;
    a = x < 0.0f ? 3.14159265358979324f - a : a;
    return copysignf(a, y);
}
// The FP64 sides take the Burgers angle difference from the FP32 atan2: it
// enters through coefficients that stay bounded as the side turns vertical,
// and an FP32 angle leaves the FP64 side's error unchanged (measured on
// random sources against the FP64 oracle); the logarithms and reciprocals,
// whose coefficients grow like cot(beta)^2, stay in FP64.
ALTAR_HD inline double f_atan2(double y, double x) { return (double)f_atan2((float)y, (float)x); }

// FP64 quotient a / b: the Newton-refined reciprocal and one residual
// correction of the product
ALTAR_HD inline double f_div_refined(double a, double b)
{
    const double r = f_rcp(b);
    const double q = a * r;
    return fma(fma(-q, b, a), r, q);
}

// FP64 atan2 of the every-side-FP64 mode (the reference's precision).  On
// the device: the octant reduction of f_atan2 to t = min / max in [0, 1];
// above tan(pi/8), atan t = atan c + atan u with u = (t - c) / (1 + c t),
// c = 1/2 up to t = 11/16 and c = 1 beyond (t - c is exact there), so that
// |u| <= tan(pi/8) and atan u = u + u s P(s), s = u^2, with P of degree 10
// (interpolation of (atan(sqrt s) / sqrt s - 1) / s at the 11 Chebyshev
// nodes of [0, tan(pi/8)^2] in 60-digit arithmetic, coefficients rounded to
// FP64: 7e-18 relative); the quotients by f_div_refined; atan c, pi/2 and pi
// added in two parts, with the rounding error of the leading sum kept.
// At most 1.06 ulp on 3,000 test points spread over the circle and the
// reduction boundaries (finite arguments).  The sign of y,
// including -0, selects the branch, and atan2(0, 0) = 0.  Host builds use
// atan2().
ALTAR_HD inline double f_atan2_exact(double y, double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    constexpr double PIO4_HI = 0x1.921fb54442d18p-1, PIO4_LO = 0x1.1a62633145c07p-55;
    constexpr double PIO2_HI = 0x1.921fb54442d18p+0, PIO2_LO = 0x1.1a62633145c07p-54;
    constexpr double PI_HI = 0x1.921fb54442d18p+1, PI_LO = 0x1.1a62633145c07p-53;
    constexpr double ATAN05_HI = 0x1.dac670561bb4fp-2, ATAN05_LO = 0x1.a2b7f222f65e2p-56;
    constexpr double TAN_PIO8 = 0x1.a827999fcef32p-2;
    const double ax = fabs(x), ay = fabs(y);
    const double mx = fmax(ax, ay), mn = fmin(ax, ay);
    const double t = mx > 0.0 ? f_div_refined(mn, mx) : 0.0;
    const bool red = t > TAN_PIO8, one = t > 0x1.6p-1;
    const double c = one ? 1.0 : 0.5;
    const double u = red ? f_div_refined(t - c, fma(c, t, 1.0)) : t;
    const double s = u * u;
    double p = -0x1.3a31b1c0fd3b7p-6;
    p = fma(p, s, 0x1.4162c02b1dda3p-5);
    p = fma(p, s, -0x1.a0999c632b6edp-5);
    p = fma(p, s, 0x1.dfe6497e96323p-5);
    p = fma(p, s, -0x1.10fa77b1a6d57p-4);
    p = fma(p, s, 0x1.3b1263064f6b9p-4);
    p = fma(p, s, -0x1.745d0b28a7e37p-4);
    p = fma(p, s, 0x1.c71c71853d7fap-4);
    p = fma(p, s, -0x1.2492492436201p-3);
    p = fma(p, s, 0x1.999999999934cp-3);
    p = fma(p, s, -0x1.5555555555555p-2);
    double a = fma(u * s, p, u);
    // h + a + l with |a| <= |h|: the rounding error of h + a recovered exactly
    // (Dekker's fast two-sum) and added with the constant's low part
    auto add3 = [](double h, double a, double l) {
        const double sum = h + a;
        return sum + (((h - sum) + a) + l);
    };
    a = red ? add3(one ? PIO4_HI : ATAN05_HI, a, one ? PIO4_LO : ATAN05_LO) : a;
    a = ay > ax ? add3(PIO2_HI, -a, PIO2_LO) : a;
    a = x < 0.0 ? add3(PI_HI, -a, PI_LO) : a;
    return copysign(a, y);
#else
    return atan2(y, x);
#endif
}
// The Burgers angle difference of a side: f_atan2, or with exact set (every
// side in FP64, the reference's precision) f_atan2_exact.
ALTAR_HD inline float side_angle(float y, float x, bool) { return f_atan2(y, x); }
ALTAR_HD inline double side_angle(double y, double x, bool exact) { return exact ? f_atan2_exact(y, x) : f_atan2(y, x); }

// Side constants of one angular-dislocation pair.  Configuration II uses
// the angle beta and configuration I the angle beta - pi (libcdm.py:344-362):
// sin and cos change sign, cot does not, so the configuration-II values are
// stored and configuration I flips the sign of s, c, cZ and e13.  cF, cL, cZ
// multiply the transcendental terms dFi/2, dlog(r+a), dlog(r-z3) of v1..v3.
template <typename T>
struct Side {
    T e1x, e1y;          // first ADCS axis (horizontal); ey2 = (e1y, -e1x, 0), ey3 = -ez
    T y1AB;              // ADCS y1 of PB - PA (libcdm.py:335-337); its y2 is zero, since ey2 is
                         // normal to the side's horizontal direction, so both vertices share y2
    T dz1, dz3;          // z1B - z1A and z3B - z3A (configuration II; configuration I flips both)
    T b1, b2, b3;        // Burgers vector in ADCS, divided by 2 pi
    T s, c;              // sin and cos of beta (configuration II)
    T hI, hII;           // cot of half the angle, configurations I and II
    T kt, e13, b3s;      // k cot(beta), b3 sin(beta) - b1 cos(beta) and b3 sin(beta)
    T cF[3], cL[3], cZ[3];   // coefficients of dFi/2, dlog(r+a), dlog(r-z3) in the east, north and
                             // up components (the ADCS ones rotated to the EFCS)
};

// per-chain geometry: 12 vertices (x, y, depth a = -z) and 12 sides
template <typename T>
struct Geom {
    T vx[12], vy[12], va[12];
    T va2[12];           // a^2 per vertex
    T ka2[12];           // (1 - 2 nu) a / 2 per vertex
    Side<T> side[12];
};

// side k of rectangular dislocation q joins vertex 4q+k to vertex 4q+(k+1)%4
ALTAR_HD constexpr int side_vA(int s) { return s; }
ALTAR_HD constexpr int side_vB(int s) { return (s / 4) * 4 + ((s % 4) + 1) % 4; }

// Rational part of AngDisDispSurf at one vertex (libcdm.py:394-425 with the
// Fi, log(r+a) and log(r-z3) terms removed), contracted with the Burgers
// vector.  With z1 = y1 c + a s, z3 = y1 s - a c, u = r s - y1,
// q = 1/(r (r - z3)), H = k (t + y1 / (2 (r+a))) - y1 / r and
// W1 = y2 (b3 s - b1 c) + b2 z1, the nine terms of the reference collapse to
//   v1 = b1 y2 H/(r+a) - b2 (y1 H + k a/2)/(r+a) + u q W1
//   v2 = -b1 (k y1 t - k a/2 + y2 (y2/r - k y2 / (2 (r+a))))/(r+a)
//        - b2 y2 H/(r+a) - y2 q W1
//   v3 = (2 nu + a/r)(b1 y2 - b2 y1)/(r+a) + (c + a/r)(W1 - b3 s y2)/(r - z3)
//        + b3 s y2 q (r c + a)
// (k = 1 - 2 nu; s, c the configuration's sine and cosine).
//
// vertex_p returns r - z3 as the fraction p / qd (qd = 1 when z3 <= 0; for
// z3 > 0, r - z3 = (z1^2 + y2^2) / (r + z3) without cancellation).
// vertex_rest takes irmz = 1 / (r - z3), formed by the caller (one
// reciprocal serves both vertices of a side), and returns v1, v2, v3 and the
// atan2 abscissa of Fi.  The two vertices of a side share y2 and with it
// y2^2, y2 (b3 s - b1 c) and G3 = b3 s y2, which the caller forms once.
template <typename T>
ALTAR_HD inline void vertex_p(T z1, T z3, T y2sq, T r, T &p, T &qd)
{
    const bool pos = z3 > T(0);
    const T rw = r + fabs(z3);   // r - z3 when z3 <= 0, r + z3 when z3 > 0
    p = pos ? z1 * z1 + y2sq : rw;
    qd = pos ? rw : T(1);
}

template <typename T>
ALTAR_HD inline void vertex_rest(T y1, T y2, T y2sq, T y2e13, T G3, T a, T ka2, T r, T ir, T irpa, T s,
                                 T c, T h, const Side<T> &sd, T k05, T nu2, T z1, T irmz, T &v1, T &v2,
                                 T &v3, T &fx, T &rpa)
{
    // the formulas above with B1 = b1/(r+a), B2 = b2/(r+a), qW = q W1 and
    // K = k/(2 (r+a)) - 1/r, so that H = k t + y1 K and the bracket of v2 is
    // k y1 t - k a/2 - y2^2 K
    const T q = ir * irmz;
    const T u = r * s - y1;
    const T B1 = sd.b1 * irpa, B2 = sd.b2 * irpa;
    const T K = k05 * irpa - ir;
    const T H = sd.kt + y1 * K;
    const T W1 = y2e13 + sd.b2 * z1;
    const T yH = y2 * H;
    const T t1 = y1 * H + ka2;
    const T qW = q * W1;
    v1 = B1 * yH - B2 * t1 + u * qW;
    const T inner2 = (sd.kt * y1 - ka2) - y2sq * K;
    v2 = -(B1 * inner2 + B2 * yH + y2 * qW);
    const T gg = nu2 + a * ir;
    const T ca = c + a * ir;
    v3 = gg * (B1 * y2 - B2 * y1) + (irmz * ca) * (W1 - G3) + (G3 * q) * (r * c + a);
    rpa = r + a;
    fx = rpa * h - y1;
}

// Host-side (FP64) construction of the side constants, following
// AngSetupFSC (libcdm.py:310-372) and RDdispSurf (libcdm.py:250-273).
// Returns false when the side is skipped (|beta| or |pi - beta| below the
// FP64 machine epsilon, libcdm.py:320-324).
ALTAR_HD inline bool make_side(const double PA[3], const double PB[3], const double bEF[3], double nu,
                      Side<double> &sd, double &beta_out)
{
    const double pi = 3.14159265358979323846;
    double sv[3] = {PB[0] - PA[0], PB[1] - PA[1], PB[2] - PA[2]};
    double ns = sqrt(sv[0] * sv[0] + sv[1] * sv[1] + sv[2] * sv[2]);
    double beta = acos(-sv[2] / ns);
    beta_out = beta;
    const double eps = 2.220446049250313e-16;
    if (fabs(beta) < eps || fabs(pi - beta) < eps) return false;
    double n1 = sqrt(sv[0] * sv[0] + sv[1] * sv[1]);
    double e1x = sv[0] / n1, e1y = sv[1] / n1;
    sd.e1x = e1x;
    sd.e1y = e1y;
    sd.y1AB = e1x * sv[0] + e1y * sv[1];
    // ey2 = ey3 x ey1 = (e1y, -e1x, 0); ey3 = (0, 0, -1)
    double b1 = e1x * bEF[0] + e1y * bEF[1];
    double b2 = e1y * bEF[0] - e1x * bEF[1];
    double b3 = -bEF[2];
    const double i2pi = 1.0 / (2.0 * pi);
    sd.b1 = b1 * i2pi;
    sd.b2 = b2 * i2pi;
    sd.b3 = b3 * i2pi;
    const double k = 1.0 - 2.0 * nu;
    // configuration II: the angle beta; configuration I: beta - pi
    const double sB = sin(beta), cB = cos(beta), tB = 1.0 / tan(beta);
    sd.s = sB;
    sd.c = cB;
    // from A to B: y1 changes by -y1AB and the depth a = -z by -sv[2]
    sd.dz1 = -sd.y1AB * cB - sv[2] * sB;
    sd.dz3 = -sd.y1AB * sB + sv[2] * cB;
    sd.hII = 1.0 / tan(beta / 2.0);
    sd.hI = 1.0 / tan((-pi + beta) / 2.0);
    sd.kt = k * tB;
    double B1 = sd.b1, B2 = sd.b2, B3 = sd.b3;
    sd.e13 = B3 * sB - B1 * cB;
    sd.b3s = B3 * sB;
    // Fi enters as 2 * atan2(...): the factor 2 is folded in here
    sd.cF[0] = 2.0 * (B1 * (1.0 - k * tB * tB));
    sd.cF[1] = 2.0 * (B2 * (1.0 + k * tB * tB));
    sd.cF[2] = 2.0 * (B1 * k * tB + B3);
    sd.cL[0] = -B2 * k * (0.5 - tB * tB);
    sd.cL[1] = B1 * k * (0.5 + tB * tB);
    sd.cL[2] = -B2 * k * tB;
    sd.cZ[0] = -B2 * k * tB * tB * cB;
    sd.cZ[1] = -B1 * k * tB / sB;
    sd.cZ[2] = B2 * k * tB * cB;
    // ADCS (v1, v2, v3) -> EFCS (e1x v1 + e1y v2, e1y v1 - e1x v2, -v3)
    double *cs[3] = {sd.cF, sd.cL, sd.cZ};
    for (double *cc : cs) {
        const double c0 = cc[0], c1 = cc[1], c2 = cc[2];
        cc[0] = e1x * c0 + e1y * c1;
        cc[1] = e1y * c0 - e1x * c1;
        cc[2] = -c2;
    }
    return true;
}

// One coordinate of the 12 vertices (libcdm.py:185-214): P0 the centre's
// coordinate, (c0, c1, c2) the row of the rotation matrix, full axes A.
// Contraction into fused multiply-adds is off, so every caller (the
// geometry, the model's verify, the fused trajectory kernel) rounds alike.
ALTAR_HD inline void vertex_row(double P0, double Ax, double Ay, double Az, double c0, double c1, double c2,
                                double out[12])
{
#pragma clang fp contract(off)
    const double p1 = P0 + Ay * c1 / 2.0 + Az * c2 / 2.0;
    const double p2 = p1 - Ay * c1;
    const double p3 = p2 - Az * c2;
    const double p4 = p1 - Az * c2;
    const double q1 = P0 - Ax * c0 / 2.0 + Az * c2 / 2.0;
    const double q2 = q1 + Ax * c0;
    const double q3 = q2 - Az * c2;
    const double q4 = q1 - Az * c2;
    const double r1 = P0 + Ax * c0 / 2.0 + Ay * c1 / 2.0;
    const double r2 = r1 - Ax * c0;
    const double r3 = r2 - Ay * c1;
    const double r4 = r1 - Ay * c1;
    out[0] = p1; out[1] = p2; out[2] = p3; out[3] = p4;
    out[4] = q1; out[5] = q2; out[6] = q3; out[7] = q4;
    out[8] = r1; out[9] = r2; out[10] = r3; out[11] = r4;
}

constexpr double DEG = 3.14159265358979323846 / 180.0;

// FP64 rotation and vertices of libcdm.py:163-214.  V[v] = (x, y, z).
// Returns true when any vertex is above the free surface.
ALTAR_HD inline bool make_vertices(double X0, double Y0, double depth, double ax, double ay, double az,
                          double omx, double omy, double omz, double V[12][3])
{
#pragma clang fp contract(off)
    ax *= 2.0; ay *= 2.0; az *= 2.0;
    double cx = cos(omx * DEG), sx = sin(omx * DEG);
    double cy = cos(omy * DEG), sy = sin(omy * DEG);
    double cz = cos(omz * DEG), sz = sin(omz * DEG);
    double Rx[3][3] = {{1, 0, 0}, {0, cx, sx}, {0, -sx, cx}};
    double Ry[3][3] = {{cy, 0, -sy}, {0, 1, 0}, {sy, 0, cy}};
    double Rz[3][3] = {{cz, sz, 0}, {-sz, cz, 0}, {0, 0, 1}};
    double T1[3][3], R[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            T1[i][j] = Ry[i][0] * Rx[0][j] + Ry[i][1] * Rx[1][j] + Ry[i][2] * Rx[2][j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i][j] = Rz[i][0] * T1[0][j] + Rz[i][1] * T1[1][j] + Rz[i][2] * T1[2][j];
        }
    double P0[3] = {X0, Y0, -depth};
    for (int i = 0; i < 3; ++i) {
        double row[12];
        vertex_row(P0[i], ax, ay, az, R[i][0], R[i][1], R[i][2], row);
        for (int v = 0; v < 12; ++v) V[v][i] = row[v];
    }
    bool above = false;
    for (int v = 0; v < 12; ++v) above = above || (V[v][2] > 0.0);
    return above;
}

// The free-surface check of make_vertices without the rest of the geometry:
// the heights need only the rotation's last row, which is (sin wY,
// -cos wY sin wX, cos wY cos wX) whatever wZ (Rz leaves z alone).  Products
// by the rotation matrices' zeros and ones are exact, so make_vertices'
// row is these three values bit for bit, and the heights come from the same
// vertex_row: the two functions agree for every source.
ALTAR_HD inline bool above_free_surface(double depth, double ax, double ay, double az, double omx, double omy)
{
#pragma clang fp contract(off)
    const double cx = cos(omx * DEG), sx = sin(omx * DEG);
    const double cy = cos(omy * DEG), sy = sin(omy * DEG);
    double z[12];
    vertex_row(-depth, ax * 2.0, ay * 2.0, az * 2.0, sy, cy * -sx, cy * cx, z);
    bool above = false;
    for (int v = 0; v < 12; ++v) above = above || (z[v] > 0.0);
    return above;
}

// flags word of a chain's geometry
constexpr unsigned FLAG_INVALID = 1u << 31;   // pierces the free surface
constexpr unsigned FLAG_NAN = 1u << 30;       // degenerate rectangle (zero-length side)
ALTAR_HD inline constexpr unsigned skip_bit(int s) { return 1u << s; }
ALTAR_HD inline constexpr unsigned fp64_bit(int s) { return 1u << (12 + s); }

// Builds the FP64 geometry and flags of one source.  |sin(beta)| below
// fp64_sin marks a side for FP64 evaluation: the side coefficients grow
// like cot(beta)^2 and the FP32 differences across the pair lose accuracy.
// Side s of the CDM whose vertices are V: its constants into sd (zero when
// the side is skipped) and its flag bits (skip, FP64, degenerate rectangle).
ALTAR_HD inline unsigned make_side_of(const double V[12][3], double ax, double ay, double az, double opening,
                                      double nu, double fp64_sin, int s, Side<double> &sd)
{
    sd = Side<double>{};
    // which rectangles contribute (libcdm.py:226-243, on the full axes)
    const double Ax = 2.0 * ax, Ay = 2.0 * ay, Az = 2.0 * az;
    bool on[3] = {true, true, true};
    if (Ax == 0 && Ay == 0 && Az == 0) { on[0] = on[1] = on[2] = false; }
    else if (Ax == 0 && Ay != 0 && Az != 0) { on[1] = on[2] = false; }
    else if (Ax != 0 && Ay == 0 && Az != 0) { on[0] = on[2] = false; }
    else if (Ax != 0 && Ay != 0 && Az == 0) { on[0] = on[1] = false; }
    const int q = s / 4;
    if (!on[q]) return skip_bit(s);
    const double *P1 = V[4 * q], *P2 = V[4 * q + 1], *P4 = V[4 * q + 3];
    double e21[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
    double e41[3] = {P4[0] - P1[0], P4[1] - P1[1], P4[2] - P1[2]};
    double vn[3] = {e21[1] * e41[2] - e21[2] * e41[1], e21[2] * e41[0] - e21[0] * e41[2],
                    e21[0] * e41[1] - e21[1] * e41[0]};
    double nn = sqrt(vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
    if (!(nn > 0)) return FLAG_NAN | skip_bit(s);
    double b[3] = {opening * vn[0] / nn, opening * vn[1] / nn, opening * vn[2] / nn};
    const double *PA = V[side_vA(s)], *PB = V[side_vB(s)];
    double sl = (PB[0] - PA[0]) * (PB[0] - PA[0]) + (PB[1] - PA[1]) * (PB[1] - PA[1]) +
                (PB[2] - PA[2]) * (PB[2] - PA[2]);
    if (!(sl > 0)) return FLAG_NAN | skip_bit(s);
    double beta;
    if (!make_side(PA, PB, b, nu, sd, beta)) {
        sd = Side<double>{};
        return skip_bit(s);
    }
    return fabs(sin(beta)) < fp64_sin ? fp64_bit(s) : 0u;
}

ALTAR_HD inline unsigned make_geometry(double X0, double Y0, double depth, double ax, double ay, double az,
                              double omx, double omy, double omz, double opening, double nu,
                              double fp64_sin, Geom<double> &G)
{
    double V[12][3];
    unsigned flags = 0;
    if (make_vertices(X0, Y0, depth, ax, ay, az, omx, omy, omz, V)) flags |= FLAG_INVALID;
    for (int v = 0; v < 12; ++v) {
        G.vx[v] = V[v][0];
        G.vy[v] = V[v][1];
        G.va[v] = -V[v][2];
        G.va2[v] = G.va[v] * G.va[v];
        G.ka2[v] = 0.5 * (1.0 - 2.0 * nu) * G.va[v];
    }
    for (int s = 0; s < 12; ++s) flags |= make_side_of(V, ax, ay, az, opening, nu, fp64_sin, s, G.side[s]);
    return flags;
}

// Displacement (ue, un, uv) of one side at a point, everything in type T,
// added to ue, un, uv.
// dxA = X - xA etc.; rA, irA, irpaA are the per-vertex quantities (r, 1/r,
// 1/(r+a)).  The second vertex's ADCS coordinates follow from the first
// one's (libcdm.py:335-337): y1B = y1A - y1AB, and y2B = y2A, because the
// ADCS y2 axis is normal to the side's horizontal direction (the reference
// subtracts the rotated side vector's y2, which is rounding, about 1e-13 m);
// z1 and z3 at B are those at A plus side constants.  With one y2 at both
// vertices, both Burgers angles lie in the same half plane and their
// difference is one atan2: atan2(y2 (xA - xB), xA xB + y2^2).
//
// One reciprocal per side: D = 1 / (pA pB qB) gives 1 / (rA - z3A) =
// qA pB qB D, 1 / (rB - z3B) = qB pA qB D and the log argument
// (rB - z3B) / (rA - z3A) = pB qA pB D.  The product stays inside the FP32
// range (lengths below 1e5 m: pA pB qB < 1e25), and its three extra
// roundings leave the error of the FP32 field unchanged (a CPU study of
// 1,000 random sources x 2,000 points against the FP64 evaluation: median
// 1.5e-6 and p99 2.6e-5 of the field's scale, as before).
template <typename T>
ALTAR_HD inline void side_disp(const Side<T> &sd, T nu, T k,
                               T dxA, T dyA, T aA, T kaA, T rA, T irA, T irpaA,
                               T aB, T kaB, T rB, T irB, T irpaB,
                               T &ue, T &un, T &uv, bool exact_angle = false)
{
    const T e1x = sd.e1x, e1y = sd.e1y;
    const T y1A = e1x * dxA + e1y * dyA;
    const T y2 = e1y * dxA - e1x * dyA;
    const T y1B = y1A - sd.y1AB;
    // configuration I (y1A >= 0) uses beta - pi: sin, cos, cZ, e13 change sign
    const bool cfgI = y1A >= T(0);
    const T sg = cfgI ? T(-1) : T(1);
    const T s = sg * sd.s, c = sg * sd.c;
    const T h = cfgI ? sd.hI : sd.hII;
    const T k05 = T(0.5) * k, nu2 = T(2) * nu;
    const T sy2 = sg * y2;
    const T y2sq = y2 * y2, y2e13 = sd.e13 * sy2, G3 = sd.b3s * sy2;
    const T z1A = y1A * c + aA * s, z3A = y1A * s - aA * c;
    const T z1B = z1A + sg * sd.dz1, z3B = z3A + sg * sd.dz3;
    T pA, qA, pB, qB;
    vertex_p<T>(z1A, z3A, y2sq, rA, pA, qA);
    vertex_p<T>(z1B, z3B, y2sq, rB, pB, qB);
    const T D = f_rcp((pA * pB) * qB);
    const T u1 = pB * D;                        // 1 / (pA qB)
    const T irmzA = qA * (qB * u1);
    const T irmzB = qB * (qB * (pA * D));
    const T rz = (pB * qA) * u1;                // (rB - z3B) / (rA - z3A)
    T v1A, v2A, v3A, fxA, rpaA;
    T v1B, v2B, v3B, fxB, rpaB;
    vertex_rest<T>(y1A, y2, y2sq, y2e13, G3, aA, kaA, rA, irA, irpaA, s, c, h, sd, k05, nu2, z1A, irmzA,
                   v1A, v2A, v3A, fxA, rpaA);
    vertex_rest<T>(y1B, y2, y2sq, y2e13, G3, aB, kaB, rB, irB, irpaB, s, c, h, sd, k05, nu2, z1B, irmzB,
                   v1B, v2B, v3B, fxB, rpaB);
    // the angle from (fxA, y2) to (fxB, y2)
    const T dFi = side_angle(y2 * (fxA - fxB), fxA * fxB + y2sq, exact_angle);
    // log((rB + aB) / (rA + aA)) and log((rB - z3B) / (rA - z3A))
    const T dLr = side_log(rpaB * irpaA);
    const T dLz = sg * side_log(rz);
    // added to the running sums, one multiply-add per term
    const T d1 = v1B - v1A, d2 = v2B - v2A, d3 = v3B - v3A;
    ue = sd.cZ[0] * dLz + (sd.cL[0] * dLr + (sd.cF[0] * dFi + (e1y * d2 + (e1x * d1 + ue))));
    un = sd.cZ[1] * dLz + (sd.cL[1] * dLr + (sd.cF[1] * dFi + ((un - e1x * d2) + e1y * d1)));
    uv = sd.cZ[2] * dLz + (sd.cL[2] * dLr + (sd.cF[2] * dFi + (uv - d3)));
}

// FP64 side constants to another precision.  An FP32 target takes cL and cZ
// multiplied by ln 2 (in FP64, then rounded once), for side_log's base-2
// logarithm.
template <typename T, typename S>
ALTAR_HD inline void convert_side(const Side<S> &a, Side<T> &b)
{
    const S lscale = sizeof(T) == 4 ? S(0.693147180559945309417232121458) : S(1);
    b.e1x = (T)a.e1x; b.e1y = (T)a.e1y;
    b.y1AB = (T)a.y1AB;
    b.dz1 = (T)a.dz1; b.dz3 = (T)a.dz3;
    b.b1 = (T)a.b1; b.b2 = (T)a.b2; b.b3 = (T)a.b3;
    b.s = (T)a.s; b.c = (T)a.c;
    b.hI = (T)a.hI; b.hII = (T)a.hII;
    b.kt = (T)a.kt; b.e13 = (T)a.e13; b.b3s = (T)a.b3s;
    for (int j = 0; j < 3; ++j) {
        b.cF[j] = (T)a.cF[j];
        b.cL[j] = (T)(a.cL[j] * lscale);
        b.cZ[j] = (T)(a.cZ[j] * lscale);
    }
}

template <typename T, typename S>
ALTAR_HD inline void convert(const Geom<S> &in, Geom<T> &out)
{
    for (int v = 0; v < 12; ++v) {
        out.vx[v] = (T)in.vx[v];
        out.vy[v] = (T)in.vy[v];
        out.va[v] = (T)in.va[v];
        out.va2[v] = (T)in.va2[v];
        out.ka2[v] = (T)in.ka2[v];
    }
    for (int s = 0; s < 12; ++s) convert_side(in.side[s], out.side[s]);
}

// Full displacement at one point from a prepared geometry (host reference
// of the device kernel's arithmetic; T is the working precision).
template <typename T>
inline void point_disp(const Geom<T> &G, const Geom<double> &G64, unsigned flags, double nu,
                       double X, double Y, double &ue, double &un, double &uv)
{
    T r[12], ir[12], irpa[12], dx[12], dy[12];
    for (int v = 0; v < 12; ++v) {
        dx[v] = (T)X - G.vx[v];
        dy[v] = (T)Y - G.vy[v];
        T a = G.va[v];
        r[v] = sqrt(dx[v] * dx[v] + dy[v] * dy[v] + a * a);
        ir[v] = T(1) / r[v];
        irpa[v] = T(1) / (r[v] + a);
    }
    T sue = 0, sun = 0, suv = 0;
    double due = 0, dun = 0, duv = 0;
    T tnu = (T)nu, tk = (T)(1.0 - 2.0 * nu);
    for (int s = 0; s < 12; ++s) {
        if (flags & skip_bit(s)) continue;
        int a = side_vA(s), b = side_vB(s);
        if (flags & fp64_bit(s)) {
            double dxa = X - G64.vx[a], dya = Y - G64.vy[a], aa = G64.va[a];
            double dxb = X - G64.vx[b], dyb = Y - G64.vy[b], ab = G64.va[b];
            double ra = sqrt(dxa * dxa + dya * dya + aa * aa);
            double rb = sqrt(dxb * dxb + dyb * dyb + ab * ab);
            side_disp<double>(G64.side[s], nu, 1.0 - 2.0 * nu, dxa, dya, aa, G64.ka2[a], ra, 1.0 / ra,
                              1.0 / (ra + aa), ab, G64.ka2[b], rb, 1.0 / rb, 1.0 / (rb + ab), due, dun, duv);
        } else {
            side_disp<T>(G.side[s], tnu, tk, dx[a], dy[a], G.va[a], G.ka2[a], r[a], ir[a], irpa[a],
                         G.va[b], G.ka2[b], r[b], ir[b], irpa[b], sue, sun, suv);
        }
    }
    ue = (double)sue + due;
    un = (double)sun + dun;
    uv = (double)suv + duv;
}

}  // namespace cdm
}  // namespace altar
