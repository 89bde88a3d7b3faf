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
// reciprocal, square root and base-2 logarithm (about 1 ulp each); FP64 and
// host builds use the correctly rounded operations.
ALTAR_HD inline float f_rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
ALTAR_HD inline double f_rcp(double x) { return 1.0 / x; }
ALTAR_HD inline float f_div(float a, float b) { return a * f_rcp(b); }
ALTAR_HD inline double f_div(double a, double b) { return a / b; }
ALTAR_HD inline float f_sqrt(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
ALTAR_HD inline double f_sqrt(double x) { return sqrt(x); }
ALTAR_HD inline float f_log(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
#else
    return logf(x);
#endif
}
ALTAR_HD inline double f_log(double x) { return log(x); }

// atan2 in FP32: octant reduction to t in [0, 1] and atan(t) = t + t^3 q(t^2)
// with a degree-6 q fitted here (weighted least squares approaching the
// minimax error on [0, 1]); max error 1.2e-7 rad in FP32.  The sign of y,
// including -0, selects the branch, as atan2 does.
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
ALTAR_HD inline double f_atan2(double y, double x) { return atan2(y, x); }

// side constants of one angular-dislocation pair for one configuration
// (configuration I uses the angle beta - pi, II uses beta; libcdm.py:344-362)
template <typename T>
struct SideCfg {
    T s, c, t, h;        // sin, cos, cot of the angle, and cot of half the angle
    T cF[3], cL[3], cZ[3];  // coefficients of dFi/2, dlog(r+a), dlog(r-z3) in v1..v3
};

template <typename T>
struct Side {
    T e1x, e1y;          // first ADCS axis (horizontal); ey2 = (e1y, -e1x, 0), ey3 = -ez
    T y1AB, y2AB;        // ADCS coordinates of PB - PA (libcdm.py:337-339)
    T b1, b2, b3;        // Burgers vector in ADCS, divided by 2 pi
    SideCfg<T> cfg[2];   // [0]: configuration I, [1]: configuration II
};

// per-chain geometry: 12 vertices (x, y, depth a = -z) and 12 sides
template <typename T>
struct Geom {
    T vx[12], vy[12], va[12];
    Side<T> side[12];
};

// side k of rectangular dislocation q joins vertex 4q+k to vertex 4q+(k+1)%4
ALTAR_HD constexpr int side_vA(int s) { return s; }
ALTAR_HD constexpr int side_vB(int s) { return (s / 4) * 4 + ((s % 4) + 1) % 4; }

// Rational part of AngDisDispSurf at one vertex (libcdm.py:394-425 with the
// Fi, log(r+a) and log(r-z3) terms removed), contracted with the Burgers
// vector.  Also returns r - z3 and the two atan2 arguments of Fi.
template <typename T>
ALTAR_HD inline void vertex_terms(T y1, T y2, T a, T r, T ir, T rpa, T irpa,
                                  const SideCfg<T> &g, T nu, T k, T b1, T b2, T b3,
                                  T &v1, T &v2, T &v3, T &rmz, T &fy, T &fx)
{
    const T s = g.s, c = g.c, t = g.t;
    T z1 = y1 * c + a * s;
    T z3 = y1 * s - a * c;
    T rmz_ = z3 > T(0) ? f_div(z1 * z1 + y2 * y2, r + z3) : r - z3;
    T irmz = f_rcp(rmz_);
    T q = ir * irmz;
    T u = r * s - y1;
    T y1sq = y1 * y1, y2sq = y2 * y2;
    T h = k * (t + T(0.5) * y1 * irpa) - y1 * ir;
    T gg = T(2) * nu + a * ir;
    T ca = c + a * ir;
    T r1b1 = y2 * irpa * h - y2 * u * c * q;
    T r2b1 = -irpa * (k * (y1 * t - T(0.5) * a - T(0.5) * y2sq * irpa) + y2sq * ir) + y2sq * c * q;
    T r3b1 = y2 * irpa * gg - y2 * c * irmz * ca;
    T r1b2 = -irpa * (k * (y1 * t + T(0.5) * a + T(0.5) * y1sq * irpa) - y1sq * ir) + z1 * u * q;
    T r2b2 = -y2 * irpa * h - y2 * z1 * q;
    T r3b2 = -y1 * irpa * gg + z1 * irmz * ca;
    T sq = s * q;
    T r1b3 = y2 * u * sq;
    T r2b3 = -y2sq * sq;
    T r3b3 = y2 * (r * c + a) * sq;
    v1 = b1 * r1b1 + b2 * r1b2 + b3 * r1b3;
    v2 = b1 * r2b1 + b2 * r2b2 + b3 * r2b3;
    v3 = b1 * r3b1 + b2 * r3b2 + b3 * r3b3;
    rmz = rmz_;
    fy = y2;
    fx = rpa * g.h - y1;
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
    sd.y2AB = e1y * sv[0] - e1x * sv[1];
    // ey2 = ey3 x ey1 = (e1y, -e1x, 0); ey3 = (0, 0, -1)
    double b1 = e1x * bEF[0] + e1y * bEF[1];
    double b2 = e1y * bEF[0] - e1x * bEF[1];
    double b3 = -bEF[2];
    const double i2pi = 1.0 / (2.0 * pi);
    sd.b1 = b1 * i2pi;
    sd.b2 = b2 * i2pi;
    sd.b3 = b3 * i2pi;
    const double k = 1.0 - 2.0 * nu;
    for (int cfg = 0; cfg < 2; ++cfg) {
        double ang = cfg == 0 ? -pi + beta : beta;
        SideCfg<double> &g = sd.cfg[cfg];
        g.s = sin(ang);
        g.c = cos(ang);
        g.t = 1.0 / tan(ang);
        g.h = 1.0 / tan(ang / 2.0);
        double s = g.s, c = g.c, t = g.t;
        double B1 = sd.b1, B2 = sd.b2, B3 = sd.b3;
        // Fi enters as 2 * atan2(...): the factor 2 is folded in here
        g.cF[0] = 2.0 * (B1 * (1.0 - k * t * t));
        g.cF[1] = 2.0 * (B2 * (1.0 + k * t * t));
        g.cF[2] = 2.0 * (B1 * k * t + B3);
        g.cL[0] = -B2 * k * (0.5 - t * t);
        g.cL[1] = B1 * k * (0.5 + t * t);
        g.cL[2] = -B2 * k * t;
        g.cZ[0] = -B2 * k * t * t * c;
        g.cZ[1] = -B1 * k * t / s;
        g.cZ[2] = B2 * k * t * c;
    }
    return true;
}

// FP64 rotation and vertices of libcdm.py:163-214.  V[v] = (x, y, z).
// Returns true when any vertex is above the free surface.
ALTAR_HD inline bool make_vertices(double X0, double Y0, double depth, double ax, double ay, double az,
                          double omx, double omy, double omz, double V[12][3])
{
    const double d2r = 3.14159265358979323846 / 180.0;
    ax *= 2.0; ay *= 2.0; az *= 2.0;
    double cx = cos(omx * d2r), sx = sin(omx * d2r);
    double cy = cos(omy * d2r), sy = sin(omy * d2r);
    double cz = cos(omz * d2r), sz = sin(omz * d2r);
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
        double c0 = R[i][0], c1 = R[i][1], c2 = R[i][2];
        double p1 = P0[i] + ay * c1 / 2.0 + az * c2 / 2.0;
        double p2 = p1 - ay * c1;
        double p3 = p2 - az * c2;
        double p4 = p1 - az * c2;
        double q1 = P0[i] - ax * c0 / 2.0 + az * c2 / 2.0;
        double q2 = q1 + ax * c0;
        double q3 = q2 - az * c2;
        double q4 = q1 - az * c2;
        double r1 = P0[i] + ax * c0 / 2.0 + ay * c1 / 2.0;
        double r2 = r1 - ax * c0;
        double r3 = r2 - ay * c1;
        double r4 = r1 - ay * c1;
        V[0][i] = p1; V[1][i] = p2; V[2][i] = p3; V[3][i] = p4;
        V[4][i] = q1; V[5][i] = q2; V[6][i] = q3; V[7][i] = q4;
        V[8][i] = r1; V[9][i] = r2; V[10][i] = r3; V[11][i] = r4;
    }
    bool above = false;
    for (int v = 0; v < 12; ++v) above = above || (V[v][2] > 0.0);
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
    }
    // which rectangles contribute (libcdm.py:226-243, on the full axes)
    double Ax = 2.0 * ax, Ay = 2.0 * ay, Az = 2.0 * az;
    bool on[3] = {true, true, true};
    if (Ax == 0 && Ay == 0 && Az == 0) { on[0] = on[1] = on[2] = false; }
    else if (Ax == 0 && Ay != 0 && Az != 0) { on[1] = on[2] = false; }
    else if (Ax != 0 && Ay == 0 && Az != 0) { on[0] = on[2] = false; }
    else if (Ax != 0 && Ay != 0 && Az == 0) { on[0] = on[1] = false; }
    for (int q = 0; q < 3; ++q) {
        const double *P1 = V[4 * q], *P2 = V[4 * q + 1], *P4 = V[4 * q + 3];
        double e21[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
        double e41[3] = {P4[0] - P1[0], P4[1] - P1[1], P4[2] - P1[2]};
        double vn[3] = {e21[1] * e41[2] - e21[2] * e41[1], e21[2] * e41[0] - e21[0] * e41[2],
                        e21[0] * e41[1] - e21[1] * e41[0]};
        double nn = sqrt(vn[0] * vn[0] + vn[1] * vn[1] + vn[2] * vn[2]);
        double b[3] = {opening * vn[0] / nn, opening * vn[1] / nn, opening * vn[2] / nn};
        for (int k = 0; k < 4; ++k) {
            int s = 4 * q + k;
            Side<double> &sd = G.side[s];
            sd = Side<double>{};
            if (!on[q]) { flags |= skip_bit(s); continue; }
            if (!(nn > 0)) { flags |= FLAG_NAN; flags |= skip_bit(s); continue; }
            double beta;
            const double *PA = V[side_vA(s)], *PB = V[side_vB(s)];
            double sl = (PB[0] - PA[0]) * (PB[0] - PA[0]) + (PB[1] - PA[1]) * (PB[1] - PA[1]) +
                        (PB[2] - PA[2]) * (PB[2] - PA[2]);
            if (!(sl > 0)) { flags |= FLAG_NAN; flags |= skip_bit(s); continue; }
            if (!make_side(PA, PB, b, nu, sd, beta)) { flags |= skip_bit(s); continue; }
            if (fabs(sin(beta)) < fp64_sin) flags |= fp64_bit(s);
        }
    }
    return flags;
}

// Displacement (ue, un, uv) of one side at a point, everything in type T.
// dxA = X - xA etc.; rA, irA, rpaA, irpaA are the per-vertex quantities.
// The second vertex's ADCS coordinates are formed from the first one's as
// in libcdm.py:337-339, so both share the sign of y2 on the Burgers
// function's branch cut.
template <typename T>
ALTAR_HD inline void side_disp(const Side<T> &sd, T nu, T k,
                               T dxA, T dyA, T aA, T rA, T irA, T rpaA, T irpaA,
                               T aB, T rB, T irB, T rpaB, T irpaB,
                               T &ue, T &un, T &uv)
{
    T e1x = sd.e1x, e1y = sd.e1y;
    T y1A = e1x * dxA + e1y * dyA;
    T y2A = e1y * dxA - e1x * dyA;
    T y1B = y1A - sd.y1AB;
    T y2B = y2A - sd.y2AB;
    const SideCfg<T> &g = y1A >= T(0) ? sd.cfg[0] : sd.cfg[1];
    T v1A, v2A, v3A, rmzA, fyA, fxA;
    T v1B, v2B, v3B, rmzB, fyB, fxB;
    vertex_terms<T>(y1A, y2A, aA, rA, irA, rpaA, irpaA, g, nu, k, sd.b1, sd.b2, sd.b3,
                    v1A, v2A, v3A, rmzA, fyA, fxA);
    vertex_terms<T>(y1B, y2B, aB, rB, irB, rpaB, irpaB, g, nu, k, sd.b1, sd.b2, sd.b3,
                    v1B, v2B, v3B, rmzB, fyB, fxB);
    T dFi = f_atan2(fyB, fxB) - f_atan2(fyA, fxA);
    T dLr = f_log(f_div(rpaB, rpaA));
    T dLz = f_log(f_div(rmzB, rmzA));
    T v1 = (v1B - v1A) + g.cF[0] * dFi + g.cL[0] * dLr + g.cZ[0] * dLz;
    T v2 = (v2B - v2A) + g.cF[1] * dFi + g.cL[1] * dLr + g.cZ[1] * dLz;
    T v3 = (v3B - v3A) + g.cF[2] * dFi + g.cL[2] * dLr + g.cZ[2] * dLz;
    ue = e1x * v1 + e1y * v2;
    un = e1y * v1 - e1x * v2;
    uv = -v3;
}

template <typename T, typename S>
ALTAR_HD inline void convert(const Geom<S> &in, Geom<T> &out)
{
    for (int v = 0; v < 12; ++v) {
        out.vx[v] = (T)in.vx[v];
        out.vy[v] = (T)in.vy[v];
        out.va[v] = (T)in.va[v];
    }
    for (int s = 0; s < 12; ++s) {
        const Side<S> &a = in.side[s];
        Side<T> &b = out.side[s];
        b.e1x = (T)a.e1x; b.e1y = (T)a.e1y;
        b.y1AB = (T)a.y1AB; b.y2AB = (T)a.y2AB;
        b.b1 = (T)a.b1; b.b2 = (T)a.b2; b.b3 = (T)a.b3;
        for (int c = 0; c < 2; ++c) {
            b.cfg[c].s = (T)a.cfg[c].s; b.cfg[c].c = (T)a.cfg[c].c;
            b.cfg[c].t = (T)a.cfg[c].t; b.cfg[c].h = (T)a.cfg[c].h;
            for (int j = 0; j < 3; ++j) {
                b.cfg[c].cF[j] = (T)a.cfg[c].cF[j];
                b.cfg[c].cL[j] = (T)a.cfg[c].cL[j];
                b.cfg[c].cZ[j] = (T)a.cfg[c].cZ[j];
            }
        }
    }
}

// Full displacement at one point from a prepared geometry (host reference
// of the device kernel's arithmetic; T is the working precision).
template <typename T>
inline void point_disp(const Geom<T> &G, const Geom<double> &G64, unsigned flags, double nu,
                       double X, double Y, double &ue, double &un, double &uv)
{
    T r[12], ir[12], rpa[12], irpa[12], dx[12], dy[12];
    for (int v = 0; v < 12; ++v) {
        dx[v] = (T)X - G.vx[v];
        dy[v] = (T)Y - G.vy[v];
        T a = G.va[v];
        r[v] = sqrt(dx[v] * dx[v] + dy[v] * dy[v] + a * a);
        ir[v] = T(1) / r[v];
        rpa[v] = r[v] + a;
        irpa[v] = T(1) / rpa[v];
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
            double e, n, u;
            side_disp<double>(G64.side[s], nu, 1.0 - 2.0 * nu, dxa, dya, aa, ra, 1.0 / ra, ra + aa,
                              1.0 / (ra + aa), ab, rb, 1.0 / rb, rb + ab, 1.0 / (rb + ab),
                              e, n, u);
            due += e; dun += n; duv += u;
        } else {
            T e, n, u;
            side_disp<T>(G.side[s], tnu, tk, dx[a], dy[a], G.va[a], r[a], ir[a], rpa[a], irpa[a],
                         G.va[b], r[b], ir[b], rpa[b], irpa[b], e, n, u);
            sue += e; sun += n; suv += u;
        }
    }
    ue = (double)sue + due;
    un = (double)sun + dun;
    uv = (double)suv + duv;
}

}  // namespace cdm
}  // namespace altar
