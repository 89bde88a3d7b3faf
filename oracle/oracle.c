/*
 * FP64 CPU oracle for the CATMIP inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the
 * smoke() entry point and bench.py's CPU-baseline leg may load it.  The
 * product (altar_amd/) never links or calls it.
 *
 * Every routine restates the reference algorithm in plain FP64 C, literally
 * and in the reference's operation order, so that it can be checked against
 * the reference's own golden vectors and against closed forms.  File:line
 * citations are relative to the reference tree.
 *
 * Parity pins (checked in tests/test_oracle_*.py):
 *   mogi_*        closed form of Mogi (1958), radial symmetry, dV linearity
 *   cdm_*         MATLAB golden vectors models/cdm/tests/libcdm.py:27-46,
 *                 zero-opening, linearity in opening, far-field point-source
 *                 limit, plus vectors written by tests/golden/make_cdm_vectors.py
 *   l2_*          textbook quadratic form r^T Cd^-1 r via LAPACK (numpy)
 *   philox        Random123 known-answer vectors (Salmon et al. 2011)
 *   cov_*         brute-force dbeta scan, closed-form weighted moments
 *   mh_*          acceptance bookkeeping A+R+U = C*S, analytic posteriors
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* the all-cores CPU baseline build parallelises the loops over samples */
#ifdef ORACLE_OMP
#define OMP_FOR _Pragma("omp parallel for schedule(dynamic, 4)")
#else
#define OMP_FOR
#endif

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  The reference draws
 * from GSL's ranlxs2 (altar/altar/simulations/GSLRNG.py:28); that stream is
 * not reproducible here, so both the oracle and the GPU path use this
 * counter-based generator with the same (key, counter) layout.            */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void oracle_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int round = 0; round < 10; ++round) {
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n1 = lo1;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += PHILOX_W0; k1 += PHILOX_W1;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* counter layout: (chain, tag<<24 | word, step, stage); key = seed */
static void draw(uint64_t seed, uint32_t stage, uint32_t step, uint32_t chain,
                 uint32_t tag, uint32_t word, uint32_t out[4])
{
    uint32_t ctr[4] = {chain, (tag << 24) | word, step, stage};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    oracle_philox4x32_10(ctr, key, out);
}

/* 53-bit uniform on (0, 1] from two words: exact in FP64 */
static double u53(uint32_t hi, uint32_t lo)
{
    uint64_t m = (((uint64_t)hi << 32) | lo) >> 11;
    return ((double)m + 1.0) * (1.0 / 9007199254740992.0);
}

/* P standard normals per chain by Box-Muller: one Philox block gives two */
void oracle_normals(uint64_t seed, uint32_t stage, uint32_t step, uint32_t tag,
                    int C, int P, double *out /* C x P */)
{
    for (int c = 0; c < C; ++c) {
        for (int k = 0; k < P; k += 2) {
            uint32_t w[4];
            draw(seed, stage, step, (uint32_t)c, tag, (uint32_t)(k / 2), w);
            double u1 = u53(w[0], w[1]);
            double u2 = u53(w[2], w[3]);
            double rad = sqrt(-2.0 * log(u1));
            double ang = 2.0 * M_PI * u2;
            out[(size_t)c * P + k] = rad * cos(ang);
            if (k + 1 < P) out[(size_t)c * P + k + 1] = rad * sin(ang);
        }
    }
}

/* one uniform on (0,1] per chain (GSL uniform_pos semantics, Metropolis.py:54) */
void oracle_uniforms(uint64_t seed, uint32_t stage, uint32_t step, uint32_t tag,
                     int C, double *out)
{
    for (int c = 0; c < C; ++c) {
        uint32_t w[4];
        draw(seed, stage, step, (uint32_t)c, tag, 0u, w);
        out[c] = u53(w[0], w[1]);
    }
}

/* two uniforms on (0,1] per (chain, word): words (0,1) and (2,3) of the block */
void oracle_uniform_pairs(uint64_t seed, uint32_t stage, uint32_t step, uint32_t tag,
                          int C, int W, double *out /* C x W x 2 */)
{
    for (int c = 0; c < C; ++c)
        for (int k = 0; k < W; ++k) {
            uint32_t w[4];
            draw(seed, stage, step, (uint32_t)c, tag, (uint32_t)k, w);
            out[((size_t)c * W + k) * 2] = u53(w[0], w[1]);
            out[((size_t)c * W + k) * 2 + 1] = u53(w[2], w[3]);
        }
}

/* ------------------------------------------------------------------------ */
/* Mogi point source: models/mogi/lib/libmogi/Source.cc:65-118 and residuals
 * :121-154.  theta is C x P row-major; strength is log10(dV) (:78).        */
void oracle_mogi_residuals(const double *theta, int C, int P,
                           int xIdx, int dIdx, int sIdx, int offsetIdx,
                           const double *ox, const double *oy, const double *los /* N x 3 */,
                           const double *d, const int *oid, int N, double nu,
                           double *res /* C x N */)
{
    OMP_FOR
    for (int s = 0; s < C; ++s) {
        const double *t = theta + (size_t)s * P;
        double xs = t[xIdx], ys = t[xIdx + 1], ds = t[dIdx];
        double S = pow(10.0, t[sIdx]);
        for (int n = 0; n < N; ++n) {
            double x = xs - ox[n];
            double y = ys - oy[n];
            double dd = ds;
            double R = sqrt(x * x + y * y + dd * dd);
            double Cc = (nu - 1.0) * S / M_PI;
            double CR3 = Cc / (R * R * R);
            double u = (x * los[3 * n] + y * los[3 * n + 1] - dd * los[3 * n + 2]) * CR3;
            double shift = offsetIdx >= 0 ? t[offsetIdx + oid[n]] : 0.0;
            u -= shift;
            res[(size_t)s * N + n] = u - d[n];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* CDM (Nikkhoo et al. 2017) as ported in models/cdm/cdm/libcdm.py:39-430.  */

static double dot3(const double a[3], const double b[3]) { return a[0]*b[0] + a[1]*b[1] + a[2]*b[2]; }
static double norm3(const double a[3]) { return sqrt(dot3(a, a)); }
static void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1]*b[2] - a[2]*b[1];
    o[1] = a[2]*b[0] - a[0]*b[2];
    o[2] = a[0]*b[1] - a[1]*b[0];
}

/* libcdm.py:375-430 */
void oracle_ang_dis_disp_surf(double y1, double y2, double beta, double b1, double b2, double b3,
                              double nu, double a, double v[3])
{
    double sinB = sin(beta);
    double cosB = cos(beta);
    double cotB = 1.0 / tan(beta);
    double z1 = y1 * cosB + a * sinB;
    double z3 = y1 * sinB - a * cosB;
    double r2 = y1 * y1 + y2 * y2 + a * a;
    double r = sqrt(r2);
    double Fi = 2.0 * atan2(y2, (r + a) / tan(beta / 2.0) - y1);
    double k = 1.0 - 2.0 * nu;

    double v1b1 = b1 / 2.0 / M_PI * (
        (1.0 - k * cotB * cotB) * Fi +
        y2 / (r + a) * (k * (cotB + y1 / 2.0 / (r + a)) - y1 / r) -
        y2 * (r * sinB - y1) * cosB / r / (r - z3));
    double v2b1 = b1 / 2.0 / M_PI * (
        k * ((0.5 + cotB * cotB) * log(r + a) - cotB / sinB * log(r - z3)) -
        1.0 / (r + a) * (k * (y1 * cotB - a / 2.0 - y2 * y2 / 2.0 / (r + a)) + y2 * y2 / r) +
        y2 * y2 * cosB / r / (r - z3));
    double v3b1 = b1 / 2.0 / M_PI * (
        k * Fi * cotB + y2 / (r + a) * (2.0 * nu + a / r) -
        y2 * cosB / (r - z3) * (cosB + a / r));

    double v1b2 = b2 / 2.0 / M_PI * (
        -k * ((0.5 - cotB * cotB) * log(r + a) + cotB * cotB * cosB * log(r - z3)) -
        1.0 / (r + a) * (k * (y1 * cotB + 0.5 * a + y1 * y1 / 2.0 / (r + a)) - y1 * y1 / r) +
        z1 * (r * sinB - y1) / r / (r - z3));
    double v2b2 = b2 / 2.0 / M_PI * (
        (1.0 + k * cotB * cotB) * Fi -
        y2 / (r + a) * (k * (cotB + y1 / 2.0 / (r + a)) - y1 / r) -
        y2 * z1 / r / (r - z3));
    double v3b2 = b2 / 2.0 / M_PI * (
        -k * cotB * (log(r + a) - cosB * log(r - z3)) -
        y1 / (r + a) * (2.0 * nu + a / r) +
        z1 / (r - z3) * (cosB + a / r));

    double v1b3 = b3 / 2.0 / M_PI * (y2 * (r * sinB - y1) * sinB / r / (r - z3));
    double v2b3 = b3 / 2.0 / M_PI * (-y2 * y2 * sinB / r / (r - z3));
    double v3b3 = b3 / 2.0 / M_PI * (Fi + y2 * (r * cosB + a) * sinB / r / (r - z3));

    v[0] = v1b1 + v1b2 + v1b3;
    v[1] = v2b1 + v2b2 + v2b3;
    v[2] = v3b1 + v3b2 + v3b3;
}

/* libcdm.py:310-372 (one observation point) */
static void ang_setup_fsc(double X, double Y, const double b[3],
                          const double PA[3], const double PB[3], double nu, double u[3])
{
    double side[3] = {PB[0] - PA[0], PB[1] - PA[1], PB[2] - PA[2]};
    double beta = acos(-side[2] / norm3(side));
    u[0] = u[1] = u[2] = 0.0;
    if (fabs(beta) < DBL_EPSILON || fabs(M_PI - beta) < DBL_EPSILON) return;

    double ey1[3] = {side[0], side[1], 0.0};
    double n1 = norm3(ey1);
    ey1[0] /= n1; ey1[1] /= n1;
    double ey3[3] = {0.0, 0.0, -1.0};
    double ey2[3];
    cross3(ey3, ey1, ey2);

    double pA[3] = {X - PA[0], Y - PA[1], 0.0 - PA[2]};
    double y1A = dot3(ey1, pA), y2A = dot3(ey2, pA);
    double y1AB = dot3(ey1, side), y2AB = dot3(ey2, side);
    double y1B = y1A - y1AB, y2B = y2A - y2AB;
    double b1 = dot3(ey1, b), b2 = dot3(ey2, b), b3 = dot3(ey3, b);

    double vA[3], vB[3];
    if (beta * y1A >= 0.0) {
        /* configuration I */
        oracle_ang_dis_disp_surf(y1A, y2A, -M_PI + beta, b1, b2, b3, nu, -PA[2], vA);
        oracle_ang_dis_disp_surf(y1B, y2B, -M_PI + beta, b1, b2, b3, nu, -PB[2], vB);
    } else {
        /* configuration II */
        oracle_ang_dis_disp_surf(y1A, y2A, beta, b1, b2, b3, nu, -PA[2], vA);
        oracle_ang_dis_disp_surf(y1B, y2B, beta, b1, b2, b3, nu, -PB[2], vB);
    }
    double v1 = vB[0] - vA[0], v2 = vB[1] - vA[1], v3 = vB[2] - vA[2];
    /* A^T v: columns of A^T are ey1, ey2, ey3 */
    u[0] = ey1[0] * v1 + ey2[0] * v2 + ey3[0] * v3;
    u[1] = ey1[1] * v1 + ey2[1] * v2 + ey3[1] * v3;
    u[2] = ey1[2] * v1 + ey2[2] * v2 + ey3[2] * v3;
}

/* libcdm.py:250-273 */
static void rd_disp_surf(double X, double Y, const double P1[3], const double P2[3],
                         const double P3[3], const double P4[3], double opening, double nu,
                         double u[3])
{
    double e21[3] = {P2[0] - P1[0], P2[1] - P1[1], P2[2] - P1[2]};
    double e41[3] = {P4[0] - P1[0], P4[1] - P1[1], P4[2] - P1[2]};
    double vn[3];
    cross3(e21, e41, vn);
    double nn = norm3(vn);
    double b[3] = {opening * vn[0] / nn, opening * vn[1] / nn, opening * vn[2] / nn};
    double s1[3], s2[3], s3[3], s4[3];
    ang_setup_fsc(X, Y, b, P1, P2, nu, s1);
    ang_setup_fsc(X, Y, b, P2, P3, nu, s2);
    ang_setup_fsc(X, Y, b, P3, P4, nu, s3);
    ang_setup_fsc(X, Y, b, P4, P1, nu, s4);
    for (int i = 0; i < 3; ++i) u[i] = s1[i] + s2[i] + s3[i] + s4[i];
}

static double cosd(double a) { return cos(a * M_PI / 180.0); }
static double sind(double a) { return sin(a * M_PI / 180.0); }

/* the 12 vertices of libcdm.py:185-214; returns 0 if all are at or below the
 * free surface, 1 otherwise (where libcdm.py:219-224 raises) */
int oracle_cdm_vertices(double X0, double Y0, double depth, double ax, double ay, double az,
                        double omegaX, double omegaY, double omegaZ, double V[12][3])
{
    ax *= 2.0; ay *= 2.0; az *= 2.0;
    double Rx[3][3] = {{1, 0, 0}, {0, cosd(omegaX), sind(omegaX)}, {0, -sind(omegaX), cosd(omegaX)}};
    double Ry[3][3] = {{cosd(omegaY), 0, -sind(omegaY)}, {0, 1, 0}, {sind(omegaY), 0, cosd(omegaY)}};
    double Rz[3][3] = {{cosd(omegaZ), sind(omegaZ), 0}, {-sind(omegaZ), cosd(omegaZ), 0}, {0, 0, 1}};
    double RyRx[3][3], R[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            RyRx[i][j] = 0;
            for (int k = 0; k < 3; ++k) RyRx[i][j] += Ry[i][k] * Rx[k][j];
        }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            R[i][j] = 0;
            for (int k = 0; k < 3; ++k) R[i][j] += Rz[i][k] * RyRx[k][j];
        }
    double c0[3] = {R[0][0], R[1][0], R[2][0]};
    double c1[3] = {R[0][1], R[1][1], R[2][1]};
    double c2[3] = {R[0][2], R[1][2], R[2][2]};
    double P0[3] = {X0, Y0, -depth};
    for (int i = 0; i < 3; ++i) {
        double p1 = P0[i] + ay * c1[i] / 2.0 + az * c2[i] / 2.0;
        double p2 = p1 - ay * c1[i];
        double p3 = p2 - az * c2[i];
        double p4 = p1 - az * c2[i];
        double q1 = P0[i] - ax * c0[i] / 2.0 + az * c2[i] / 2.0;
        double q2 = q1 + ax * c0[i];
        double q3 = q2 - az * c2[i];
        double q4 = q1 - az * c2[i];
        double r1 = P0[i] + ax * c0[i] / 2.0 + ay * c1[i] / 2.0;
        double r2 = r1 - ax * c0[i];
        double r3 = r2 - ay * c1[i];
        double r4 = r1 - ay * c1[i];
        V[0][i] = p1; V[1][i] = p2; V[2][i] = p3; V[3][i] = p4;
        V[4][i] = q1; V[5][i] = q2; V[6][i] = q3; V[7][i] = q4;
        V[8][i] = r1; V[9][i] = r2; V[10][i] = r3; V[11][i] = r4;
    }
    for (int v = 0; v < 12; ++v)
        if (V[v][2] > 0.0) return 1;
    return 0;
}

/* libcdm.py:39-247: displacements (ue, un, uv) at M points; returns 1 when the
 * source pierces the free surface (the reference raises ValueError there) */
int oracle_cdm(const double *X, const double *Y, int M, double X0, double Y0, double depth,
               double ax, double ay, double az, double omegaX, double omegaY, double omegaZ,
               double opening, double nu, double *ue, double *un, double *uv)
{
    double V[12][3];
    if (oracle_cdm_vertices(X0, Y0, depth, ax, ay, az, omegaX, omegaY, omegaZ, V)) return 1;
    double Ax = 2.0 * ax, Ay = 2.0 * ay, Az = 2.0 * az;
    for (int m = 0; m < M; ++m) {
        double u[3] = {0, 0, 0}, t[3];
        if (Ax == 0 && Ay == 0 && Az == 0) {
            /* zero field */
        } else if (Ax == 0 && Ay != 0 && Az != 0) {
            rd_disp_surf(X[m], Y[m], V[0], V[1], V[2], V[3], opening, nu, u);
        } else if (Ax != 0 && Ay == 0 && Az != 0) {
            rd_disp_surf(X[m], Y[m], V[4], V[5], V[6], V[7], opening, nu, u);
        } else if (Ax != 0 && Ay != 0 && Az == 0) {
            rd_disp_surf(X[m], Y[m], V[8], V[9], V[10], V[11], opening, nu, u);
        } else {
            rd_disp_surf(X[m], Y[m], V[0], V[1], V[2], V[3], opening, nu, t);
            for (int i = 0; i < 3; ++i) u[i] += t[i];
            rd_disp_surf(X[m], Y[m], V[4], V[5], V[6], V[7], opening, nu, t);
            for (int i = 0; i < 3; ++i) u[i] += t[i];
            rd_disp_surf(X[m], Y[m], V[8], V[9], V[10], V[11], opening, nu, t);
            for (int i = 0; i < 3; ++i) u[i] += t[i];
        }
        ue[m] = u[0]; un[m] = u[1]; uv[m] = u[2];
    }
    return 0;
}

/* CDM residuals for a sample set: models/cdm/cdm/Native.py:72-105 (LOS
 * projection of cdm/Source.py:76-88, minus the data and the dataset shift).
 * Layout: x, y, depth, opening, aX..aZ, omegaX..omegaZ at the given indices
 * (models/cdm/cdm/CDM.py:228-240).  Returns the number of samples whose
 * source pierces the free surface; their rows are filled with NaN.        */
int oracle_cdm_residuals(const double *theta, int C, int P,
                         int xIdx, int dIdx, int openingIdx, int aXIdx, int omegaXIdx, int offsetIdx,
                         const double *ox, const double *oy, const double *los,
                         const double *d, const int *oid, int N, double nu, double *res)
{
    int bad = 0;
    OMP_FOR
    for (int s = 0; s < C; ++s) {
        double *ue = (double *)malloc(sizeof(double) * N * 3);
        double *un = ue + N, *uv = ue + 2 * N;
        const double *t = theta + (size_t)s * P;
        int rc = oracle_cdm(ox, oy, N, t[xIdx], t[xIdx + 1], t[dIdx],
                            t[aXIdx], t[aXIdx + 1], t[aXIdx + 2],
                            t[omegaXIdx], t[omegaXIdx + 1], t[omegaXIdx + 2],
                            t[openingIdx], nu, ue, un, uv);
        for (int n = 0; n < N; ++n) {
            if (rc) { res[(size_t)s * N + n] = NAN; continue; }
            double u = ue[n] * los[3 * n] + un[n] * los[3 * n + 1] + uv[n] * los[3 * n + 2];
            u -= d[n];
            if (offsetIdx >= 0) u -= t[offsetIdx + oid[n]];
            res[(size_t)s * N + n] = u;
        }
        free(ue);
#ifdef ORACLE_OMP
#pragma omp atomic
#endif
        bad += rc;
    }
    return bad;
}

/* ------------------------------------------------------------------------ */
/* Linear model: models/linear/linear/Linear.py:151-185 — residuals
 * R = G theta^T - d, stored here as C x N (one row per sample).            */
void oracle_linear_residuals(const double *theta, int C, int P, const double *G /* N x P */,
                             const double *d, int N, double *res)
{
    OMP_FOR
    for (int s = 0; s < C; ++s) {
        const double *t = theta + (size_t)s * P;
        for (int n = 0; n < N; ++n) {
            double acc = 0.0;
            const double *g = G + (size_t)n * P;
            for (int p = 0; p < P; ++p) acc += g[p] * t[p];
            res[(size_t)s * N + n] = acc - d[n];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* L2 data log-likelihood: altar/altar/norms/L2.py:39-52 with the square of
 * models/mogi/mogi/Fast.py:84-92:  llk = normalization - |W r|^2 / 2, where W
 * is lower triangular (dtrmv Lower, NoTrans).  kind 0: W is N x N; kind 1:
 * W is the diagonal only (length N).                                       */
void oracle_l2_llk(const double *res, int C, int N, const double *W, int kind,
                   double normalization, double *llk)
{
    OMP_FOR
    for (int s = 0; s < C; ++s) {
        const double *r = res + (size_t)s * N;
        double acc = 0.0;
        if (kind == 1) {
            for (int i = 0; i < N; ++i) { double v = W[i] * r[i]; acc += v * v; }
        } else {
            for (int i = 0; i < N; ++i) {
                const double *w = W + (size_t)i * N;
                double v = 0.0;
                for (int j = 0; j <= i; ++j) v += w[j] * r[j];
                acc += v * v;
            }
        }
        llk[s] = normalization - acc / 2.0;
    }
}

/* ------------------------------------------------------------------------ */
/* Priors: altar/altar/distributions/Uniform.py:44-72 (verify, closed support)
 * and Base.py:60-79 (sum of log densities).  Each parameter j has kind
 * (0 uniform on [lo,hi], 1 gaussian mean lo, sigma hi).                    */
void oracle_prior(const double *theta, int C, int P, const int *kind,
                  const double *lo, const double *hi, double *prior, int *mask)
{
    for (int s = 0; s < C; ++s) {
        const double *t = theta + (size_t)s * P;
        double acc = 0.0;
        int bad = 0;
        for (int j = 0; j < P; ++j) {
            if (kind[j] == 0) {
                if (!(lo[j] <= t[j] && t[j] <= hi[j])) bad = 1;
                acc += log(1.0 / (hi[j] - lo[j]));
            } else {
                double z = (t[j] - lo[j]) / hi[j];
                acc += -0.5 * z * z - log(hi[j] * sqrt(2.0 * M_PI));
            }
        }
        prior[s] = acc;
        if (mask) mask[s] = bad;
    }
}

/* ------------------------------------------------------------------------ */
/* Metropolis: altar/altar/bayesian/Metropolis.py:93-111 (L = chol(s^2 Sigma))
 * and :236-259 (theta + (L xi)^T).  L is P x P lower, row-major.          */
void oracle_mh_propose(const double *theta, int C, int P, const double *L,
                       const double *xi, double *cand)
{
    for (int s = 0; s < C; ++s) {
        for (int i = 0; i < P; ++i) {
            double acc = 0.0;
            for (int j = 0; j <= i; ++j) acc += L[i * P + j] * xi[(size_t)s * P + j];
            cand[(size_t)s * P + i] = theta[(size_t)s * P + i] + acc;
        }
    }
}

/* Metropolis.py:189-223: returns accepted count; decisions[s] = 0 accepted,
 * 1 rejected (mask), 2 unlikely; state rows updated in place.             */
int oracle_mh_accept(double *theta, double *prior, double *data, double *post,
                     const double *ctheta, const double *cprior, const double *cdata,
                     const double *cpost, const int *mask, const double *dice,
                     int C, int P, int *decisions)
{
    int accepted = 0;
    for (int s = 0; s < C; ++s) {
        if (mask[s]) { decisions[s] = 1; continue; }
        double diff = cpost[s] - post[s];
        if (log(dice[s]) > diff) { decisions[s] = 2; continue; }
        decisions[s] = 0;
        ++accepted;
        memcpy(theta + (size_t)s * P, ctheta + (size_t)s * P, sizeof(double) * P);
        prior[s] = cprior[s];
        data[s] = cdata[s];
        post[s] = cpost[s];
    }
    return accepted;
}

/* Metropolis.py:262-281 */
double oracle_mh_adjust_scaling(double aw, double rw, long accepted, long rejected, long unlikely)
{
    double acceptance = (double)accepted / (double)(accepted + rejected + unlikely);
    double kc = (aw * acceptance + rw) / (aw + rw);
    if (kc < 0.1) kc = 0.1;
    if (kc > 1.0) kc = 1.0;
    return kc;
}

/* ------------------------------------------------------------------------ */
/* COV scheduler: altar/lib/libaltar/bayesian/COV.cc:334-368 (cov::cov) and
 * :222-326 (dbeta_grid).                                                   */
static double mean_of(const double *x, int n)
{
    double m = 0.0;
    for (int i = 0; i < n; ++i) m += x[i];
    return m / n;
}

double oracle_cov_eval(double dbeta, const double *llk, int C, double median, double target,
                       double *w, double *cov_out)
{
    for (int i = 0; i < C; ++i) w[i] = exp(dbeta * (llk[i] - median));
    double wsum = C * mean_of(w, C);
    for (int i = 0; i < C; ++i) w[i] *= 1.0 / wsum;
    double mean = mean_of(w, C);
    double ss = 0.0;
    for (int i = 0; i < C; ++i) { double dv = w[i] - mean; ss += dv * dv; }
    double sdev = sqrt(ss / (C - 1));
    double cov = sdev / mean;
    if (isinf(cov) || isnan(cov)) { *cov_out = 1e100; return 1e100; }
    *cov_out = cov;
    return (cov - target) * (cov - target);
}

/* returns dbeta; beta_io is the current beta (updated), cov_out the COV */
double oracle_dbeta_grid(const double *llk, int C, double median, double *w,
                         double *beta_io, double tolerance, double target, double *cov_out)
{
    double beta_low = 0.0;
    double beta_high = 1.0 - *beta_io;
    double beta_guess = 0.0 + 5.0e-5;
    double cov;
    oracle_cov_eval(beta_high, llk, C, median, target, w, &cov);
    if (cov < target || fabs(cov - target) < tolerance) {
        *beta_io = 1.0;
        *cov_out = cov;
        return beta_high;
    }
    oracle_cov_eval(beta_low, llk, C, median, target, w, &cov);
    oracle_cov_eval(beta_guess, llk, C, median, target, w, &cov);
    const int Nbeta = 10;
    int nbeta = 0, Nloop = 0, found = 0;
    double beta_step;
    do {
        ++Nloop;
        beta_step = (beta_high - beta_low) / Nbeta;
        for (beta_guess = beta_low, nbeta = 0; nbeta <= Nbeta; beta_guess += beta_step, ++nbeta) {
            oracle_cov_eval(beta_guess, llk, C, median, target, w, &cov);
            if (fabs(cov - target) < tolerance) { found = 1; break; }
            if (cov >= target) { if (nbeta == 0) found = 1; break; }
        }
        if (nbeta > Nbeta) beta_guess -= beta_step;
        beta_high = beta_guess;
        beta_low = beta_guess - beta_step;
    } while (Nloop <= Nbeta && !found);
    double dbeta = beta_guess;
    oracle_cov_eval(dbeta, llk, C, median, target, w, &cov);
    *cov_out = cov;
    *beta_io += dbeta;
    return dbeta;
}

/* COV.cc:58-219 (dbeta_brent): minimize f(dbeta) = (COV(dbeta) - target)^2
 * over [0, 1 - beta] from the guess 5e-5, stopping when the COV of the last
 * evaluated point is within tolerance of the target (gsl_root_test_residual)
 * or after maxiter iterations; then dbeta = the best point.  The reference
 * runs GSL's Brent minimizer (gsl_min_fminimizer_brent, GSL version not
 * pinned); this is Brent's method as published (Brent 1973, "Algorithms for
 * Minimization without Derivatives", ch. 5, localmin): golden-section steps
 * with parabolic interpolation through the three best points.  The skip to
 * beta = 1 is COV.cc's.                                                     */
static double brent_f(double dbeta, const double *llk, int C, double median, double target, double *w,
                      double *cov)
{
    return oracle_cov_eval(dbeta, llk, C, median, target, w, cov);
}

double oracle_dbeta_brent(const double *llk, int C, double median, double *w, double *beta_io,
                          double tolerance, double target, int maxiter, double *cov_out)
{
    double a = 0.0, b = 1.0 - *beta_io;
    double cov;
    brent_f(b, llk, C, median, target, w, &cov);
    if (cov < target || fabs(cov - target) < tolerance) {
        *beta_io = 1.0;
        *cov_out = cov;
        return b;
    }
    const double gold = 0.5 * (3.0 - sqrt(5.0));
    const double reps = 1.4901161193847656e-08, aeps = 1e-12;
    double x = 5.0e-5, fx = brent_f(x, llk, C, median, target, w, &cov);
    double v = x, w_ = x, fv = fx, fw = fx, d = 0.0, e = 0.0, last_cov = cov;
    int iter = 0;
    while (iter < maxiter && !(fabs(last_cov - target) < tolerance)) {
        ++iter;
        const double m = 0.5 * (a + b);
        const double tol1 = reps * fabs(x) + aeps, tol2 = 2.0 * tol1;
        if (fabs(x - m) <= tol2 - 0.5 * (b - a)) break;
        double p = 0.0, q = 0.0, r = 0.0;
        if (fabs(e) > tol1) {
            r = (x - w_) * (fx - fv);
            q = (x - v) * (fx - fw);
            p = (x - v) * q - (x - w_) * r;
            q = 2.0 * (q - r);
            if (q > 0.0) p = -p; else q = -q;
            r = e;
            e = d;
        }
        if (fabs(p) < fabs(0.5 * q * r) && p > q * (a - x) && p < q * (b - x)) {
            d = p / q;
            const double u = x + d;
            if (u - a < tol2 || b - u < tol2) d = x < m ? tol1 : -tol1;
        } else {
            e = x < m ? b - x : a - x;
            d = gold * e;
        }
        const double u = fabs(d) >= tol1 ? x + d : x + (d > 0.0 ? tol1 : -tol1);
        const double fu = brent_f(u, llk, C, median, target, w, &cov);
        last_cov = cov;
        if (fu <= fx) {
            if (u < x) b = x; else a = x;
            v = w_; fv = fw;
            w_ = x; fw = fx;
            x = u; fx = fu;
        } else {
            if (u < x) a = u; else b = u;
            if (fu <= fw || w_ == x) {
                v = w_; fv = fw;
                w_ = u; fw = fu;
            } else if (fu <= fv || v == x || v == w_) {
                v = u; fv = fu;
            }
        }
    }
    brent_f(x, llk, C, median, target, w, &cov);
    *cov_out = cov;
    *beta_io += x;
    return x;
}

/* altar/altar/bayesian/COV.py:105-156: Sigma = sum w theta theta^T - mean mean^T */
void oracle_weighted_covariance(const double *theta, const double *w, int C, int P,
                                double *mean, double *sigma)
{
    double wsum = 0.0;
    for (int i = 0; i < C; ++i) wsum += w[i];
    for (int j = 0; j < P; ++j) {
        double acc = 0.0;
        for (int i = 0; i < C; ++i) acc += w[i] * theta[(size_t)i * P + j];
        mean[j] = acc / wsum;
    }
    for (int a = 0; a < P; ++a)
        for (int b = 0; b < P; ++b) {
            double acc = 0.0;
            for (int i = 0; i < C; ++i) acc += w[i] * theta[(size_t)i * P + a] * theta[(size_t)i * P + b];
            sigma[a * P + b] = acc - mean[a] * mean[b];
        }
}

/* COV.py:222-251 multinomial multiplicities: r_i ~ U(0,1], bin edges
 * {0, cumsum(w)} with the last edge forced to 1 (so every draw lands in a
 * bin).  Bins are (edge_k, edge_{k+1}].                                     */
void oracle_multiplicities(const double *w, int C, const double *r, int *count)
{
    double *edge = (double *)malloc(sizeof(double) * (C + 1));
    edge[0] = 0.0;
    for (int i = 0; i < C; ++i) edge[i + 1] = edge[i] + w[i];
    edge[C] = 1.0;
    for (int i = 0; i < C; ++i) count[i] = 0;
    for (int i = 0; i < C; ++i) {
        /* smallest k with r <= edge[k+1] */
        int lo = 0, hi = C - 1;
        while (lo < hi) {
            int mid = (lo + hi) / 2;
            if (r[i] <= edge[mid + 1]) hi = mid; else lo = mid + 1;
        }
        count[lo] += 1;
    }
    free(edge);
}
