"""FP64 CPU oracle for the CATMIP inner loop (ctypes front end of oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``, always as the checker.  The
product package ``altar_amd`` never imports it.

Besides wrapping the C restatements, this module restates in numpy the pieces
the reference delegates to LAPACK through GSL:

* ``covariance_inverse_cholesky``: models/mogi/mogi/Mogi.py:343-356
  (chol(LU^-1(Cd)), the reference's whitening operator, lower triangular);
* ``exact_whitener``: inv(chol(Cd)), for which |W r|^2 = r^T Cd^-1 r exactly;
* ``normalization``: Mogi.py:327-340, -(N ln 2 pi + ln det Cd) / 2.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
# where the host-tuned baseline builds go (set by baseline_build)
_NATIVE_DIR = None

_d = ctypes.POINTER(ctypes.c_double)
_i = ctypes.POINTER(ctypes.c_int)
_u32 = ctypes.POINTER(ctypes.c_uint32)


def build():
    """Compile liboracle.so (and the OpenMP baseline variant) in place."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def baseline_build(outdir):
    """Compile the CPU-baseline builds for the host that runs them:
    gcc -O3 -march=native, one thread (liboracle_native.so) and OpenMP over
    chains (liboracle_native_omp.so), into outdir.  Afterwards the omp
    argument of the forward and likelihood wrappers selects "native" or
    "native_omp" as well."""
    global _NATIVE_DIR
    subprocess.run(["make", "-s", "-C", _HERE, "native", f"OUTDIR={os.path.abspath(outdir)}"], check=True)
    _NATIVE_DIR = os.path.abspath(outdir)
    for k in ("native", "native_omp"):
        _LIBS.pop(k, None)


def _declare(lib):
    c_int, c_double, c_u64, c_u32 = ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32
    sig = {
        "oracle_philox4x32_10": (None, [_u32, _u32, _u32]),
        "oracle_normals": (None, [c_u64, c_u32, c_u32, c_u32, c_int, c_int, _d]),
        "oracle_uniforms": (None, [c_u64, c_u32, c_u32, c_u32, c_int, _d]),
        "oracle_uniform_pairs": (None, [c_u64, c_u32, c_u32, c_u32, c_int, c_int, _d]),
        "oracle_mogi_residuals": (None, [_d, c_int, c_int, c_int, c_int, c_int, c_int,
                                         _d, _d, _d, _d, _i, c_int, c_double, _d]),
        "oracle_ang_dis_disp_surf": (None, [c_double] * 8 + [_d]),
        "oracle_cdm_vertices": (c_int, [c_double] * 9 + [_d]),
        "oracle_cdm": (c_int, [_d, _d, c_int] + [c_double] * 11 + [_d, _d, _d]),
        "oracle_cdm_residuals": (c_int, [_d, c_int, c_int] + [c_int] * 6 +
                                 [_d, _d, _d, _d, _i, c_int, c_double, _d]),
        "oracle_linear_residuals": (None, [_d, c_int, c_int, _d, _d, c_int, _d]),
        "oracle_l2_llk": (None, [_d, c_int, c_int, _d, c_int, c_double, _d]),
        "oracle_prior": (None, [_d, c_int, c_int, _i, _d, _d, _d, _i]),
        "oracle_mh_propose": (None, [_d, c_int, c_int, _d, _d, _d]),
        "oracle_mh_accept": (c_int, [_d, _d, _d, _d, _d, _d, _d, _d, _i, _d, c_int, c_int, _i]),
        "oracle_mh_adjust_scaling": (c_double, [c_double, c_double, ctypes.c_long,
                                                ctypes.c_long, ctypes.c_long]),
        "oracle_cov_eval": (c_double, [c_double, _d, c_int, c_double, c_double, _d, _d]),
        "oracle_dbeta_grid": (c_double, [_d, c_int, c_double, _d, _d, c_double, c_double, _d]),
        "oracle_dbeta_brent": (c_double, [_d, c_int, c_double, _d, _d, c_double, c_double, c_int, _d]),
        "oracle_weighted_covariance": (None, [_d, _d, c_int, c_int, _d, _d]),
        "oracle_multiplicities": (None, [_d, c_int, _d, _i]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def lib(omp=False):
    """The oracle library.  omp=False: the test build; True: the -O3 OpenMP
    build; "native" / "native_omp": the -O3 -march=native builds made by
    baseline_build for the CPU baseline."""
    key = {False: "plain", True: "omp"}.get(omp, omp)
    if key not in _LIBS:
        if key in ("native", "native_omp"):
            if _NATIVE_DIR is None:
                raise RuntimeError("baseline_build(outdir) has not run")
            path = os.path.join(_NATIVE_DIR, f"liboracle_{key}.so")
        else:
            path = os.path.join(_HERE, "liboracle_omp.so" if key == "omp" else "liboracle.so")
            if not os.path.exists(path):
                build()
        _LIBS[key] = _declare(ctypes.CDLL(path))
    return _LIBS[key]


def _p(a, t=_d):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- random streams
def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32).copy()
    k = np.asarray(key, dtype=np.uint32).copy()
    out = np.zeros(4, dtype=np.uint32)
    lib().oracle_philox4x32_10(_p(c, _u32), _p(k, _u32), _p(out, _u32))
    return out


def normals(seed, stage, step, tag, C, P):
    out = np.zeros((C, P))
    lib().oracle_normals(seed, stage, step, tag, C, P, _p(out))
    return out


def uniforms(seed, stage, step, tag, C):
    out = np.zeros(C)
    lib().oracle_uniforms(seed, stage, step, tag, C, _p(out))
    return out


def uniform_pairs(seed, stage, step, tag, C, W):
    out = np.zeros((C, W, 2))
    lib().oracle_uniform_pairs(seed, stage, step, tag, C, W, _p(out))
    return out


def prep_sample(seed, attempt, kinds, lo, hi, C):
    """Initial draws (uniform: a + (b - a)(1 - u1); gaussian: Box-Muller) from
    the init stream (stage 0, step = attempt, tag 3)."""
    kinds = np.asarray(kinds)
    P = len(kinds)
    u = uniform_pairs(seed, 0, attempt, 3, C, P)
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    uni = lo + (hi - lo) * (1.0 - u[:, :, 0])
    gau = lo + hi * np.sqrt(-2.0 * np.log(u[:, :, 0])) * np.cos(2.0 * np.pi * u[:, :, 1])
    return np.where(kinds == 0, uni, gau)


# ---------------------------------------------------------------- forward models
def mogi_residuals(theta, layout, ox, oy, los, d, oid, nu, omp=False):
    theta = _f64(theta)
    C, P = theta.shape
    N = len(ox)
    res = np.zeros((C, N))
    xIdx, dIdx, sIdx, offsetIdx = layout
    lib(omp).oracle_mogi_residuals(_p(theta), C, P, xIdx, dIdx, sIdx, offsetIdx,
                                   _p(_f64(ox)), _p(_f64(oy)), _p(_f64(los)), _p(_f64(d)),
                                   _p(_i32(oid), _i), N, nu, _p(res))
    return res


def cdm(X, Y, X0, Y0, depth, ax, ay, az, omegaX, omegaY, omegaZ, opening, nu):
    """libcdm.CDM with its argument order (libcdm.py:39); returns (ue, un, uv)."""
    X, Y = _f64(X), _f64(Y)
    M = len(X)
    ue, un, uv = np.zeros(M), np.zeros(M), np.zeros(M)
    rc = lib().oracle_cdm(_p(X), _p(Y), M, X0, Y0, depth, ax, ay, az, omegaX, omegaY, omegaZ,
                          opening, nu, _p(ue), _p(un), _p(uv))
    if rc:
        raise ValueError("Half-space solution: The CDM must be under the free surface!")
    return ue, un, uv


def cdm_vertices(X0, Y0, depth, ax, ay, az, omegaX, omegaY, omegaZ):
    V = np.zeros((12, 3))
    rc = lib().oracle_cdm_vertices(X0, Y0, depth, ax, ay, az, omegaX, omegaY, omegaZ, _p(V))
    return V, bool(rc)


def ang_dis_disp_surf(y1, y2, beta, b1, b2, b3, nu, a):
    v = np.zeros(3)
    lib().oracle_ang_dis_disp_surf(y1, y2, beta, b1, b2, b3, nu, a, _p(v))
    return v


def cdm_residuals(theta, layout, ox, oy, los, d, oid, nu, omp=False):
    """Rows of sources that pierce the free surface come back as NaN."""
    theta = _f64(theta)
    C, P = theta.shape
    N = len(ox)
    res = np.zeros((C, N))
    xIdx, dIdx, openingIdx, aXIdx, omegaXIdx, offsetIdx = layout
    lib(omp).oracle_cdm_residuals(_p(theta), C, P, xIdx, dIdx, openingIdx, aXIdx, omegaXIdx,
                                  offsetIdx, _p(_f64(ox)), _p(_f64(oy)), _p(_f64(los)),
                                  _p(_f64(d)), _p(_i32(oid), _i), N, nu, _p(res))
    return res


def linear_residuals(theta, G, d, omp=False):
    theta, G = _f64(theta), _f64(G)
    C, P = theta.shape
    N = G.shape[0]
    res = np.zeros((C, N))
    lib(omp).oracle_linear_residuals(_p(theta), C, P, _p(G), _p(_f64(d)), N, _p(res))
    return res


# ---------------------------------------------------------------- likelihood
def l2_llk(res, W, normalization, omp=False):
    """normalization - |W r|^2/2 per row; W is N x N lower triangular or a diagonal vector."""
    res = _f64(res)
    C, N = res.shape
    W = _f64(W)
    kind = 1 if W.ndim == 1 else 0
    llk = np.zeros(C)
    lib(omp).oracle_l2_llk(_p(res), C, N, _p(W), kind, normalization, _p(llk))
    return llk


def covariance_inverse_cholesky(Cd):
    """Mogi.py:343-356: chol(inverse(Cd)), lower triangular (reference whitening)."""
    return np.linalg.cholesky(np.linalg.inv(_f64(Cd)))


def exact_whitener(Cd):
    """inv(chol(Cd)), lower triangular: |W r|^2 = r^T Cd^-1 r."""
    L = np.linalg.cholesky(_f64(Cd))
    return np.linalg.solve(L, np.eye(L.shape[0]))


def normalization(Cd):
    """Mogi.py:327-340."""
    sign, logdet = np.linalg.slogdet(_f64(Cd))
    assert sign > 0
    N = Cd.shape[0]
    return -0.5 * (np.log(2 * np.pi) * N + logdet)


# ---------------------------------------------------------------- priors
def prior(theta, kinds, lo, hi):
    theta = _f64(theta)
    C, P = theta.shape
    out = np.zeros(C)
    mask = np.zeros(C, dtype=np.int32)
    lib().oracle_prior(_p(theta), C, P, _p(_i32(kinds), _i), _p(_f64(lo)), _p(_f64(hi)),
                       _p(out), _p(mask, _i))
    return out, mask


# ---------------------------------------------------------------- Metropolis
def mh_propose(theta, L, xi):
    theta, L, xi = _f64(theta), _f64(L), _f64(xi)
    C, P = theta.shape
    cand = np.zeros_like(theta)
    lib().oracle_mh_propose(_p(theta), C, P, _p(L), _p(xi), _p(cand))
    return cand


def mh_accept(theta, prior_, data, post, ctheta, cprior, cdata, cpost, mask, dice):
    """In place on (theta, prior_, data, post); returns (accepted, decisions)."""
    C, P = theta.shape
    dec = np.zeros(C, dtype=np.int32)
    acc = lib().oracle_mh_accept(_p(theta), _p(prior_), _p(data), _p(post), _p(_f64(ctheta)),
                                 _p(_f64(cprior)), _p(_f64(cdata)), _p(_f64(cpost)),
                                 _p(_i32(mask), _i), _p(_f64(dice)), C, P, _p(dec, _i))
    return acc, dec


def mh_adjust_scaling(accepted, rejected, unlikely, aw=8.0, rw=1.0):
    return lib().oracle_mh_adjust_scaling(aw, rw, accepted, rejected, unlikely)


# ---------------------------------------------------------------- COV
def median(llk):
    """GSL median of sorted data (altar/altar/bayesian/Grid.py:57): mean of the middle pair."""
    return float(np.median(llk))


def cov_eval(dbeta, llk, med, target=1.0):
    llk = _f64(llk)
    w = np.zeros_like(llk)
    cov = ctypes.c_double()
    metric = lib().oracle_cov_eval(dbeta, _p(llk), len(llk), med, target, _p(w),
                                   ctypes.byref(cov))
    return metric, cov.value, w


def dbeta_grid(llk, beta, tolerance=0.01, target=1.0):
    """Returns (new beta, cov, w) per COV.cc:222-326."""
    llk = _f64(llk)
    w = np.zeros_like(llk)
    b = ctypes.c_double(beta)
    cov = ctypes.c_double()
    lib().oracle_dbeta_grid(_p(llk), len(llk), median(llk), _p(w), ctypes.byref(b),
                            tolerance, target, ctypes.byref(cov))
    return b.value, cov.value, w


def dbeta_brent(llk, beta, tolerance=0.01, target=1.0, maxiter=1000):
    """Returns (new beta, cov, w) per COV.cc:58-219 (Brent's method)."""
    llk = _f64(llk)
    w = np.zeros_like(llk)
    b = ctypes.c_double(beta)
    cov = ctypes.c_double()
    lib().oracle_dbeta_brent(_p(llk), len(llk), median(llk), _p(w), ctypes.byref(b), tolerance, target, maxiter,
                             ctypes.byref(cov))
    return b.value, cov.value, w


def weighted_covariance(theta, w):
    theta = _f64(theta)
    C, P = theta.shape
    mean = np.zeros(P)
    sigma = np.zeros((P, P))
    lib().oracle_weighted_covariance(_p(theta), _p(_f64(w)), C, P, _p(mean), _p(sigma))
    return mean, sigma


def multiplicities(w, r):
    w = _f64(w)
    count = np.zeros(len(w), dtype=np.int32)
    lib().oracle_multiplicities(_p(w), len(w), _p(_f64(r)), _p(count, _i))
    return count
