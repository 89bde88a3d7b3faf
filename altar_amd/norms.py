"""L2 norm with data-covariance whitening (altar/altar/norms/L2.py).

The reference whitens a residual r with L = chol(inverse(Cd)) (lower) and
takes |L r| (L2.py:39-52, cd_inv from models/mogi/mogi/Mogi.py:343-356).
|L r|^2 equals r^T Cd^-1 r only when Cd is diagonal; for a full Cd the
quadratic form needs L^T.  This module builds both operators as lower
triangular matrices W with llk = normalization - |W r|^2 / 2:

* ``"exact"`` (default): W = inverse(chol(Cd)), so |W r|^2 = r^T Cd^-1 r;
* ``"reference"``: W = chol(inverse(Cd)), the reference's operator.

For a diagonal Cd both are diag(1/sqrt(Cd_ii)).  These are one-time FP64
set-up factorizations on the host (LAPACK through numpy/scipy), as in the
reference; the per-sample whitening runs on the GPU.
"""
import numpy as np


def is_diagonal(cd):
    cd = np.asarray(cd)
    if cd.ndim == 1:
        return True
    return not np.any(cd - np.diag(np.diag(cd)))


def normalization(cd, chol=None):
    """-(N ln 2 pi + ln det Cd) / 2 (Mogi.py:327-340; Linear.py:315-328).
    ``chol`` may pass a Cholesky factor of ``cd`` already at hand."""
    cd = np.asarray(cd, dtype=np.float64)
    if cd.ndim == 1:
        logdet = float(np.sum(np.log(cd)))
        N = cd.shape[0]
    else:
        L = np.linalg.cholesky(cd) if chol is None else chol
        logdet = 2.0 * float(np.sum(np.log(np.diag(L))))
        N = cd.shape[0]
    return -0.5 * (np.log(2.0 * np.pi) * N + logdet)


def whitener(cd, mode="exact", chol=None):
    """(kind, W): kind "diag" with W the N inverse standard deviations, or
    "lower" with W an N x N lower-triangular FP64 matrix.  ``chol`` may pass
    a Cholesky factor of ``cd`` already at hand (exact mode)."""
    cd = np.asarray(cd, dtype=np.float64)
    if is_diagonal(cd):
        var = cd if cd.ndim == 1 else np.diag(cd)
        return "diag", 1.0 / np.sqrt(var)
    if mode == "exact":
        from scipy.linalg import lapack
        L = np.linalg.cholesky(cd) if chol is None else chol
        W, info = lapack.dtrtri(L, lower=1)
        if info != 0:
            raise np.linalg.LinAlgError(f"triangular inverse failed (info {info})")
        return "lower", np.ascontiguousarray(np.tril(W))
    if mode == "reference":
        return "lower", np.ascontiguousarray(np.linalg.cholesky(np.linalg.inv(cd)))
    raise ValueError(f"unknown whitening mode {mode!r}")


def exponential_covariance(xy, sill, rng_length, nugget=0.0):
    """Cd_ij = sill exp(-|s_i - s_j| / range) + nugget delta_ij over station
    positions xy (N x 2)."""
    from scipy.spatial.distance import cdist
    xy = np.asarray(xy, dtype=np.float64)
    cd = cdist(xy, xy)
    np.multiply(cd, -1.0 / rng_length, out=cd)
    np.exp(cd, out=cd)
    cd *= sill
    if nugget:
        cd[np.diag_indices_from(cd)] += nugget
    return cd
