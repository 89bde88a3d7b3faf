"""Synthetic inversion cases at the benchmark sizes (SI units: metres, m^3).

* ``mogi_case``: 1,000 LOS points on a 20 x 20 km square, source at (0, 0),
  depth 3 km, dV = 1e6 m^3, nu = 0.25, LOS (-0.38, 0.09, 0.92), sigma 5 mm,
  diagonal Cd, 4 + 1 parameters.
* ``linear_case``: G 4,000 x 400 ~ N(0, 1) 1e-2, slip ~ U(0, 2) m, full
  exponential Cd (sill 25 mm^2, range 5 km over random stations in a 50 km
  square), Gaussian prior N(0, 1).
* ``cdm_case``: 5,000 LOS points on a 100 x 50 lattice over 30 x 30 km,
  source at (1, -0.5) km, depth 4 km, rotations (5, -8, 30) deg, semi-axes
  (1.0, 1.2, 0.8) km, opening 0.1 m, sigma 8 mm with a full exponential Cd
  (sill 64 mm^2, range 5 km), 10 parameters.

Each returns a dict with observations, covariance, parameter sets and the
true parameter vector.  The CDM data are computed by the given forward model
(default: this package's GPU kernel).
"""
import numpy as np

from . import norms

LOS = np.array([-0.38, 0.09, 0.92])


def unit_los(N):
    n = LOS / np.linalg.norm(LOS)
    return np.tile(n, (N, 1))


def mogi_displacement(x, y, los, xs, ys, ds, logdv, nu=0.25, shift=0.0):
    """Closed form of Mogi (1958) as parametrised in libmogi/Source.cc:65-118."""
    dx, dy = xs - x, ys - y
    R = np.sqrt(dx * dx + dy * dy + ds * ds)
    C = (nu - 1.0) * 10.0 ** logdv / np.pi
    return (dx * los[:, 0] + dy * los[:, 1] - ds * los[:, 2]) * C / R ** 3 - shift


def mogi_case(N=1000, seed=20261203, sigma=5e-3):
    from .mogi import default_psets
    rng = np.random.default_rng(seed)
    x = rng.uniform(-10e3, 10e3, N)
    y = rng.uniform(-10e3, 10e3, N)
    los = unit_los(N)
    truth = np.array([0.0, 0.0, 3000.0, 6.0, 0.0])
    u = mogi_displacement(x, y, los, *truth[:4])
    d = u + sigma * rng.standard_normal(N)
    return dict(x=x, y=y, los=los, d=d, oid=np.zeros(N, np.int32), cd=np.full(N, sigma ** 2),
                psets=default_psets(), truth=truth, nu=0.25)


def linear_case(N=4000, P=400, seed=20261204, sill=25e-6, rng_length=5e3):
    from .linear import default_psets
    rng = np.random.default_rng(seed)
    G = rng.standard_normal((N, P)) * 1e-2
    truth = rng.uniform(0.0, 2.0, P)
    xy = rng.uniform(0.0, 50e3, (N, 2))
    cd = norms.exponential_covariance(xy, sill, rng_length)
    L = np.linalg.cholesky(cd)
    d = G @ truth + L @ rng.standard_normal(N)
    return dict(G=G, d=d, cd=cd, psets=default_psets(P), truth=truth, stations=xy)


def cdm_lattice(N=5000):
    nx, ny = (100, 50) if N == 5000 else (int(np.ceil(np.sqrt(2 * N))), 0)
    if N != 5000:
        ny = int(np.ceil(N / nx))
    gx = np.linspace(-15e3, 15e3, nx)
    gy = np.linspace(-15e3, 15e3, ny)
    X, Y = np.meshgrid(gx, gy)
    return X.ravel()[:N], Y.ravel()[:N]


CDM_TRUTH = np.array([1000.0, -500.0, 4000.0, 0.1, 1000.0, 1200.0, 800.0, 5.0, -8.0, 30.0])


def cdm_case(N=5000, seed=20261205, sigma=8e-3, sill=64e-6, rng_length=5e3, forward=None, full_cd=True):
    """``forward(theta (1 x 10), x, y, los) -> u (N)``; default: the GPU kernel."""
    from .cdm import default_psets
    rng = np.random.default_rng(seed)
    x, y = cdm_lattice(N)
    los = unit_los(N)
    truth = CDM_TRUTH.copy()
    if forward is None:
        forward = gpu_cdm_forward
    u = np.asarray(forward(truth[None, :], x, y, los)).reshape(-1)
    if full_cd:
        cd = norms.exponential_covariance(np.stack([x, y], 1), sill, rng_length)
        noise = np.linalg.cholesky(cd) @ rng.standard_normal(N)
    else:
        cd = np.full(N, sigma ** 2)
        noise = sigma * rng.standard_normal(N)
    return dict(x=x, y=y, los=los, d=u + noise, oid=np.zeros(N, np.int32), cd=cd,
                psets=default_psets(), truth=truth, nu=0.25)


def gpu_cdm_forward(theta, x, y, los, nu=0.25):
    import torch
    from . import _lib
    from .source import Source
    N = len(x)
    src = Source(_lib.CDM, theta.shape[1], N, nu)
    src.layout([0, 2, 3, 4, 7, -1, 0])
    src.observations(x, y, los, np.zeros(N), np.zeros(N, np.int32))
    t = torch.from_numpy(np.ascontiguousarray(theta, dtype=np.float64)).cuda()
    return src.forward(t, "fp32").cpu().numpy()
