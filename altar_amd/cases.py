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

* ``joint_case``: a Mogi source and a CDM observed by an ascending and a
  descending InSAR track (20,000 points each on offset lattices over
  40 x 40 km; exponential Cd, sill 48 mm^2, range 5 km, plus a 16 mm^2
  nugget) and 200 three-component GNSS stations (sigma 2 mm horizontal,
  5 mm vertical, diagonal); 14 source parameters and 3 dataset offsets.

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


LOS_ASC = np.array([-0.60, -0.11, 0.79])
LOS_DESC = np.array([0.60, -0.11, 0.79])
JOINT_TRUTH = np.array([-6000.0, 4000.0, 4000.0, 6.7,                       # Mogi x, y, depth, log10 dV
                        5000.0, -3000.0, 4000.0, 0.1, 1000.0, 1200.0, 800.0,  # CDM x, y, depth, opening, a
                        5.0, -8.0, 30.0,                                      # CDM omega
                        0.01, -0.02, 0.0])                                    # dataset offsets


def joint_case(n_insar=20000, n_gnss=200, seed=20261207, forward=None):
    """Datasets (ascending, descending, GNSS) and truth of the joint case.
    ``forward(theta (1 x 17), x, y, los, oid) -> u (N)`` gives the noise-free
    prediction (default: this package's GPU kernels)."""
    from .joint import Dataset, default_psets
    rng = np.random.default_rng(seed)
    nx = int(np.ceil(np.sqrt(2 * n_insar)))
    ny = int(np.ceil(n_insar / nx))
    gx, gy = np.linspace(-20e3, 20e3, nx), np.linspace(-20e3, 20e3, ny)
    X, Y = np.meshgrid(gx, gy)
    xa, ya = X.ravel()[:n_insar], Y.ravel()[:n_insar]
    xd, yd = xa + 0.5 * (gx[1] - gx[0]), ya + 0.5 * (gy[1] - gy[0])
    sg = rng.uniform(-25e3, 25e3, (n_gnss, 2))
    xg, yg = np.repeat(sg[:, 0], 3), np.repeat(sg[:, 1], 3)
    lg = np.tile(np.eye(3), (n_gnss, 1))
    parts = [(xa, ya, np.tile(LOS_ASC / np.linalg.norm(LOS_ASC), (n_insar, 1))),
             (xd, yd, np.tile(LOS_DESC / np.linalg.norm(LOS_DESC), (n_insar, 1))),
             (xg, yg, lg)]
    truth = JOINT_TRUTH.copy()
    if forward is None:
        forward = gpu_joint_forward
    datasets = []
    for k, (x, y, los) in enumerate(parts):
        u = np.asarray(forward(truth[None, :], x, y, los, np.full(len(x), k, np.int32))).reshape(-1)
        chol = None
        if k < 2:
            cd = norms.exponential_covariance(np.stack([x, y], 1), 48e-6, 5e3, nugget=16e-6)
            chol = np.linalg.cholesky(cd)
            noise = chol @ rng.standard_normal(len(x))
        else:
            cd = np.tile([4e-6, 4e-6, 25e-6], n_gnss)
            noise = np.sqrt(cd) * rng.standard_normal(len(x))
        datasets.append(Dataset(x=x, y=y, los=los, d=u + noise, cd=cd, name=("asc", "desc", "gnss")[k],
                                chol=chol))
    return dict(datasets=datasets, psets=default_psets(), truth=truth, nu=0.25)


def gpu_joint_forward(theta, x, y, los, oid, nu=0.25):
    """Mogi + CDM prediction minus the dataset shift (parameters in the
    joint layout), on the GPU."""
    import torch
    from . import _lib
    from .source import Source
    N = len(x)
    src = Source(_lib.JOINT, theta.shape[1], N, nu)
    src.layout([0, 2, 3, 4, 6, 7, 8, 11, 14, 3])
    src.observations(x, y, los, np.zeros(N), oid)
    t = torch.from_numpy(np.ascontiguousarray(theta, dtype=np.float64)).cuda()
    return src.forward(t, "fp32").cpu().numpy()
