"""COV annealing schedule (altar/altar/bayesian/COV.py:21-255, Grid.py:19-63).

Two placements of the stage:

* ``resident=True`` (default): the median, the dbeta grid search, the
  weights, the weighted covariance and the multinomial resampling all run on
  the GPU (altar_cov_update_dev, altar_cov_resample_dev); only the new
  temperature and COV come back to the host.
* ``resident=False``: the median and the grid search run in FP64 host C++
  (libaltar's dbeta_grid, COV.cc:222-368) on a copy of the data
  log-likelihoods, bit for bit as the oracle; covariance and gather on the GPU.

Solvers: ``Grid`` (the reference's default, Solver.py:45-53) and ``Brent``
(Brent.py, libaltar dbeta_brent, COV.cc:58-219; host FP64 in either
placement, its weights then feed the device stage).

Resampled rows are laid out in the order of the original samples (each
sample repeated by its multiplicity) rather than by decreasing multiplicity
(COV.py:159-208); the resampled set is the same multiset.
"""
import ctypes

import numpy as np
import torch

from .. import _lib


class Grid:
    """dbeta solver by iterative grid search (Grid.py:19-63)."""

    def __init__(self, tolerance=0.01, maxiter=1000):
        self.tolerance = tolerance
        self.maxiter = maxiter

    def initialize(self, application=None, scheduler=None):
        self.target = scheduler.target if scheduler is not None else 1.0
        return self

    def solve(self, llk, beta):
        """llk: host FP64 array.  Returns (new beta, cov, normalized weights)."""
        llk = np.ascontiguousarray(llk, dtype=np.float64)
        median = float(np.median(llk))
        w = np.empty_like(llk)
        b, c = ctypes.c_double(), ctypes.c_double()
        _lib.call("altar_cov_dbeta_grid", _lib.ptr(llk), len(llk), median, float(beta), self.tolerance,
                  self.target, _lib.ptr(w), ctypes.byref(b), ctypes.byref(c))
        return b.value, c.value, w


class Brent:
    """dbeta solver by Brent's minimization of (COV - target)^2 (Brent.py:19-63,
    libaltar dbeta_brent, COV.cc:58-219), FP64 on the host."""

    def __init__(self, tolerance=0.01, maxiter=1000):
        self.tolerance = tolerance
        self.maxiter = maxiter

    def initialize(self, application=None, scheduler=None):
        self.target = scheduler.target if scheduler is not None else 1.0
        return self

    def solve(self, llk, beta):
        """llk: host FP64 array.  Returns (new beta, cov, normalized weights)."""
        llk = np.ascontiguousarray(llk, dtype=np.float64)
        median = float(np.median(llk))
        w = np.empty_like(llk)
        b, c = ctypes.c_double(), ctypes.c_double()
        _lib.call("altar_cov_dbeta_brent", _lib.ptr(llk), len(llk), median, float(beta), self.tolerance,
                  self.target, int(self.maxiter), _lib.ptr(w), ctypes.byref(b), ctypes.byref(c))
        return b.value, c.value, w


class COV:
    def __init__(self, target=1.0, solver=None, resident=True):
        self.target = target
        self.solver = solver or Grid()
        self.resident = resident
        self.w = None
        self.cov = 0.0
        self._ws = None

    def initialize(self, application=None):
        self.solver.initialize(application=application, scheduler=self)
        return self

    def update(self, step, annealer=None):
        """COV.py:65-87."""
        beta = self.updateTemperature(step)
        sigma = self.computeCovariance(step)
        stage = annealer.worker.iteration if annealer is not None else 0
        seed = annealer.seed if annealer is not None else 0
        self.rank(step, seed, stage)
        step.beta = beta
        step.sigma.copy_(sigma)
        step.computePosterior()
        return step

    def _workspace(self, C, dev):
        ws = self._ws
        if ws is None or ws["w"].shape[0] != C:
            ws = {"w": torch.empty(C, dtype=torch.float64, device=dev),
                  "out": torch.zeros(3, dtype=torch.float64, device=dev),
                  "edges": torch.empty(C + 1, dtype=torch.float64, device=dev),
                  "count": torch.empty(C, dtype=torch.int32, device=dev),
                  "offset": torch.empty(C, dtype=torch.int32, device=dev),
                  "src": torch.empty(C, dtype=torch.int32, device=dev)}
            self._ws = ws
        return ws

    def updateTemperature(self, step):
        # the device stage implements the grid solver; another solver runs on
        # the host, and its weights go back to the device for the covariance
        # and the resampling
        if self.resident and isinstance(self.solver, Grid):
            ws = self._workspace(step.samples, step.data.device)
            _lib.call("altar_cov_update_dev", _lib.ptr(step.data), step.samples, float(step.beta),
                      self.solver.tolerance, self.target, _lib.ptr(ws["w"]), _lib.ptr(ws["out"]), _lib.stream())
            beta, cov, median = ws["out"].cpu().tolist()
            self.cov = cov
            self.w_dev = ws["w"]
            self.w = None
            return beta
        llk = step.data.cpu().numpy()
        beta, self.cov, self.w = self.solver.solve(llk, step.beta)
        self.w_dev = torch.from_numpy(self.w).to(step.data.device)
        if self.resident:
            self._workspace(step.samples, step.data.device)["w"].copy_(self.w_dev)
            self.w_dev = self._ws["w"]
        return beta

    def computeCovariance(self, step):
        """Sigma = sum_i w_i theta_i theta_i^T - mean mean^T (COV.py:105-156)."""
        C, P = step.samples, step.parameters
        dev = step.theta.device
        mean = torch.empty(P, dtype=torch.float64, device=dev)
        sigma = torch.empty((P, P), dtype=torch.float64, device=dev)
        _lib.call("altar_cov_covariance", _lib.ptr(step.theta), _lib.ptr(self.w_dev), C, P, _lib.ptr(mean),
                  _lib.ptr(sigma), _lib.stream())
        self.mean = mean
        return sigma

    def computeSampleMultiplicities(self, seed, stage):
        C = len(self.w)
        count = np.empty(C, dtype=np.int32)
        _lib.call("altar_cov_multiplicities", _lib.ptr(self.w), C, int(seed), int(stage), _lib.ptr(count))
        return count

    def rank(self, step, seed=0, stage=0):
        """Resample by the multiplicities (COV.py:159-251)."""
        C, P = step.samples, step.parameters
        if self.resident:
            ws = self._workspace(C, step.theta.device)
            _lib.call("altar_cov_resample_dev", _lib.ptr(self.w_dev), C, int(seed), int(stage), _lib.ptr(ws["edges"]),
                      _lib.ptr(ws["count"]), _lib.ptr(ws["offset"]), _lib.ptr(ws["src"]), _lib.stream())
            src_dev = ws["src"]
            count = None
        else:
            count = self.computeSampleMultiplicities(seed, stage)
            src = np.repeat(np.arange(C, dtype=np.int32), count)
            assert len(src) == C
            src_dev = torch.from_numpy(src).to(step.theta.device)
        old = (step.theta.clone(), step.prior.clone(), step.data.clone(), step.posterior.clone())
        _lib.call("altar_resample_gather", *(_lib.ptr(t) for t in old), _lib.ptr(src_dev), C, P,
                  _lib.ptr(step.theta), _lib.ptr(step.prior), _lib.ptr(step.data), _lib.ptr(step.posterior),
                  _lib.stream())
        self.multiplicities = count
        return step
