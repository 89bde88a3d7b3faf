"""COV annealing schedule (altar/altar/bayesian/COV.py:21-255, Grid.py:19-63).

updateTemperature: the median of the data log-likelihoods and the grid
search for dbeta run in FP64 host C++ (libaltar's dbeta_grid,
COV.cc:222-368).  computeCovariance and the resampling gather run on the
GPU.  Resampled rows are laid out in the order of the original samples
(each sample repeated by its multiplicity) rather than by decreasing
multiplicity (COV.py:159-208); the resampled set is the same multiset.
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


class COV:
    def __init__(self, target=1.0, solver=None):
        self.target = target
        self.solver = solver or Grid()
        self.w = None
        self.cov = 0.0

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

    def updateTemperature(self, step):
        llk = step.data.cpu().numpy()
        beta, self.cov, self.w = self.solver.solve(llk, step.beta)
        return beta

    def computeCovariance(self, step):
        """Sigma = sum_i w_i theta_i theta_i^T - mean mean^T (COV.py:105-156)."""
        C, P = step.samples, step.parameters
        dev = step.theta.device
        self.w_dev = torch.from_numpy(self.w).to(dev)
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
