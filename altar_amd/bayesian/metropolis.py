"""Metropolis sampler (altar/altar/bayesian/Metropolis.py:21-290) on the GPU.

Per step: propose theta' = theta + (L xi)^T with L = chol(scaling^2 Sigma)
and xi from the Philox stream (stage, step); verify; revert rejected rows;
likelihoods; accept if log u <= post' - post.  Counters stay on the device
until the end of the walk.
"""
import numpy as np
import torch

from .. import _lib
from .cooling_step import CoolingStep


class Metropolis:
    def __init__(self, scaling=0.1, acceptanceWeight=8.0, rejectionWeight=1.0, steps=20):
        self.scaling = scaling
        self.acceptanceWeight = acceptanceWeight
        self.rejectionWeight = rejectionWeight
        self.steps = steps
        self.sigma_chol = None

    def initialize(self, application=None, job=None):
        if job is not None:
            self.steps = job.steps
        return self

    def samplePosterior(self, annealer, step):
        d = annealer.dispatcher
        d.notify(d.samplePosteriorStart, annealer)
        self.prepareSamplingPDF(annealer, step)
        stats = self.walkChains(annealer, step)
        d.notify(d.samplePosteriorFinish, annealer)
        return stats

    def resample(self, annealer, statistics):
        self.adjustCovarianceScaling(*statistics)

    def prepareSamplingPDF(self, annealer, step):
        """L = chol(scaling^2 Sigma), FP64 (Metropolis.py:93-111)."""
        d = annealer.dispatcher
        d.notify(d.prepareSamplingPDFStart, annealer)
        sigma = step.sigma.cpu().numpy() * self.scaling ** 2
        try:
            L = np.linalg.cholesky(sigma)
        except np.linalg.LinAlgError:
            # a degenerate sample set (e.g. a single surviving chain); the
            # reference has no conditioning (COV.py:212-219) and would fail here
            eps = 1e-12 * max(float(np.max(np.abs(np.diag(sigma)))), 1e-300)
            L = np.linalg.cholesky(sigma + eps * np.eye(sigma.shape[0]))
        self.sigma_chol = torch.from_numpy(np.ascontiguousarray(L)).to(step.theta.device)
        d.notify(d.prepareSamplingPDFFinish, annealer)

    def walkChains(self, annealer, step):
        """Metropolis.py:114-233; returns (accepted, rejected, unlikely)."""
        model = annealer.model
        d = annealer.dispatcher
        C, P = step.samples, step.parameters
        dev = step.theta.device
        seed = annealer.seed
        stage = annealer.worker.iteration
        f64 = lambda *s: torch.empty(s, dtype=torch.float64, device=dev)
        ws = getattr(self, "_ws", None)
        if ws is None or ws["cand"].shape != (C, P):
            ws = {"cand": f64(C, P), "xi": f64(C, P), "cprior": f64(C), "cdata": f64(C), "cpost": f64(C),
                  "mask": torch.empty(C, dtype=torch.int32, device=dev)}
            self._ws = ws
        counters = torch.zeros(3, dtype=torch.int64, device=dev)
        candidate = CoolingStep(step.beta, ws["cand"], (ws["cprior"], ws["cdata"], ws["cpost"]), sigma=step.sigma)
        s = _lib.stream()
        for k in range(self.steps):
            d.notify(d.chainAdvanceStart, annealer)
            _lib.call("altar_mh_propose", _lib.ptr(step.theta), C, P, _lib.ptr(self.sigma_chol), seed, stage, k,
                      _lib.ptr(ws["xi"]), _lib.ptr(ws["cand"]), s)
            d.notify(d.verifyStart, annealer)
            ws["mask"].zero_()
            model.verify(candidate, ws["mask"])
            _lib.call("altar_mh_restore", _lib.ptr(ws["cand"]), _lib.ptr(step.theta), _lib.ptr(ws["mask"]),
                      C, P, s)
            d.notify(d.verifyFinish, annealer)
            model.likelihoods(annealer, candidate)
            d.notify(d.acceptStart, annealer)
            _lib.call("altar_mh_accept", _lib.ptr(step.theta), _lib.ptr(step.prior), _lib.ptr(step.data),
                      _lib.ptr(step.posterior), _lib.ptr(ws["cand"]), _lib.ptr(ws["cprior"]),
                      _lib.ptr(ws["cdata"]), None, _lib.ptr(ws["mask"]), float(step.beta), C, P, seed, stage,
                      k, None, None, _lib.ptr(counters), s)
            d.notify(d.acceptFinish, annealer)
            d.notify(d.chainAdvanceFinish, annealer)
        acc, rej, unl = (int(v) for v in counters.cpu().tolist())
        return acc, rej, unl

    def adjustCovarianceScaling(self, accepted, rejected, unlikely):
        """Metropolis.py:262-281."""
        self.scaling = _lib.lib().altar_mh_scaling(self.acceptanceWeight, self.rejectionWeight,
                                                    accepted, rejected, unlikely)
        return self
