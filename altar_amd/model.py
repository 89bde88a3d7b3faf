"""The Model protocol and the Bayesian base model.

Mirrors altar/altar/models/Model.py:16-101 (the protocol the annealer calls)
and altar/altar/models/Bayesian.py:19-214 (posterior = prior + beta data,
``likelihoods`` = prior, data, posterior).  Sample state lives on the GPU as
FP64 torch tensors; every method dispatches to the C ABI of libaltar_amd.

``forwardModelBatched(theta) -> predicted`` is the batched forward model
(one row of LOS displacements per sample) that the data likelihood is built
on (Source::displacements, models/mogi/lib/libmogi/Source.cc:65-118).
"""
import ctypes
from collections import OrderedDict

import numpy as np
import torch

from . import _lib
from .distributions import flatten


class Model:
    """Protocol obligations (Model.py:24-89)."""

    def initialize(self, application=None):
        raise NotImplementedError

    def initializeSample(self, step):
        raise NotImplementedError

    def priorLikelihood(self, step):
        raise NotImplementedError

    def dataLikelihood(self, step):
        raise NotImplementedError

    def posteriorLikelihood(self, step):
        raise NotImplementedError

    def likelihoods(self, annealer, step):
        raise NotImplementedError

    def verify(self, step, mask):
        raise NotImplementedError

    def forwardModelBatched(self, theta):
        raise NotImplementedError

    def top(self, step):
        return self

    def bottom(self, step):
        return self


class Prior:
    """Device handle of the concatenated per-parameter distributions."""

    def __init__(self, kind, a, b):
        self.P = len(kind)
        h = ctypes.c_void_p()
        k = np.ascontiguousarray(kind, dtype=np.int32)
        aa = np.ascontiguousarray(a, dtype=np.float64)
        bb = np.ascontiguousarray(b, dtype=np.float64)
        _lib.call("altar_prior_create", ctypes.byref(h), self.P, _lib.ptr(k), _lib.ptr(aa), _lib.ptr(bb))
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().altar_prior_destroy(h)
            except Exception:
                pass
            self.handle = None


class Bayesian(Model):
    """Base model: parameter sets, priors, the posterior at a temperature.

    Subclasses provide ``strategy`` (an object with ``dataLikelihood(model,
    step)``), ``observations`` and, for models with geometric constraints,
    ``verifyModel(theta, mask)``.
    """

    offset = 0
    max_init_attempts = 100

    def __init__(self, psets, job=None, device="cuda"):
        self.psets = OrderedDict(psets)
        self.job = job
        self.device = torch.device(device)
        self.parameters = 0
        self.strategy = None
        self._prior = None
        self._prep = None

    # -- layout
    def initializeParameterSets(self):
        offset = 0
        for pset in self.psets.values():
            offset += pset.initialize(model=self, offset=offset)
        self.parameters = offset
        self._prior = Prior(*flatten(self.psets, "prior"))
        self._prep = Prior(*flatten(self.psets, "prep"))
        return self

    def restrict(self, theta):
        """My columns of the sample matrix (Bayesian.py:184-203)."""
        return theta[:, self.offset:self.offset + self.parameters]

    # -- protocol
    def initializeSample(self, step):
        """Fill step.theta from the prep distributions; rows the model rejects
        (e.g. a CDM source above the free surface) are drawn again, so the
        initial sample follows the prior truncated to valid geometries."""
        theta = step.theta
        C = theta.shape[0]
        seed = self.job.seed if self.job is not None else 0
        s = _lib.stream()
        _lib.call("altar_prior_sample", self._prep.handle, _lib.ptr(theta), C, seed, 0, 0, None, s)
        mask = torch.zeros(C, dtype=torch.int32, device=theta.device)
        for attempt in range(1, self.max_init_attempts + 1):
            mask.zero_()
            self.verifyModel(theta, mask)
            if int(mask.sum().item()) == 0:
                return self
            _lib.call("altar_prior_sample", self._prep.handle, _lib.ptr(theta), C, seed, 0, attempt,
                      _lib.ptr(mask), s)
        raise RuntimeError("initializeSample: could not draw valid samples from the prep distribution")

    def priorLikelihood(self, step):
        _lib.call("altar_prior_llk", self._prior.handle, _lib.ptr(step.theta), step.samples,
                  _lib.ptr(step.prior), _lib.stream())
        return self

    def dataLikelihood(self, step):
        self.strategy.dataLikelihood(model=self, step=step)
        return self

    def posteriorLikelihood(self, step):
        """posterior = prior + beta data (Bayesian.py:104-115)."""
        _lib.call("altar_posterior", _lib.ptr(step.prior), _lib.ptr(step.data), float(step.beta),
                  step.samples, _lib.ptr(step.posterior), _lib.stream())
        return self

    def likelihoods(self, annealer, step):
        """Bayesian.py:118-150 with the same event names."""
        d = getattr(annealer, "dispatcher", None)
        if d: d.notify(d.priorStart, annealer)
        self.priorLikelihood(step)
        if d: d.notify(d.priorFinish, annealer)
        if d: d.notify(d.dataStart, annealer)
        self.dataLikelihood(step)
        if d: d.notify(d.dataFinish, annealer)
        if d: d.notify(d.posteriorStart, annealer)
        self.posteriorLikelihood(step)
        if d: d.notify(d.posteriorFinish, annealer)
        return self

    def verify(self, step, mask):
        """Support of every parameter set, then the model's own constraints."""
        _lib.call("altar_prior_verify", self._prior.handle, _lib.ptr(step.theta), step.samples,
                  _lib.ptr(mask), _lib.stream())
        self.verifyModel(step.theta, mask)
        return mask

    def verifyModel(self, theta, mask):
        return mask
