"""Linear model d = G theta (models/linear/linear/Linear.py).

The reference forms R = G theta^T - d with dgemm and takes the whitened
norm of every column (Linear.py:151-185).  Two device paths:

* ``mode="direct"``: the forward GEMM writes R (FP32 MFMA), then the
  whitening GEMM reduces |W r|^2 — the reference's two steps;
* ``mode="prewhitened"`` (default): G' = W G and d' = W d are formed once in
  FP64 on the host, and one GEMM with a fused bias-and-square epilogue gives
  |G' theta - d'|^2 = |W (G theta - d)|^2.  Its flop count is 2 N P per
  evaluation instead of 2 N P + N^2.

The reference omits the square of the norm (Linear.py:182); here llk =
normalization - |W r|^2 / 2 as in the Mogi and CDM models.
"""
from collections import OrderedDict

import numpy as np

from . import _lib, norms
from .model import Bayesian
from .source import HIP, Source


class Linear(Bayesian):
    def __init__(self, psets, G, d, cd, mode="prewhitened", whiten="exact", precision="fp32",
                 job=None, device="cuda"):
        super().__init__(psets, job=job, device=device)
        self.G = np.asarray(G, np.float64)
        self.d = np.asarray(d, np.float64)
        self.cd = cd
        self.mode = mode
        self.whiten = whiten
        self.precision = precision
        self.observations = self.G.shape[0]

    def initialize(self, application=None):
        self.initializeParameterSets()
        assert self.parameters == self.G.shape[1], "parameter sets must cover the columns of G"
        self.normalization = norms.normalization(self.cd)
        kind, W = norms.whitener(self.cd, self.whiten)
        self.whitening_kind, self.W = kind, W
        self.source = Source(_lib.LINEAR, self.parameters, self.observations)
        if self.mode == "prewhitened":
            if kind == "diag":
                Gw, dw = self.G * W[:, None], self.d * W
            else:
                Gw, dw = W @ self.G, W @ self.d
            self.source.green(Gw, dw, prewhitened=True)
            self.source.whitening("diag", np.ones(self.observations), self.normalization)
        elif self.mode == "direct":
            self.source.green(self.G, self.d, prewhitened=False)
            if kind == "diag":
                Wfull = np.diag(W)
            else:
                Wfull = W
            self.source.whitening("lower", Wfull, self.normalization)
        else:
            raise ValueError(f"unknown linear mode {self.mode!r}")
        self.strategy = HIP().initialize(application=application, model=self)
        return self

    def forwardModelBatched(self, theta, precision=None):
        """G theta (C x N); in prewhitened mode this is W G theta."""
        return self.source.forward(self.restrict(theta), "fp32")


def default_psets(P, sigma=1.0):
    from .distributions import Contiguous, Gaussian
    ps = OrderedDict()
    ps["slip"] = Contiguous(P, Gaussian(0.0, sigma))
    return ps
