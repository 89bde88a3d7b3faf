"""Mogi (1958) point pressure source (models/mogi/mogi/Mogi.py).

Parameter sets, in order: ``location`` (x, y), ``depth``, ``source``
(log10 of the volume change, libmogi/Source.cc:78) and ``offsets`` (one
shift per dataset), as in models/mogi/examples/mogi.pfg:34-72.
"""
from collections import OrderedDict

import numpy as np

from . import _lib
from .model import Bayesian
from .source import HIP, Source, set_covariance


class Mogi(Bayesian):
    def __init__(self, psets, x, y, los, d, oid, cd, nu=0.25, whiten="exact", precision="fp32",
                 job=None, device="cuda"):
        super().__init__(psets, job=job, device=device)
        self.x, self.y = np.asarray(x, float), np.asarray(y, float)
        self.los, self.d = np.asarray(los, float), np.asarray(d, float)
        self.oid = np.asarray(oid, np.int32)
        self.cd = cd
        self.nu = nu
        self.whiten = whiten
        self.precision = precision
        self.observations = len(self.x)

    def initialize(self, application=None):
        self.initializeParameterSets()
        ps = self.psets
        self.xIdx = ps["location"].offset
        self.yIdx = self.xIdx + 1
        self.dIdx = ps["depth"].offset
        self.sIdx = ps["source"].offset
        if "offsets" in ps and ps["offsets"].count > 0:
            self.offsetIdx, nsets = ps["offsets"].offset, ps["offsets"].count
        else:
            self.offsetIdx, nsets = -1, 0
        self.source = Source(_lib.MOGI, self.parameters, self.observations, self.nu)
        self.source.layout([self.xIdx, self.dIdx, self.sIdx, self.offsetIdx, nsets])
        self.source.observations(self.x, self.y, self.los, self.d, self.oid)
        self.whitening_kind, self.W, self.normalization = set_covariance(self.source, self.cd, self.whiten)
        self.strategy = HIP().initialize(application=application, model=self)
        return self

    def forwardModelBatched(self, theta, precision=None):
        return self.source.forward(self.restrict(theta), precision or self.precision)


def default_psets(xy=10e3, depth=(500.0, 8000.0), logdv=(5.0, 7.0), offsets=1, offset_range=0.1):
    from .distributions import Contiguous, Uniform
    ps = OrderedDict()
    ps["location"] = Contiguous(2, Uniform((-xy, xy)))
    ps["depth"] = Contiguous(1, Uniform(depth))
    ps["source"] = Contiguous(1, Uniform(logdv))
    if offsets:
        ps["offsets"] = Contiguous(offsets, Uniform((-offset_range, offset_range)))
    return ps
