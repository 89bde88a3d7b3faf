"""Compound dislocation model (models/cdm/cdm/CDM.py, Nikkhoo et al. 2017).

Parameter sets, in order: ``location`` (x, y), ``depth``, ``opening``, ``a``
(semi-axes ax, ay, az), ``omega`` (rotations about x, y, z in degrees) and
``offsets``, as in models/cdm/examples/cdm.pfg:19-82.  A source whose
rectangles reach above the free surface is rejected by ``verify`` (the
reference raises there, libcdm.py:219-224).
"""
from collections import OrderedDict

import numpy as np

from . import _lib
from .model import Bayesian
from .source import HIP, Source, set_covariance


class CDM(Bayesian):
    def __init__(self, psets, x, y, los, d, oid, cd, nu=0.25, whiten="exact", precision="fp32",
                 job=None, device="cuda", fp64_sides=0.2):
        super().__init__(psets, job=job, device=device)
        self.x, self.y = np.asarray(x, float), np.asarray(y, float)
        self.los, self.d = np.asarray(los, float), np.asarray(d, float)
        self.oid = np.asarray(oid, np.int32)
        self.cd = cd
        self.nu = nu
        self.whiten = whiten
        self.precision = precision
        # sides with |sin(beta)| below this are evaluated in FP64
        self.fp64_sides = fp64_sides
        self.observations = len(self.x)

    def initialize(self, application=None):
        self.initializeParameterSets()
        ps = self.psets
        self.xIdx = ps["location"].offset
        self.yIdx = self.xIdx + 1
        self.dIdx = ps["depth"].offset
        self.openingIdx = ps["opening"].offset
        self.aXIdx = ps["a"].offset
        self.omegaXIdx = ps["omega"].offset
        if "offsets" in ps and ps["offsets"].count > 0:
            self.offsetIdx, nsets = ps["offsets"].offset, ps["offsets"].count
        else:
            self.offsetIdx, nsets = -1, 0
        self.source = Source(_lib.CDM, self.parameters, self.observations, self.nu)
        self.source.layout([self.xIdx, self.dIdx, self.openingIdx, self.aXIdx, self.omegaXIdx,
                            self.offsetIdx, nsets])
        self.source.observations(self.x, self.y, self.los, self.d, self.oid)
        self.source.fp64_sides(self.fp64_sides)
        if self.cd is not None:
            self.whitening_kind, self.W, self.normalization = set_covariance(self.source, self.cd, self.whiten)
        self.strategy = HIP().initialize(application=application, model=self)
        return self

    def verifyModel(self, theta, mask):
        self.source.verify(self.restrict(theta), mask)
        return mask

    def forwardModelBatched(self, theta, precision=None):
        return self.source.forward(self.restrict(theta), precision or self.precision)


def default_psets(xy=10e3, depth=(1e3, 6e3), opening=(1e-3, 1.0), axes=(100.0, 3000.0),
                  omega=(-90.0, 90.0), offsets=0, offset_range=0.1):
    """Priors of the 5,000-point CDM inversion (10 parameters, no offsets)."""
    from .distributions import Contiguous, Uniform
    ps = OrderedDict()
    ps["location"] = Contiguous(2, Uniform((-xy, xy)))
    ps["depth"] = Contiguous(1, Uniform(depth))
    ps["opening"] = Contiguous(1, Uniform(opening))
    ps["a"] = Contiguous(3, Uniform(axes))
    ps["omega"] = Contiguous(3, Uniform(omega))
    if offsets:
        ps["offsets"] = Contiguous(offsets, Uniform((-offset_range, offset_range)))
    return ps
