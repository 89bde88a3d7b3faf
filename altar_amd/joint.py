"""Joint Mogi + CDM model over several datasets with a block-diagonal Cd.

The reference's attempt at combining models (altar/altar/models/Ensemble.py
:17-98) cannot run, so this model has no working counterpart.  The
prediction at an observation is the sum of a Mogi point source
(models/mogi/mogi/Mogi.py) and a compound dislocation
(models/cdm/cdm/CDM.py), minus the shift of the observation's dataset.
That superposition is exact in linear elasticity.

Parameter sets, in order:
* ``mogi.location`` (x, y), ``mogi.depth``, ``mogi.source`` (log10 dV);
* ``cdm.location`` (x, y), ``cdm.depth``, ``cdm.opening``, ``cdm.a``
  (semi-axes), ``cdm.omega`` (rotations, degrees);
* ``offsets``: one shift per dataset.

Each dataset (an InSAR track, the GNSS components) has its own covariance
and is whitened by its own block.  On the device every dataset starts at a
multiple of 128 observations.  Padding observations fill the gaps: they lie
far from both sources, have a zero LOS vector and lie outside every
whitening block, so they never enter the likelihood.
"""
from collections import OrderedDict
from dataclasses import dataclass

import numpy as np

from . import _lib, norms
from .model import Bayesian
from .source import HIP, Source

PAD_XY = 1.0e7   # metres: where padding observations sit


@dataclass
class Dataset:
    """Observations of one dataset: positions (N), LOS unit vectors (N x 3),
    data (N) and covariance (N x N, or N variances for a diagonal Cd)."""
    x: np.ndarray
    y: np.ndarray
    los: np.ndarray
    d: np.ndarray
    cd: np.ndarray
    name: str = ""
    chol: np.ndarray = None   # Cholesky factor of a full cd, when already computed


def layout_datasets(datasets):
    """Padded observation arrays and the block offsets: (x, y, los, d, oid,
    offsets, sizes, real), where ``real`` indexes the true observations."""
    xs, ys, ls, ds, ids, offsets, sizes, real = [], [], [], [], [], [], [], []
    at = 0
    for k, ds_ in enumerate(datasets):
        n = len(ds_.x)
        start = (at + 127) // 128 * 128
        pad = start - at
        if pad:
            xs.append(PAD_XY + np.arange(pad, dtype=float))
            ys.append(np.full(pad, PAD_XY))
            ls.append(np.zeros((pad, 3)))
            ds.append(np.zeros(pad))
            ids.append(np.full(pad, max(k - 1, 0), np.int32))
        xs.append(np.asarray(ds_.x, float))
        ys.append(np.asarray(ds_.y, float))
        ls.append(np.asarray(ds_.los, float).reshape(n, 3))
        ds.append(np.asarray(ds_.d, float))
        ids.append(np.full(n, k, np.int32))
        offsets.append(start)
        sizes.append(n)
        real.append(np.arange(start, start + n))
        at = start + n
    cat = np.concatenate
    return (cat(xs), cat(ys), cat(ls), cat(ds), cat(ids), np.array(offsets), np.array(sizes), cat(real))


class Joint(Bayesian):
    def __init__(self, psets, datasets, nu=0.25, whiten="exact", precision="fp32", job=None, device="cuda",
                 fp64_sides=0.2):
        super().__init__(psets, job=job, device=device)
        self.datasets = list(datasets)
        self.nu = nu
        self.whiten = whiten
        self.precision = precision
        self.fp64_sides = fp64_sides
        (self.x, self.y, self.los, self.d, self.oid, self.block_offsets, self.block_sizes,
         self.real) = layout_datasets(self.datasets)
        self.observations = len(self.x)          # padded count on the device
        self.data_count = len(self.real)          # true observations

    def initialize(self, application=None):
        self.initializeParameterSets()
        ps = self.psets
        mx, md, ms = ps["mogi.location"].offset, ps["mogi.depth"].offset, ps["mogi.source"].offset
        cx, cd_, co = ps["cdm.location"].offset, ps["cdm.depth"].offset, ps["cdm.opening"].offset
        ca, cw = ps["cdm.a"].offset, ps["cdm.omega"].offset
        if "offsets" in ps and ps["offsets"].count > 0:
            off, nsets = ps["offsets"].offset, ps["offsets"].count
            assert nsets == len(self.datasets), "one offset per dataset"
        else:
            off, nsets = -1, 0
        self.source = Source(_lib.JOINT, self.parameters, self.observations, self.nu)
        self.source.layout([mx, md, ms, cx, cd_, co, ca, cw, off, nsets])
        oid = self.oid if nsets else np.zeros_like(self.oid)
        self.source.observations(self.x, self.y, self.los, self.d, oid)
        self.source.fp64_sides(self.fp64_sides)
        blocks, norm = [], 0.0
        for ds_ in self.datasets:
            chol = ds_.chol if self.whiten == "exact" else None
            kind, W = norms.whitener(ds_.cd, self.whiten, chol=chol)
            blocks.append(np.diag(W) if kind == "diag" else W)
            norm += norms.normalization(ds_.cd, chol=ds_.chol)
        self.normalization = norm
        self.source.whitening_blocks(self.block_offsets, blocks, norm)
        self.strategy = HIP().initialize(application=application, model=self)
        return self

    def verifyModel(self, theta, mask):
        self.source.verify(self.restrict(theta), mask)
        return mask

    def forwardModelBatched(self, theta, precision=None):
        """Predicted displacements minus the dataset shifts at the true
        observations (datasets in order), C x N FP64."""
        full = self.source.forward(self.restrict(theta), precision or self.precision)
        return full[:, self.real]


def default_psets(xy=20e3, mogi_depth=(500.0, 10e3), logdv=(5.0, 8.0), cdm_depth=(1e3, 6e3),
                  opening=(1e-3, 1.0), axes=(100.0, 3000.0), omega=(-90.0, 90.0), datasets=3,
                  offset_range=0.1):
    """Priors of the joint inversion: 14 source parameters and one offset per dataset."""
    from .distributions import Contiguous, Uniform
    ps = OrderedDict()
    ps["mogi.location"] = Contiguous(2, Uniform((-xy, xy)))
    ps["mogi.depth"] = Contiguous(1, Uniform(mogi_depth))
    ps["mogi.source"] = Contiguous(1, Uniform(logdv))
    ps["cdm.location"] = Contiguous(2, Uniform((-xy, xy)))
    ps["cdm.depth"] = Contiguous(1, Uniform(cdm_depth))
    ps["cdm.opening"] = Contiguous(1, Uniform(opening))
    ps["cdm.a"] = Contiguous(3, Uniform(axes))
    ps["cdm.omega"] = Contiguous(3, Uniform(omega))
    if datasets:
        ps["offsets"] = Contiguous(datasets, Uniform((-offset_range, offset_range)))
    return ps
