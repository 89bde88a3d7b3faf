"""Device forward-model handle and the HIP likelihood strategy.

``Source`` wraps one ``altar_source`` of the C ABI (the counterpart of the
reference's libmogi/libcdm/libcudamogi capsules, models/mogi/ext/mogi/
source.cc:44-571).  ``HIP`` is the data-likelihood strategy that the models
select, beside the reference's Native/Fast/CUDA strategies
(models/mogi/mogi/Mogi.py:122-152).
"""
import ctypes

import numpy as np
import torch

from . import _lib, norms


class Source:
    def __init__(self, model, P, N, nu=0.25):
        h = ctypes.c_void_p()
        _lib.call("altar_source_create", ctypes.byref(h), int(model), int(P), int(N), float(nu))
        self.handle = h
        self.model, self.P, self.N = model, P, N

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().altar_source_destroy(h)
            except Exception:
                pass
            self.handle = None

    def layout(self, idx):
        arr = np.ascontiguousarray(idx, dtype=np.int32)
        _lib.call("altar_source_layout", self.handle, _lib.ptr(arr), len(arr))

    def observations(self, x, y, los, d, oid):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        x, y, los, d = f(x), f(y), f(los), f(d)
        oid = np.ascontiguousarray(oid, dtype=np.int32)
        assert los.shape == (self.N, 3) and x.shape == (self.N,) and oid.shape == (self.N,)
        _lib.call("altar_source_observations", self.handle, _lib.ptr(x), _lib.ptr(y), _lib.ptr(los),
                  _lib.ptr(d), _lib.ptr(oid))

    def green(self, G, d, prewhitened=False):
        G = np.ascontiguousarray(G, dtype=np.float64)
        d = np.ascontiguousarray(d, dtype=np.float64)
        assert G.shape == (self.N, self.P) and d.shape == (self.N,)
        _lib.call("altar_source_green", self.handle, _lib.ptr(G), _lib.ptr(d), int(bool(prewhitened)))

    def whitening(self, kind, W, normalization):
        W = np.ascontiguousarray(W, dtype=np.float64)
        code = {"diag": _lib.WHITEN_DIAG, "lower": _lib.WHITEN_LOWER}[kind]
        _lib.call("altar_source_whitening", self.handle, code, _lib.ptr(W), float(normalization))

    def whitening_blocks(self, offsets, blocks, normalization):
        """Block-diagonal whitener: block b is the lower-triangular ``blocks[b]``
        over the observations starting at ``offsets[b]`` (a multiple of 128)."""
        off = np.ascontiguousarray(offsets, dtype=np.int32)
        sizes = np.ascontiguousarray([np.asarray(W).shape[0] for W in blocks], dtype=np.int32)
        W = np.concatenate([np.ascontiguousarray(B, dtype=np.float64).ravel() for B in blocks])
        _lib.call("altar_source_whitening_blocks", self.handle, len(blocks), _lib.ptr(off), _lib.ptr(sizes),
                  _lib.ptr(W), float(normalization))

    def _rows(self, theta, precision, fn):
        theta = theta.contiguous()
        C = theta.shape[0]
        out = torch.empty((C, self.N), dtype=torch.float64, device=theta.device)
        _lib.call(fn, self.handle, _lib.ptr(theta), C, _lib.ptr(out), self.N,
                  _lib.PRECISIONS[precision], _lib.stream())
        return out

    def forward(self, theta, precision="fp32"):
        """Predicted displacements (minus the dataset shifts), C x N FP64."""
        return self._rows(theta, precision, "altar_source_forward")

    def residuals(self, theta, precision="fp32"):
        return self._rows(theta, precision, "altar_source_residuals")

    def data_llk(self, theta, out, precision="fp32"):
        theta = theta.contiguous()
        _lib.call("altar_source_data_llk", self.handle, _lib.ptr(theta), theta.shape[0], _lib.ptr(out),
                  _lib.PRECISIONS[precision], _lib.stream())
        return out

    def fp64_sides(self, threshold):
        """CDM: sides with |sin(beta)| below threshold are evaluated in FP64."""
        _lib.call("altar_source_fp64_sides", self.handle, float(threshold))

    def side_census(self, C):
        """CDM: (FP32 sides, FP64 sides, skipped sides, invalid chains) of the
        C chains prepared by the last call."""
        counts = (ctypes.c_longlong * 4)()
        _lib.call("altar_source_side_census", self.handle, int(C), counts, _lib.stream())
        return tuple(counts)

    def verify(self, theta, mask):
        theta = theta.contiguous()
        _lib.call("altar_source_verify", self.handle, _lib.ptr(theta), theta.shape[0], _lib.ptr(mask),
                  _lib.stream())
        return mask


class HIP:
    """Data-likelihood strategy on the GPU (Native.py/Fast.py/CUDA.py's role)."""

    def initialize(self, application=None, model=None):
        return self

    def dataLikelihood(self, model, step):
        theta = model.restrict(step.theta)
        model.source.data_llk(theta, step.data, model.precision)
        return self


def set_covariance(source, cd, mode="exact"):
    """Whitening operator and normalization of a source from its Cd."""
    kind, W = norms.whitener(cd, mode)
    norm = norms.normalization(cd)
    source.whitening(kind, W, norm)
    return kind, W, norm


def los_from_angles(theta, phi):
    """LOS unit vectors from incidence/azimuth angles (Mogi.py:112-117)."""
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    return np.stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)], axis=1)
