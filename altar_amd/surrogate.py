"""Learned surrogate of the forward model for delayed-acceptance screening.

No reference counterpart: an addition to CATMIP.  A
dense network (10 -> 256 -> 256 -> 256 -> 64, GELU, BF16 weights, FP32
accumulation; kernels in csrc/mlp.hip) maps source parameters to the leading
principal components of the *whitened* residual field W r(theta), where W is
the likelihood's whitener.  In that space the surrogate log-likelihood is a
64-term sum:

    |W r|^2 ~= cperp + sum_k (e_k + s_k z_k)^2,
    e = V^T mean,  cperp = |(I - V V^T) mean|^2,

with V the orthonormal PCA basis, ``mean`` the mean whitened residual of the
training set, s the coefficient scales and z the network output.  The
training fields are the whitened residuals that the annealer's own exact
evaluations produce (altar_source_data_llk_fields), so no extra forward
model runs are needed to build the training set.

The network is trained on the GPU by its own forward and backward kernels
(Adam).  Validation reports the RMSE of the predicted LOS field in metres on
held-out samples; the sampler falls back to the exact path above a bound.
"""
import ctypes

import numpy as np
import torch

from . import _lib


class MLP:
    """Handle of an altar_mlp network."""

    def __init__(self, nin, nhid=256, nlayers=4, nout=64, max_batch=16384):
        h = ctypes.c_void_p()
        _lib.call("altar_mlp_create", ctypes.byref(h), nin, nhid, nlayers, nout, max_batch)
        self.handle = h
        self.nin, self.nhid, self.nlayers, self.nout, self.max_batch = nin, nhid, nlayers, nout, max_batch
        self.shapes = [((nhid if l < nlayers - 1 else nout), (nin if l == 0 else nhid)) for l in range(nlayers)]

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                _lib.lib().altar_mlp_destroy(h)
            except Exception:
                pass
            self.handle = None

    def init(self, seed=0):
        """He-normal weights (GELU layers), LeCun-normal output layer, zero biases."""
        rng = np.random.default_rng(seed)
        for l, (o, k) in enumerate(self.shapes):
            gain = 2.0 if l < self.nlayers - 1 else 1.0
            self.set_layer(l, rng.normal(0.0, np.sqrt(gain / k), (o, k)), np.zeros(o))
        return self

    def input_scaling(self, lo, hi):
        lo = np.ascontiguousarray(lo, dtype=np.float64)
        hi = np.ascontiguousarray(hi, dtype=np.float64)
        _lib.call("altar_mlp_input_scaling", self.handle, _lib.ptr(lo), _lib.ptr(hi))

    def set_layer(self, l, W, b):
        W = np.ascontiguousarray(W, dtype=np.float32)
        b = np.ascontiguousarray(b, dtype=np.float32)
        assert W.shape == self.shapes[l] and b.shape == (self.shapes[l][0],)
        _lib.call("altar_mlp_set_layer", self.handle, l, _lib.ptr(W), _lib.ptr(b))

    def layer(self, l, grads=False):
        W = np.empty(self.shapes[l], dtype=np.float32)
        b = np.empty(self.shapes[l][0], dtype=np.float32)
        _lib.call("altar_mlp_get_layer", self.handle, l, _lib.ptr(W), _lib.ptr(b), int(grads))
        return W, b

    def forward(self, theta, out=None):
        theta = theta.contiguous()
        B = theta.shape[0]
        if out is None:
            out = torch.empty((B, self.nout), dtype=torch.float32, device=theta.device)
        _lib.call("altar_mlp_forward", self.handle, _lib.ptr(theta), B, _lib.ptr(out), out.shape[1], _lib.stream())
        return out

    def backward(self, theta, target, weights, loss=None):
        theta = theta.contiguous()
        _lib.call("altar_mlp_backward", self.handle, _lib.ptr(theta), _lib.ptr(target), target.shape[1],
                  _lib.ptr(weights), theta.shape[0], _lib.ptr(loss), _lib.stream())

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        _lib.call("altar_mlp_adam", self.handle, float(lr), float(beta1), float(beta2), float(eps), _lib.stream())


def whitened_pca(wr, K, oversample=32, iterations=3, seed=0):
    """Mean and top-K right singular vectors of the rows of ``wr`` (host,
    randomized SVD with power iterations, Halko, Martinsson and Tropp 2011)."""
    wr = np.asarray(wr, dtype=np.float64)
    mean = wr.mean(axis=0)
    X = wr - mean
    rng = np.random.default_rng(seed)
    q = min(K + oversample, min(X.shape))
    Q = np.linalg.qr(X @ rng.standard_normal((X.shape[1], q)))[0]
    for _ in range(iterations):
        Q = np.linalg.qr(X.T @ Q)[0]
        Q = np.linalg.qr(X @ Q)[0]
    _, sv, Vt = np.linalg.svd(Q.T @ X, full_matrices=False)
    return mean, Vt[:K].T.copy(), sv[:K]


def device_pca(store, ld, N, rows, K, oversample=32, iterations=3, seed=0):
    """Mean and top-K right singular vectors of the field rows store[rows, :N]
    (device FP32, row pitch ``ld``, a multiple of 128): the randomized SVD of
    ``whitened_pca`` with the products against the field matrix X on the
    device (FP32 MFMA GEMMs over X and its transpose) and the QR and SVD
    factorizations of the thin factors in FP64 on the host.  ``len(rows)``
    must be a multiple of 128 and K + oversample at most 128."""
    dev = store.device
    st = _lib.stream()
    p = _lib.ptr
    M = len(rows)
    q = K + oversample
    assert M % 128 == 0 and ld % 128 == 0 and q <= 128
    X = torch.zeros((M, ld), dtype=torch.float32, device=dev)
    idx = torch.from_numpy(np.asarray(rows, dtype=np.int32)).to(dev)
    _lib.call("altar_gather_rows_f32", p(store), ld, p(idx), M, N, p(X), ld, st)
    XT = torch.empty((ld, M), dtype=torch.float32, device=dev)
    _lib.call("altar_transpose_f32", p(X), ld, M, ld, p(XT), M, st)
    # column means: a block of ones times X (row 0 of the product)
    ones = torch.ones((128, M), dtype=torch.float32, device=dev)
    sums = torch.empty((128, ld), dtype=torch.float32, device=dev)
    _lib.call("altar_gemm_f32", p(ones), M, p(XT), M, 128, ld, M, None, p(sums), ld, st)
    mean = sums[0].double().cpu().numpy() / M
    del ones, sums

    def x_times(Bt):
        """(X - 1 mean^T) Bt^T for Bt (q x ld): M x q"""
        B = torch.zeros((128, ld), dtype=torch.float32, device=dev)
        B[:q] = torch.from_numpy(Bt.astype(np.float32)).to(dev)
        bias = torch.zeros(128, dtype=torch.float32, device=dev)
        bias[:q] = torch.from_numpy((Bt @ mean).astype(np.float32)).to(dev)
        out = torch.empty((M, 128), dtype=torch.float32, device=dev)
        _lib.call("altar_gemm_f32", p(X), ld, p(B), ld, M, 128, ld, p(bias), p(out), 128, st)
        return out[:, :q].double().cpu().numpy()

    def xt_times(Q):
        """(X - 1 mean^T)^T Q for Q (M x q): ld x q"""
        B = torch.zeros((128, M), dtype=torch.float32, device=dev)
        B[:q] = torch.from_numpy(np.ascontiguousarray(Q.T, dtype=np.float32)).to(dev)
        out = torch.empty((ld, 128), dtype=torch.float32, device=dev)
        _lib.call("altar_gemm_f32", p(XT), M, p(B), M, ld, 128, M, None, p(out), 128, st)
        return out[:, :q].double().cpu().numpy() - np.outer(mean, Q.sum(axis=0))

    rng = np.random.default_rng(seed)
    omega = rng.standard_normal((q, ld))
    omega[:, N:] = 0.0
    Q = np.linalg.qr(x_times(omega))[0]
    for _ in range(iterations):
        Qn = np.linalg.qr(xt_times(Q))[0]
        Q = np.linalg.qr(x_times(Qn.T))[0]
    _, sv, Vt = np.linalg.svd(xt_times(Q).T, full_matrices=False)
    return mean[:N].copy(), Vt[:K, :N].T.copy(), sv[:K]


class Surrogate:
    """PCA + MLP surrogate of a source's whitened residual field."""

    def __init__(self, model, K=64, hidden=256, layers=4, capacity=200_000, max_batch=16384, seed=0):
        self.model = model
        self.K = K
        self.P = model.parameters
        self.N = model.observations
        self.Npad = (self.N + 127) // 128 * 128
        self.capacity = (int(capacity) + 127) // 128 * 128
        self.max_batch = max_batch
        self.seed = seed
        dev = model.device
        self.theta = torch.empty((self.capacity, self.P), dtype=torch.float64, device=dev)
        self.wr = torch.empty((self.capacity, self.Npad), dtype=torch.float32, device=dev)
        self.count = 0
        self.mlp = MLP(self.P, hidden, layers, K, max_batch).init(seed)
        from .distributions import flatten
        kinds, a, b = flatten(model.psets, "prior")
        kinds, a, b = np.array(kinds), np.array(a), np.array(b)
        # network inputs scaled to [-1, 1] over a uniform support, or over
        # mean +- 3 sigma of a gaussian prior
        self.lo = np.where(kinds == 0, a, a - 3.0 * b)
        self.hi = np.where(kinds == 0, b, a + 3.0 * b)
        self.mlp.input_scaling(self.lo, self.hi)
        self.trained = False
        self.validation_rmse = None
        self.history = []

    # -- training data
    def space(self):
        return self.capacity - self.count

    def collect_llk(self, theta, llk_out, precision="fp32"):
        """Exact data llk of ``theta`` (C rows, C % 128 == 0) that also appends the
        whitened residuals to the training set.  Returns the number stored.
        The fields are FP32: a model that computes in FP64 gets them from an
        FP32 evaluation beside its own FP64 llk."""
        C = theta.shape[0]
        if self.space() < C:
            self.model.source.data_llk(theta, llk_out, precision)
            return 0
        theta = theta.contiguous()
        dst = self.wr[self.count:self.count + C]
        fields_llk = llk_out if precision != "fp64" else torch.empty_like(llk_out)
        _lib.call("altar_source_data_llk_fields", self.model.source.handle, _lib.ptr(theta), C,
                  _lib.ptr(fields_llk), _lib.PRECISIONS[self.fields_precision(precision)], _lib.ptr(dst), self.Npad,
                  _lib.stream())
        if precision == "fp64":
            self.model.source.data_llk(theta, llk_out, precision)
        self.theta[self.count:self.count + C].copy_(theta)
        self.count += C
        return C

    @staticmethod
    def fields_precision(precision):
        """The precision of a fields evaluation (FP32 fields for FP64 models)."""
        return "fp32" if precision == "fp64" else precision

    # -- fitting
    def fit(self, epochs=40, batch=4096, lr=2e-3, holdout=0.05, pca_sample=40_000, lr_final=1e-4):
        """PCA of the collected whitened residuals, then Adam on the MLP."""
        n = self.count
        assert n >= 1024, "too few training fields"
        dev = self.theta.device
        rng = np.random.default_rng(self.seed)
        perm = rng.permutation(n)
        nval = max(256, int(holdout * n))
        val_idx, tr_idx = perm[:nval], perm[nval:]
        # PCA of a sample of the training rows (whitened residuals), products on the device
        m = max(128, min(pca_sample, len(tr_idx)) // 128 * 128)
        samp = np.sort(rng.choice(tr_idx, size=m, replace=len(tr_idx) < m))
        mean, V, sv = device_pca(self.wr, self.Npad, self.N, samp, self.K, seed=self.seed)
        self.mean, self.V = mean, V
        self.e = V.T @ mean
        self.cperp = float(mean @ mean - self.e @ self.e)
        # coefficients of every stored field: c = V^T (wr - mean), FP32 MFMA GEMM
        Kp = 128
        Vt = np.zeros((Kp, self.Npad), dtype=np.float32)
        Vt[:self.K, :self.N] = V.T
        bias = np.zeros(Kp, dtype=np.float32)
        bias[:self.K] = self.e
        Vt_d = torch.from_numpy(Vt).to(dev)
        bias_d = torch.from_numpy(bias).to(dev)
        rows = (n + 127) // 128 * 128
        coef = torch.zeros((rows, Kp), dtype=torch.float32, device=dev)
        _lib.call("altar_gemm_f32", _lib.ptr(self.wr), self.Npad, _lib.ptr(Vt_d), self.Npad, rows, Kp, self.Npad,
                  _lib.ptr(bias_d), _lib.ptr(coef), Kp, _lib.stream())
        c_host = coef[:n, :self.K].cpu().numpy().astype(np.float64)
        s = c_host.std(axis=0) + 1e-12
        self.s = s
        target_all = torch.from_numpy((c_host / s).astype(np.float32)).to(dev)
        del coef
        # loss weights: the share of whitened variance of each component
        w = s ** 2 / np.sum(s ** 2)
        w_d = torch.from_numpy(w.astype(np.float32)).to(dev)
        self.e_d = torch.from_numpy(self.e).to(dev)
        self.s_d = torch.from_numpy(s).to(dev)
        # training: one native call per epoch over shuffled mini-batches
        # (gathered rows, backward pass, Adam with a cosine-decayed rate)
        tr = np.asarray(tr_idx)
        batch = min(batch, self.max_batch)
        nb = max(1, len(tr) // batch)
        # Each epoch's last-batch loss stays on the device until the end, so the
        # host draws the next epoch's order while the device runs this one.
        losses = torch.zeros(epochs, dtype=torch.float32, device=dev)
        total = epochs * nb
        for ep in range(epochs):
            order = torch.from_numpy(rng.permutation(tr).astype(np.int32)).to(dev)
            _lib.call("altar_mlp_train_epoch", self.mlp.handle, _lib.ptr(self.theta), _lib.ptr(target_all), self.K,
                      _lib.ptr(w_d), _lib.ptr(order), nb, batch, ep * nb, total, float(lr), float(lr_final), 0.9,
                      0.999, 1e-8, _lib.ptr(losses[ep:ep + 1]), _lib.stream())
        self.history.extend(float(v) for v in losses.cpu().numpy())
        self.trained = True
        self.val_idx = val_idx
        self.validation_rmse = self.validate_stored(val_idx)
        return self

    # -- inference
    def data_llk(self, theta, out, z=None):
        z = self.mlp.forward(theta, z)
        _lib.call("altar_surrogate_llk", _lib.ptr(z), z.shape[1], theta.shape[0], self.K, _lib.ptr(self.e_d),
                  _lib.ptr(self.s_d), self.cperp, float(self.model.normalization), _lib.ptr(out), _lib.stream())
        return out

    def predict_wr(self, theta):
        """Predicted whitened residuals (host, FP64) of the given samples."""
        z = self.mlp.forward(theta).double().cpu().numpy()
        return self.mean + (z * self.s) @ self.V.T

    def field_rmse(self, wr_pred, wr_true):
        """RMSE of the LOS field in metres: W^-1 maps whitened residuals back."""
        if getattr(self.model, "block_offsets", None) is not None:
            return self._blocked_rmse(wr_pred - wr_true)
        L = getattr(self, "_L", None)
        if L is None and self.model.whitening_kind == "lower":
            # W = chol(Cd)^-1 for the exact whitener, so W^-1 = chol(Cd)
            if getattr(self.model, "whiten", "exact") == "exact":
                L = np.linalg.cholesky(np.asarray(self.model.cd, dtype=np.float64))
            else:
                L = np.linalg.inv(self.model.W)
            self._L = L
        diff = wr_pred - wr_true
        if L is None:
            f = diff / self.model.W[None, :]
        else:
            f = diff @ L.T
        return float(np.sqrt(np.mean(f ** 2)))

    def _blocked_rmse(self, diff):
        """Block-diagonal whitener (one block per dataset): each dataset's
        columns map back through its own W_b^-1 = chol(Cd_b) (exact mode)."""
        from . import norms
        blocks = getattr(self, "_Lb", None)
        if blocks is None:
            blocks = []
            for off, ds in zip(self.model.block_offsets, self.model.datasets):
                cd = np.asarray(ds.cd, dtype=np.float64)
                if cd.ndim == 1:
                    blocks.append((off, len(ds.x), np.sqrt(cd)))
                else:
                    L = ds.chol if ds.chol is not None else np.linalg.cholesky(cd)
                    if getattr(self.model, "whiten", "exact") != "exact":
                        L = np.linalg.inv(norms.whitener(cd, self.model.whiten)[1])
                    blocks.append((off, len(ds.x), L))
            self._Lb = blocks
        sq, count = 0.0, 0
        for off, n, L in blocks:
            d = diff[:, off:off + n]
            f = d * L[None, :] if L.ndim == 1 else d @ L.T
            sq += float(np.sum(f ** 2))
            count += f.size
        return float(np.sqrt(sq / count))

    # -- field RMSE on the device (the host forms above are the reference)
    def _rmse_blocks(self):
        """(offset, n, map) per dataset on the device, where map takes a
        whitened-residual difference back to the LOS field: chol(Cd) (padded
        to a multiple of 128, FP32) for a full Cd in exact mode, W^-1 in the
        other modes, the standard deviations for a diagonal Cd."""
        blocks = getattr(self, "_Lb_dev", None)
        if blocks is not None:
            return blocks
        from . import norms
        dev = self.theta.device
        if getattr(self.model, "block_offsets", None) is not None:
            parts = []
            for off, ds in zip(self.model.block_offsets, self.model.datasets):
                cd = np.asarray(ds.cd, dtype=np.float64)
                if cd.ndim == 1:
                    parts.append((int(off), len(ds.x), np.sqrt(cd)))
                else:
                    L = ds.chol if ds.chol is not None else np.linalg.cholesky(cd)
                    if getattr(self.model, "whiten", "exact") != "exact":
                        L = np.linalg.inv(norms.whitener(cd, self.model.whiten)[1])
                    parts.append((int(off), len(ds.x), L))
        elif self.model.whitening_kind == "lower":
            if getattr(self.model, "whiten", "exact") == "exact":
                L = np.linalg.cholesky(np.asarray(self.model.cd, dtype=np.float64))
            else:
                L = np.linalg.inv(self.model.W)
            parts = [(0, self.N, L)]
        else:
            parts = [(0, self.N, 1.0 / np.asarray(self.model.W, dtype=np.float64))]
        blocks = []
        for off, n, L in parts:
            if L.ndim == 1:
                blocks.append((off, n, torch.from_numpy(L.astype(np.float32)).to(dev)))
            else:
                npad = (n + 127) // 128 * 128
                Lp = np.zeros((npad, npad), dtype=np.float32)
                Lp[:n, :n] = L
                blocks.append((off, n, torch.from_numpy(Lp).to(dev)))
        self._Lb_dev = blocks
        return blocks

    def _v_device(self):
        """The basis V (Npad x K, FP32, zero rows past N) and -mean on the device."""
        cached = getattr(self, "_Vdev", None)
        if cached is not None and cached[0] is self.V:
            return cached[1], cached[2]
        dev = self.theta.device
        Vp = np.zeros((self.Npad, self.K), dtype=np.float32)
        Vp[:self.N] = self.V
        negmean = np.zeros(self.Npad, dtype=np.float32)
        negmean[:self.N] = -self.mean
        self._Vdev = (self.V, torch.from_numpy(Vp).to(dev), torch.from_numpy(negmean).to(dev))
        return self._Vdev[1], self._Vdev[2]

    def field_rmse_device(self, theta, wr_true):
        """RMSE in metres of the predicted LOS field against the exact one:
        predicted whitened residuals mean + V (s * z) and the mapping back
        through W^-1 are FP32 GEMMs on the device.  theta: rows x P (FP64),
        wr_true: rows x (>= N) FP32 whitened residuals, both on the device;
        taken in chunks of the network's batch size."""
        sq, count = 0.0, 0
        step = self.mlp.max_batch
        for r0 in range(0, theta.shape[0], step):
            s_, c_ = self._field_sq(theta[r0:r0 + step], wr_true[r0:r0 + step])
            sq += s_
            count += c_
        return float(np.sqrt(sq / count))

    def _field_sq(self, theta, wr_true):
        """Sum of squared LOS-field errors and their count for one chunk."""
        rows = theta.shape[0]
        rp = (rows + 127) // 128 * 128
        dev = theta.device
        p, st = _lib.ptr, _lib.stream()
        V, negmean = self._v_device()
        z = self.mlp.forward(theta.contiguous())
        A = torch.zeros((rp, self.K), dtype=torch.float32, device=dev)
        A[:rows] = z[:, :self.K] * self.s_d.float()
        pred = torch.empty((rp, self.Npad), dtype=torch.float32, device=dev)
        _lib.call("altar_gemm_f32", p(A), self.K, p(V), self.K, rp, self.Npad, self.K, p(negmean), p(pred),
                  self.Npad, st)
        diff = pred[:rows, :self.N] - wr_true[:rows, :self.N]
        sq, count = 0.0, 0
        for off, n, L in self._rmse_blocks():
            d = diff[:, off:off + n]
            if L.dim() == 1:
                sq += float(((d * L) ** 2).sum())
            else:
                npad = L.shape[0]
                Ad = torch.zeros((rp, npad), dtype=torch.float32, device=dev)
                Ad[:rows, :n] = d
                f = torch.empty((rp, npad), dtype=torch.float32, device=dev)
                _lib.call("altar_gemm_f32", p(Ad), npad, p(L), npad, rp, npad, npad, None, p(f), npad, st)
                sq += float((f[:rows, :n].double() ** 2).sum())
            count += rows * n
        return sq, count

    def validate_stored(self, idx):
        idx = np.sort(np.asarray(idx))[:4096]
        dev = self.theta.device
        idx_d = torch.from_numpy(idx.astype(np.int32)).to(dev)
        th = torch.empty((len(idx), self.P), dtype=torch.float64, device=dev)
        _lib.call("altar_da_pack", _lib.ptr(self.theta), _lib.ptr(idx_d), len(idx), self.P, _lib.ptr(th),
                  _lib.stream())
        wr = torch.empty((len(idx), self.N), dtype=torch.float32, device=dev)
        _lib.call("altar_gather_rows_f32", _lib.ptr(self.wr), self.Npad, _lib.ptr(idx_d), len(idx), self.N,
                  _lib.ptr(wr), self.N, _lib.stream())
        return self.field_rmse_device(th, wr)
