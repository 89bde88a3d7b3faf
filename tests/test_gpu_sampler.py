"""GPU parity of the Metropolis, prior and COV kernels against the oracle.

* Philox normals: within 1e-12 (Box-Muller through device vs host libm).
* proposals: within 1e-12 relative of the oracle's theta + (L xi)^T.
* accept/reject: decisions and updated state bit-exact given the same
  uniforms and the same candidate likelihoods.
* prior: verify masks bit-exact; log densities within 1e-14 relative.
* weighted covariance: within 1e-12 relative; resampling gather exact.
"""
import ctypes

import numpy as np
import pytest
import torch

from altar_amd import _lib

pytestmark = pytest.mark.gpu


def test_normals_match_oracle(cuda, oracle):
    C, P = 1000, 7
    out = torch.empty((C, P), dtype=torch.float64, device=cuda)
    _lib.call("altar_random_normals", _lib.ptr(out), C, P, 1234, 3, 5, _lib.TAG_PROPOSAL, None)
    ref = oracle.normals(1234, 3, 5, _lib.TAG_PROPOSAL, C, P)
    assert np.max(np.abs(out.cpu().numpy() - ref)) < 1e-12


@pytest.mark.parametrize("C,P", [(2048, 10), (1000, 100), (2048, 400)])
def test_propose_restore(cuda, oracle, C, P):
    """Proposals against the oracle: bit for bit given the same normals for
    P <= 32; the tiled kernel for P > 32 uses fused multiply-adds (same order,
    1e-14 of the row scale)."""
    print(f"case: propose C={C} P={P}", flush=True)
    rng = np.random.default_rng(7)
    theta = rng.normal(size=(C, P))
    A = rng.normal(size=(P, P))
    L = np.linalg.cholesky(A @ A.T + P * np.eye(P))
    t = torch.from_numpy(theta).to(cuda)
    Ld = torch.from_numpy(L).to(cuda)
    xi = torch.empty_like(t)
    cand = torch.empty_like(t)
    _lib.call("altar_mh_propose", _lib.ptr(t), C, P, _lib.ptr(Ld), 99, 2, 4, _lib.ptr(xi), _lib.ptr(cand), None)
    ref = oracle.mh_propose(theta, L, oracle.normals(99, 2, 4, 0, C, P))
    assert np.max(np.abs(cand.cpu().numpy() - ref) / (1 + np.abs(ref))) < 1e-12
    same = oracle.mh_propose(theta, L, xi.cpu().numpy())
    if P <= 32:
        assert np.array_equal(cand.cpu().numpy(), same)
    else:
        scale = np.abs(xi.cpu().numpy()) @ np.abs(L).T + np.abs(theta)
        assert np.max(np.abs(cand.cpu().numpy() - same) / scale) < 1e-14
    mask = torch.from_numpy(rng.integers(0, 2, C).astype(np.int32)).to(cuda)
    _lib.call("altar_mh_restore", _lib.ptr(cand), _lib.ptr(t), _lib.ptr(mask), C, P, None)
    m = mask.cpu().numpy().astype(bool)
    c = cand.cpu().numpy()
    assert np.array_equal(c[m], theta[m])


@pytest.mark.parametrize("C,P", [(4096, 6), (1001, 150)])
def test_accept_bit_exact(cuda, oracle, C, P):
    print(f"case: accept C={C} P={P}", flush=True)
    rng = np.random.default_rng(8)
    theta = rng.normal(size=(C, P))
    prior, data = rng.normal(size=C), rng.normal(size=C) * 10
    beta = 0.37
    post = prior + beta * data
    ctheta = rng.normal(size=(C, P))
    cprior, cdata = rng.normal(size=C), rng.normal(size=C) * 10
    cpost = cprior + beta * cdata
    mask = (rng.random(C) < 0.2).astype(np.int32)
    dice = oracle.uniforms(5, 1, 2, 1, C)
    # make some exact ties log(u) == diff impossible to mis-order: keep as drawn
    ref_state = [a.copy() for a in (theta, prior, data, post)]
    acc, dec = oracle.mh_accept(*ref_state, ctheta, cprior, cdata, cpost, mask, dice)
    dev = [torch.from_numpy(a.copy()).to(cuda) for a in (theta, prior, data, post)]
    cands = [torch.from_numpy(a).to(cuda) for a in (ctheta, cprior, cdata)]
    md = torch.from_numpy(mask).to(cuda)
    decision = torch.empty(C, dtype=torch.int32, device=cuda)
    counters = torch.zeros(3, dtype=torch.int64, device=cuda)
    for use_dice in (True, False):
        for a, b in zip(dev, (theta, prior, data, post)):
            a.copy_(torch.from_numpy(b))
        counters.zero_()
        dd = torch.from_numpy(dice).to(cuda) if use_dice else None
        _lib.call("altar_mh_accept", *(_lib.ptr(a) for a in dev), *(_lib.ptr(a) for a in cands), None,
                  _lib.ptr(md), beta, C, P, 5, 1, 2, _lib.ptr(dd), _lib.ptr(decision), _lib.ptr(counters), None)
        got = decision.cpu().numpy()
        assert np.array_equal(got, dec)
        cnt = counters.cpu().numpy()
        assert cnt.sum() == C and cnt[0] == acc
        for a, b in zip(dev, ref_state):
            assert np.array_equal(a.cpu().numpy(), b)


@pytest.mark.parametrize("rep", [1, 40])
def test_prior_verify_llk_sample(cuda, oracle, rep):
    """P = 5 (one thread per chain) and P = 200 (one wavefront per chain)."""
    rng = np.random.default_rng(9)
    C, P = 3000, 5 * rep
    print(f"case: prior C={C} P={P}", flush=True)
    kind = np.tile(np.array([0, 0, 1, 0, 1], dtype=np.int32), rep)
    a = np.tile(np.array([-1.0, 0.0, 0.5, 10.0, -2.0]), rep)
    b = np.tile(np.array([1.0, 2.0, 2.0, 20.0, 0.1]), rep)
    theta = rng.normal(size=(C, P)) * np.tile([1.2, 1.5, 2, 6, 1], rep) + np.tile([0, 1, 0, 15, 0], rep)
    theta[0, 0] = 1.0  # the closed upper bound is inside (Uniform.py:66)
    ref_llk, ref_mask = oracle.prior(theta, kind, a, b)
    from altar_amd.model import Prior
    pr = Prior(kind, a, b)
    t = torch.from_numpy(theta).to(cuda)
    llk = torch.empty(C, dtype=torch.float64, device=cuda)
    mask = torch.zeros(C, dtype=torch.int32, device=cuda)
    _lib.call("altar_prior_llk", pr.handle, _lib.ptr(t), C, _lib.ptr(llk), None)
    _lib.call("altar_prior_verify", pr.handle, _lib.ptr(t), C, _lib.ptr(mask), None)
    assert np.array_equal(mask.cpu().numpy(), ref_mask)
    assert np.max(np.abs(llk.cpu().numpy() - ref_llk) / np.abs(ref_llk)) < 1e-14
    # initial draws stay inside the uniform supports
    _lib.call("altar_prior_sample", pr.handle, _lib.ptr(t), C, 11, 0, 0, None, None)
    s = t.cpu().numpy()
    for j in (0, 1, 3, P - 2):
        assert s[:, j].min() >= a[j] and s[:, j].max() < b[j]
    assert abs(s[:, 2].mean() - a[2]) < 5 * b[2] / np.sqrt(C)


@pytest.mark.parametrize("C,P", [(5000, 8), (3001, 37)])
def test_covariance_and_gather(cuda, oracle, C, P):
    print(f"case: weighted covariance C={C} P={P}", flush=True)
    rng = np.random.default_rng(10)
    theta = rng.normal(size=(C, P)) @ rng.normal(size=(P, P)) + rng.normal(size=P)
    w = rng.random(C)
    w /= w.sum()
    t = torch.from_numpy(theta).to(cuda)
    wd = torch.from_numpy(w).to(cuda)
    mean = torch.empty(P, dtype=torch.float64, device=cuda)
    sig = torch.empty((P, P), dtype=torch.float64, device=cuda)
    _lib.call("altar_cov_covariance", _lib.ptr(t), _lib.ptr(wd), C, P, _lib.ptr(mean), _lib.ptr(sig), None)
    rm, rs = oracle.weighted_covariance(theta, w)
    assert np.allclose(mean.cpu().numpy(), rm, rtol=1e-12, atol=1e-12)
    assert np.allclose(sig.cpu().numpy(), rs, rtol=1e-10, atol=1e-12)
    # same summation order as the oracle: bit for bit on the lower triangle
    assert np.array_equal(mean.cpu().numpy(), rm)
    assert np.array_equal(np.tril(sig.cpu().numpy()), np.tril(rs))
    # closed form: numpy's weighted covariance with bias=True
    assert np.allclose(rs, np.cov(theta.T, aweights=w, bias=True), rtol=1e-9, atol=1e-12)
    src = rng.integers(0, C, C).astype(np.int32)
    vecs = [torch.from_numpy(rng.normal(size=C)).to(cuda) for _ in range(3)]
    outs = [torch.empty_like(t)] + [torch.empty_like(v) for v in vecs]
    _lib.call("altar_resample_gather", _lib.ptr(t), *(_lib.ptr(v) for v in vecs),
              _lib.ptr(torch.from_numpy(src).to(cuda)), C, P, *(_lib.ptr(o) for o in outs), None)
    assert np.array_equal(outs[0].cpu().numpy(), theta[src])
    for v, o in zip(vecs, outs[1:]):
        assert np.array_equal(o.cpu().numpy(), v.cpu().numpy()[src])


@pytest.mark.parametrize("C", [16384, 4096, 1001])
def test_resident_cov_matches_oracle(cuda, oracle, C):
    """GPU-resident COV stage against the oracle on identical llk: median
    exact, beta and COV to 1e-12, weights to 1e-12, multiplicities and the
    resampling map identical on the same uniforms."""
    rng = np.random.default_rng(C)
    for beta0, scale in ((0.0, 30.0), (0.2, 3.0), (0.7, 0.3)):
        llk = rng.normal(-800, scale, C)
        d = torch.from_numpy(llk).to(cuda)
        w = torch.empty(C, dtype=torch.float64, device=cuda)
        out = torch.zeros(3, dtype=torch.float64, device=cuda)
        _lib.call("altar_cov_update_dev", _lib.ptr(d), C, beta0, 0.01, 1.0, _lib.ptr(w), _lib.ptr(out), None)
        beta, cov, med = out.cpu().numpy()
        rb, rc, rw = oracle.dbeta_grid(llk, beta0)
        assert med == np.median(llk)
        assert abs(beta - rb) <= 1e-12 * max(1.0, rb) and abs(cov - rc) <= 1e-12 * rc
        wg = w.cpu().numpy()
        assert np.allclose(wg, rw, rtol=1e-12, atol=1e-300)
        edges = torch.empty(C + 1, dtype=torch.float64, device=cuda)
        count = torch.empty(C, dtype=torch.int32, device=cuda)
        offset = torch.empty(C, dtype=torch.int32, device=cuda)
        src = torch.empty(C, dtype=torch.int32, device=cuda)
        _lib.call("altar_cov_resample_dev", _lib.ptr(w), C, 77, 5, _lib.ptr(edges), _lib.ptr(count), _lib.ptr(offset),
                  _lib.ptr(src), None)
        want = oracle.multiplicities(wg, oracle.uniforms(77, 5, 0, _lib.TAG_RESAMPLE, C))
        assert np.array_equal(count.cpu().numpy(), want)
        assert np.array_equal(src.cpu().numpy(), np.repeat(np.arange(C), want))
