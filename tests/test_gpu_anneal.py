"""End-to-end CATMIP on the GPU against the oracle's CPU driver
(oracle/anneal.py) on the same Philox streams.

* Mogi, FP64 path, diagonal Cd: the beta schedule agrees to 1e-9, the
  acceptance counts of every stage agree exactly, and the final sample agrees
  to 1e-8 (relative), row by row.
* Linear-Gaussian: the annealed posterior moments agree with the closed form
  Sigma = (G^T Cd^-1 G + I)^-1, mu = Sigma G^T Cd^-1 d within Monte-Carlo
  error (5 standard errors with an effective sample size of C / 10).
"""
import numpy as np
import pytest
import torch

from altar_amd import Job, norms
from altar_amd.bayesian import Annealer, COV, Metropolis
from altar_amd.cases import mogi_case
from altar_amd.distributions import flatten

pytestmark = pytest.mark.gpu


def test_mogi_anneal_matches_oracle(cuda, oracle):
    from oracle.anneal import anneal
    from altar_amd.mogi import Mogi, default_psets
    print("case: mogi anneal C=512 N=200 steps=3", flush=True)
    case = mogi_case(N=200)
    job = Job(chains=512, steps=3, seed=11, tolerance=1e-3)
    model = Mogi(default_psets(), case["x"], case["y"], case["los"], case["d"], case["oid"], case["cd"],
                 precision="fp64", job=job).initialize()
    # the host COV placement is the bit-for-bit restatement of libaltar
    annealer = Annealer(Metropolis(steps=3), COV(resident=False), seed=11).initialize(job)
    annealer.posterior(model)
    gpu = annealer.worker.step
    W = 1.0 / np.sqrt(case["cd"])
    norm = norms.normalization(case["cd"])

    def llk(th):
        r = oracle.mogi_residuals(th, (0, 2, 3, 4), case["x"], case["y"], case["los"], case["d"], case["oid"],
                                  0.25)
        return oracle.l2_llk(r, W, norm)

    kinds, lo, hi = flatten(model.psets, "prior")
    ref = anneal(llk, kinds, lo, hi, lo, hi, 512, 3, 11)
    gb = np.array([h["beta"] for h in annealer.history])
    rb = np.array([h["beta"] for h in ref["history"]])
    print("stages gpu", len(gb), "oracle", len(rb))
    assert len(gb) == len(rb)
    assert np.allclose(gb, rb, rtol=1e-9, atol=0)
    for h, r in zip(annealer.history, ref["history"]):
        assert (h["accepted"], h["rejected"], h["unlikely"]) == (r["accepted"], r["rejected"], r["unlikely"])
    th = gpu.theta.cpu().numpy()
    assert np.allclose(th, ref["theta"], rtol=1e-8, atol=1e-10)
    assert np.allclose(gpu.data.cpu().numpy(), ref["data"], rtol=1e-9)


def test_linear_gaussian_posterior(cuda):
    from altar_amd.linear import Linear, default_psets
    rng = np.random.default_rng(12)
    N, P, C = 60, 3, 8192
    G = rng.normal(size=(N, P))
    truth = rng.normal(size=P)
    xy = rng.uniform(0, 2e4, (N, 2))
    cd = norms.exponential_covariance(xy, 0.5, 5e3) + 0.5 * np.eye(N)
    d = G @ truth + np.linalg.cholesky(cd) @ rng.normal(size=N)
    job = Job(chains=C, steps=20, seed=3, tolerance=1e-3)
    model = Linear(default_psets(P), G, d, cd, job=job).initialize()
    annealer = Annealer(Metropolis(steps=20), COV(), seed=3).initialize(job)
    annealer.posterior(model)
    th = annealer.worker.step.theta.cpu().numpy()
    ci = np.linalg.inv(cd)
    S = np.linalg.inv(G.T @ ci @ G + np.eye(P))
    mu = S @ G.T @ ci @ d
    sd = np.sqrt(np.diag(S))
    se = sd / np.sqrt(C / 10)
    print("posterior mean", th.mean(0), "closed form", mu, "se", se)
    assert np.all(np.abs(th.mean(0) - mu) < 5 * se)
    emp = np.cov(th.T)
    # covariance: relative error of the standard deviations within 10%
    assert np.all(np.abs(np.sqrt(np.diag(emp)) / sd - 1) < 0.1)
