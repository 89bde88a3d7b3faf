"""End-to-end CATMIP on the GPU against the oracle's CPU driver
(oracle/anneal.py) on the same Philox streams.

* Mogi, FP64 path, diagonal Cd, at 512 chains x 200 points x 3 steps and at
  the benchmark's own size, 1,024 chains x 1,000 points x 5 steps per stage:
  the beta schedule agrees to 1e-9, the acceptance counts of every stage
  agree exactly, and the final sample agrees to 1e-8 (relative), row by row.
  Both COV placements run: on the host (the bit-for-bit restatement of
  libaltar's dbeta_grid) and resident on the GPU (COV(resident=True), the
  default: median, grid search, weights, resampling and the weighted
  covariance in device kernels); and the Brent solver (libaltar's
  dbeta_brent) against the oracle's Brent.
* Linear-Gaussian: the annealed posterior moments of 6 replicate anneals
  (independent seeds) agree with the closed form Sigma = (G^T Cd^-1 G + I)^-1,
  mu = Sigma G^T Cd^-1 d within 3 standard errors measured from the scatter
  of the replicates (tests/replicates.py).
"""
import numpy as np
import pytest

from altar_amd import Job, norms
from altar_amd.bayesian import Annealer, COV, Metropolis
from altar_amd.cases import mogi_case
from altar_amd.distributions import flatten
from replicates import assert_moments, replicate_moments, summary

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("C,N,steps,resident,solver", [(512, 200, 3, False, "grid"), (1024, 1000, 5, False, "grid"),
                                                       (1024, 1000, 5, True, "grid"),
                                                       (1024, 1000, 5, True, "brent")])
def test_mogi_anneal_matches_oracle(cuda, oracle, C, N, steps, resident, solver):
    from oracle.anneal import anneal
    from altar_amd.bayesian import Brent, Grid
    from altar_amd.mogi import Mogi, default_psets
    print(f"case: mogi anneal C={C} N={N} steps={steps} resident COV={resident} solver={solver}", flush=True)
    case = mogi_case(N=N)
    job = Job(chains=C, steps=steps, seed=11, tolerance=1e-3)
    model = Mogi(default_psets(), case["x"], case["y"], case["los"], case["d"], case["oid"], case["cd"],
                 precision="fp64", job=job).initialize()
    annealer = Annealer(Metropolis(steps=steps), COV(resident=resident, solver=Brent() if solver == "brent" else Grid()),
                        seed=11).initialize(job)
    annealer.posterior(model)
    gpu = annealer.worker.step
    W = 1.0 / np.sqrt(case["cd"])
    norm = norms.normalization(case["cd"])

    def llk(th):
        r = oracle.mogi_residuals(th, (0, 2, 3, 4), case["x"], case["y"], case["los"], case["d"], case["oid"],
                                  0.25)
        return oracle.l2_llk(r, W, norm)

    kinds, lo, hi = flatten(model.psets, "prior")
    ref = anneal(llk, kinds, lo, hi, lo, hi, C, steps, 11, solver=solver)
    gb = np.array([h["beta"] for h in annealer.history])
    rb = np.array([h["beta"] for h in ref["history"]])
    print("stages gpu", len(gb), "oracle", len(rb), "max beta difference",
          np.max(np.abs(gb - rb)) if len(gb) == len(rb) else None)
    assert len(gb) == len(rb)
    assert np.allclose(gb, rb, rtol=1e-9, atol=0)
    for h, r in zip(annealer.history, ref["history"]):
        assert (h["accepted"], h["rejected"], h["unlikely"]) == (r["accepted"], r["rejected"], r["unlikely"])
    th = gpu.theta.cpu().numpy()
    print("max relative theta difference", np.max(np.abs(th - ref["theta"]) / np.maximum(np.abs(ref["theta"]), 1e-10)))
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
    ci = np.linalg.inv(cd)
    S = np.linalg.inv(G.T @ ci @ G + np.eye(P))
    mu = S @ G.T @ ci @ d
    sd = np.sqrt(np.diag(S))
    thetas = []
    for seed in range(3, 9):
        job = Job(chains=C, steps=20, seed=seed, tolerance=1e-3)
        model = Linear(default_psets(P), G, d, cd, job=job).initialize()
        annealer = Annealer(Metropolis(steps=20), COV(), seed=seed).initialize(job)
        annealer.posterior(model)
        thetas.append(annealer.worker.step.theta.cpu().numpy())
    m = replicate_moments(thetas, mu, sd)
    print("linear-Gaussian:", summary(m))
    assert_moments(m)
