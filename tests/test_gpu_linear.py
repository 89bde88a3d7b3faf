"""Linear model on the GPU against the oracle, on the reference's own
example cases (tests/golden/patch9: G 108 x 18; tests/golden/patch144:
G 351 x 288; both with a diagonal Cd) and on a synthetic full-Cd case.

* data llk, pre-whitened (one GEMM with fused bias + square) and direct
  (forward GEMM, then whitening GEMM), FP32 and split BF16 + refinement,
  against the FP64 oracle (bars in LLK_BARS);
* posterior of the annealed patch-9 case with the reference's Gaussian
  prior (sigma 0.5, models/linear/examples/linear.pfg:17-25) against the
  closed form, over 6 replicate anneals: means and standard deviations within
  3 standard errors measured from the replicates (tests/replicates.py);
* patch-144 (P = 288: the tiled proposal kernel, the wave-per-chain prior):
  the Metropolis walk at beta = 1 keeps the closed-form posterior invariant.
  16,384 chains start at independent exact draws of N(mu, Sigma) and take
  300 steps through the GPU path (proposal, verify, prior, FP32 linear llk,
  accept); the end points are independent draws of the target, so each mean
  has the standard error sd / sqrt(C).  A wrong prior, likelihood or accept
  rule moves the population off the target and fails.  CATMIP's own anneal of
  this case at the reference's 20 steps per stage ends under-dispersed (see
  the printed figures): 92 stages of resampling, which 20 random-walk steps in
  288 dimensions cannot rejuvenate, and a scaling rule (Metropolis.py:262-281)
  that oscillates between acceptance 0.5 and 0.003.
"""
import os

import numpy as np
import pytest
import torch

from altar_amd import Job, io, norms
from altar_amd.bayesian import Annealer, COV, Metropolis
from altar_amd.distributions import Contiguous, Gaussian
from altar_amd.linear import Linear
from replicates import assert_moments, replicate_moments, summary

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def psets(P, sigma):
    return {"slip": Contiguous(P, Gaussian(0.0, sigma))}


def llk_ref(oracle, theta, G, d, cd):
    res = oracle.linear_residuals(theta, G, d)
    return oracle.l2_llk(res, oracle.exact_whitener(cd), oracle.normalization(cd))


# relative llk error bars per precision, 3-5x the largest measured over the
# three cases and both modes (fp32 3.1e-7, bf16r 1.4e-5)
LLK_BARS = {"fp32": 1e-6, "bf16r": 5e-5}


def load_case(case, rng):
    if case in ("patch9", "patch144"):
        c = io.load_linear_case(os.path.join(GOLDEN, case))
        return c["G"], c["d"], c["cd"]
    N, P = 700, 40
    G = rng.normal(size=(N, P)) * 1e-2
    cd = norms.exponential_covariance(rng.uniform(0, 5e4, (N, 2)), 25e-6, 5e3)
    d = G @ rng.uniform(0, 2, P) + np.linalg.cholesky(cd) @ rng.normal(size=N)
    return G, d, cd


@pytest.mark.parametrize("case", ["patch9", "patch144", "full"])
def test_linear_llk_vs_oracle(cuda, oracle, case):
    rng = np.random.default_rng(40)
    G, d, cd = load_case(case, rng)
    N, P = G.shape
    C = 1000
    print(f"case: linear llk {case} N={N} P={P} C={C}", flush=True)
    theta = rng.normal(0, 0.5, (C, P))
    if case == "full":
        theta = theta + 1.0
    ref = llk_ref(oracle, theta, G, d, cd)
    t = torch.from_numpy(theta).to(cuda)
    for mode in ("prewhitened", "direct"):
        for prec in ("fp32", "bf16r"):
            m = Linear(psets(P, 0.5), G, d, cd, mode=mode, precision=prec, job=Job(chains=C)).initialize()
            out = torch.empty(C, dtype=torch.float64, device=cuda)
            m.source.data_llk(t, out, prec)
            rel = np.abs(out.cpu().numpy() - ref) / np.abs(ref)
            print(case, mode, prec, "llk relative error: median", np.median(rel), "max", rel.max())
            assert rel.max() < LLK_BARS[prec], (case, mode, prec)


def test_patch9_posterior(cuda):
    c = io.load_linear_case(os.path.join(GOLDEN, "patch9"))
    G, d, cd = c["G"], c["d"], c["cd"]
    N, P = G.shape
    C, steps, sigma = 8192, 20, 0.5
    print(f"case: linear posterior patch9 N={N} P={P} C={C} steps={steps}, 6 replicates", flush=True)
    ci = np.linalg.inv(cd)
    S = np.linalg.inv(G.T @ ci @ G + np.eye(P) / sigma ** 2)
    mu = S @ G.T @ ci @ d
    sd = np.sqrt(np.diag(S))
    thetas, stages = [], []
    for seed in range(9, 15):
        job = Job(chains=C, steps=steps, seed=seed, tolerance=1e-3)
        model = Linear(psets(P, sigma), G, d, cd, job=job).initialize()
        ann = Annealer(Metropolis(steps=steps), COV(), seed=seed).initialize(job)
        ann.posterior(model)
        thetas.append(ann.worker.step.theta.cpu().numpy())
        stages.append(len(ann.history))
    m = replicate_moments(thetas, mu, sd)
    print("patch9 stages", stages, summary(m))
    assert_moments(m)


def test_patch144_posterior_invariance(cuda):
    from altar_amd.bayesian.cooling_step import CoolingStep
    c = io.load_linear_case(os.path.join(GOLDEN, "patch144"))
    G, d, cd = c["G"], c["d"], c["cd"]
    N, P = G.shape
    C, steps, sigma = 16384, 300, 0.5
    print(f"case: linear patch144 N={N} P={P} C={C}: {steps} Metropolis steps from exact posterior draws",
          flush=True)
    ci = np.linalg.inv(cd)
    S = np.linalg.inv(G.T @ ci @ G + np.eye(P) / sigma ** 2)
    mu = S @ G.T @ ci @ d
    sd = np.sqrt(np.diag(S))
    rng = np.random.default_rng(144)
    start = mu + rng.standard_normal((C, P)) @ np.linalg.cholesky(S).T
    job = Job(chains=C, steps=steps, seed=144, tolerance=1e-3)
    model = Linear(psets(P, sigma), G, d, cd, job=job).initialize()
    ann = Annealer(Metropolis(steps=steps), COV(), seed=144).initialize(job)
    ann.model = model
    ann.worker.iteration = 1
    z = lambda: torch.zeros(C, dtype=torch.float64, device=cuda)
    step = CoolingStep(1.0, torch.from_numpy(start).to(cuda), (z(), z(), z()),
                       sigma=torch.from_numpy(S).to(cuda))
    model.likelihoods(ann, step)
    sampler = ann.sampler
    sampler.scaling = 2.38 / np.sqrt(P)        # the random-walk scale for a Gaussian target in P dimensions
    sampler.prepareSamplingPDF(ann, step)
    acc, rej, unl = sampler.walkChains(ann, step)
    th = step.theta.cpu().numpy()
    moved = np.sqrt(np.mean(((th - start) / sd) ** 2))
    z_mean = (th.mean(0) - mu) / (sd / np.sqrt(C))
    ratio = th.std(0, ddof=1) / sd
    # the sd of a sample standard deviation of C independent normal draws is sd / sqrt(2 (C - 1))
    z_sd = (ratio - 1.0) * np.sqrt(2 * (C - 1))
    print(f"acceptance {acc / (C * steps):.3f}, rms displacement {moved:.2f} sd | max |z| mean "
          f"{np.abs(z_mean).max():.2f}, sd {np.abs(z_sd).max():.2f} | chi2 mean {np.sum(z_mean ** 2):.1f}, "
          f"sd {np.sum(z_sd ** 2):.1f} over {P} | sd ratio [{ratio.min():.4f}, {ratio.max():.4f}]")
    assert acc + rej + unl == C * steps and 0.1 < acc / (C * steps) < 0.5
    assert moved > 0.8          # the chains have travelled about a posterior sd
    from scipy import stats
    crit = stats.chi2.ppf(0.999, P)
    assert np.abs(z_mean).max() <= 4.0 and np.sum(z_mean ** 2) <= crit
    assert np.abs(z_sd).max() <= 4.0 and np.sum(z_sd ** 2) <= crit

    # the annealed CATMIP result at the reference's settings, for the record
    job = Job(chains=C, steps=20, seed=9, tolerance=1e-3)
    model = Linear(psets(P, sigma), G, d, cd, job=job).initialize()
    ann = Annealer(Metropolis(steps=20), COV(), seed=9).initialize(job)
    ann.posterior(model)
    th = ann.worker.step.theta.cpu().numpy()
    r = th.std(0) / sd
    print(f"annealed (20 steps per stage): {len(ann.history)} stages, sd ratio [{r.min():.3f}, {r.max():.3f}], "
          f"max |mean - mu| {np.abs((th.mean(0) - mu) / sd).max():.2f} sd")
