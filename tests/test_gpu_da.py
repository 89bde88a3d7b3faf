"""Delayed acceptance with a learned or deliberately wrong surrogate.

1. Exactness (stationarity): 8,192 chains start at exact draws of a 2-D
   Gaussian target; every step proposes, screens with a biased surrogate
   (shifted mean, inflated covariance) and corrects with the exact target
   through the DA kernels.  After 200 steps the sample must still have the
   target's mean (within 5 Monte-Carlo standard errors) and covariance
   (within 5%): the two-stage rule keeps the target invariant whatever the
   surrogate.
2. End to end: a Mogi inversion with a full exponential Cd is annealed with
   plain Metropolis (two seeds) and with delayed acceptance whose surrogate
   is trained on the annealer's own beta < train_beta evaluations.  The DA
   posterior means must agree with the plain ones within 5 standard errors
   and the standard deviations within the plain-vs-plain spread (+10%), and
   DA must need fewer exact evaluations per accepted sample (the gain the
   network brings).
"""
import numpy as np
import pytest
import torch

from altar_amd import Job, _lib, norms
from altar_amd.bayesian import Annealer, COV, Metropolis
from altar_amd.bayesian.da import DelayedAcceptance
from altar_amd.cases import mogi_case
from altar_amd.mogi import Mogi, default_psets
from altar_amd.surrogate import Surrogate

pytestmark = pytest.mark.gpu


def gauss_logpdf(x, mu, S):
    Si = np.linalg.inv(S)
    d = x - mu
    return -0.5 * np.einsum("ci,ij,cj->c", d, Si, d)


def test_da_kernels_keep_target_invariant(cuda):
    rng = np.random.default_rng(31)
    C, P, steps = 8192, 2, 200
    mu = np.array([1.0, -2.0])
    S = np.array([[2.0, 0.8], [0.8, 1.0]])
    mu_s, S_s = mu + np.array([0.7, -0.5]), 1.5 * S        # biased surrogate
    theta = rng.multivariate_normal(mu, S, C)
    f = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(cuda)
    th = f(theta)
    prior = f(np.zeros(C))
    data = f(gauss_logpdf(theta, mu, S))
    post = f(data.cpu().numpy())
    post_s = f(gauss_logpdf(theta, mu_s, S_s))
    L = f(0.9 * np.linalg.cholesky(S))
    xi, cand = torch.empty_like(th), torch.empty_like(th)
    mask = torch.zeros(C, dtype=torch.int32, device=cuda)
    cprior = f(np.zeros(C))
    cpost_s = torch.empty(C, dtype=torch.float64, device=cuda)
    index = torch.empty(2 * C, dtype=torch.int32, device=cuda)
    count = torch.zeros(1, dtype=torch.int32, device=cuda)
    counters = torch.zeros(4, dtype=torch.int64, device=cuda)
    for k in range(steps):
        _lib.call("altar_mh_propose", _lib.ptr(th), C, P, _lib.ptr(L), 17, 1, k, _lib.ptr(xi), _lib.ptr(cand), None)
        c_host = cand.cpu().numpy()
        cdata_s = f(gauss_logpdf(c_host, mu_s, S_s))
        _lib.call("altar_da_screen", _lib.ptr(cprior), _lib.ptr(cdata_s), _lib.ptr(post_s), _lib.ptr(mask), 1.0, C,
                  17, 1, k, _lib.ptr(cpost_s), _lib.ptr(index), _lib.ptr(count), _lib.ptr(counters), None)
        n = int(count.item())
        idx = index[:n].cpu().numpy()
        cdata = f(gauss_logpdf(c_host[idx], mu, S))
        _lib.call("altar_da_accept", _lib.ptr(th), _lib.ptr(prior), _lib.ptr(data), _lib.ptr(post), _lib.ptr(post_s),
                  _lib.ptr(cand), _lib.ptr(cprior), _lib.ptr(cdata), _lib.ptr(cpost_s), _lib.ptr(index), n, 1.0, P,
                  17, 1, k, None, _lib.ptr(counters), None)
    x = th.cpu().numpy()
    acc, rej, unl2, scr = counters.cpu().numpy()
    print("accepted", acc, "screened", scr, "stage-2 rejected", unl2, "mean", x.mean(0), "cov", np.cov(x.T))
    assert acc > 0.2 * C * steps and scr > 0 and unl2 > 0
    se = np.sqrt(np.diag(S) / C)
    assert np.all(np.abs(x.mean(0) - mu) < 5 * se)
    assert np.allclose(np.cov(x.T), S, rtol=0.05, atol=0.05 * S.max())


def build(case, cd, C, steps, seed):
    job = Job(chains=C, steps=steps, seed=seed, tolerance=1e-3)
    return Mogi(default_psets(), case["x"], case["y"], case["los"], case["d"], case["oid"], cd,
                precision="fp32", job=job).initialize(), job


def test_da_matches_plain_posterior(cuda):
    print("case: mogi DA vs plain C=4096 N=600 steps=20", flush=True)
    rng = np.random.default_rng(21)
    case = mogi_case(N=600)
    cd = norms.exponential_covariance(np.stack([case["x"], case["y"]], 1), 25e-6, 5e3)
    case["d"] = case["d"] + np.linalg.cholesky(cd) @ rng.standard_normal(600) * 0.5
    C, S = 4096, 20
    runs = []
    for seed in (5, 6):
        model, job = build(case, cd, C, S, seed)
        plain = Annealer(Metropolis(steps=S), COV(), seed=seed).initialize(job)
        plain.posterior(model)
        runs.append((plain.worker.step.theta.cpu().numpy(), sum(h["accepted"] for h in plain.history),
                     C * S * len(plain.history)))
    (th_plain, plain_acc, plain_evals), (th_plain2, _, _) = runs

    model2, job2 = build(case, cd, C, S, 5)
    sur = Surrogate(model2, K=32, capacity=C * S * 2, max_batch=C, seed=1)
    da = DelayedAcceptance(sur, train_beta=0.02, max_rmse=2e-3, steps=S,
                           fit_kwargs=dict(epochs=30, batch=2048, lr=2e-3))
    ann = Annealer(da, COV(), seed=5).initialize(job2)
    ann.posterior(model2)
    th_da = ann.worker.step.theta.cpu().numpy()
    modes = [m["mode"] for m in da.mode_history]
    da_acc = sum(h["accepted"] for h in ann.history)
    print("surrogate rmse (m)", sur.validation_rmse, "modes", modes)
    print("plain: accepted", plain_acc, "exact evals", plain_evals,
          "| DA: accepted", da_acc, "exact evals", da.exact_evals, "screened", da.screened)
    assert sur.trained and "da" in modes
    sd = th_plain.std(0)
    se = sd / np.sqrt(C / 10)
    diff = np.abs(th_da.mean(0) - th_plain.mean(0))
    spread = np.abs(th_plain2.std(0) / sd - 1)
    print("mean diff / se", diff / (se * np.sqrt(2)), "sd ratio", th_da.std(0) / sd, "plain-plain sd ratio",
          th_plain2.std(0) / sd)
    assert np.all(diff < 5 * np.sqrt(2) * se)
    assert np.all(np.abs(th_da.std(0) / sd - 1) < 2 * spread + 0.1)
    assert da.exact_evals / max(da_acc, 1) < plain_evals / max(plain_acc, 1)
