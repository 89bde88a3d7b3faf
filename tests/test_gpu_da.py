"""Delayed acceptance with a learned or deliberately wrong surrogate.

1. Exactness (stationarity): 8,192 chains start at exact draws of a 2-D
   Gaussian target; every step proposes, screens with a biased surrogate
   (shifted mean, inflated covariance) and corrects with the exact target
   through the DA kernels.  After 200 steps the sample must still have the
   target's mean (within 3 standard errors sd / sqrt(C): the chains start
   independent and exact, so their end points are C independent draws of an
   invariant target) and covariance (within 5%, 3 standard errors of a
   variance at C = 8,192): the two-stage rule keeps the target invariant
   whatever the surrogate.
   The same holds for surrogate trajectories (several surrogate Metropolis
   sub-steps per exact step, judged at the end point).
2. Linear-Gaussian, annealed end to end with a surrogate trained on the
   annealer's own beta < train_beta evaluations: the DA posterior agrees with
   the closed form Sigma = (G^T Cd^-1 G + I)^-1, mu = Sigma G^T Cd^-1 d over
   6 replicate anneals (means and sds within 3 standard errors measured from
   the replicates, tests/replicates.py), and DA needs fewer exact
   evaluations per accepted sample than plain Metropolis on the same problem
   (the gain the network brings).
"""
import numpy as np
import pytest
import torch

from altar_amd import Job, _lib, norms
from altar_amd.bayesian import Annealer, COV, Metropolis
from altar_amd.bayesian.da import DelayedAcceptance
from altar_amd.linear import Linear, default_psets
from altar_amd.surrogate import Surrogate
from replicates import assert_moments, replicate_moments, summary

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
    assert np.all(np.abs(x.mean(0) - mu) < 3 * se)
    assert np.allclose(np.cov(x.T), S, rtol=0.05, atol=0.05 * S.max())


def test_trajectory_kernels_keep_target_invariant(cuda):
    print("case: surrogate trajectories, 2-D Gaussian, C=8192, 60 steps x 4 sub-steps", flush=True)
    rng = np.random.default_rng(32)
    C, P, steps, subs = 8192, 2, 60, 4
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
    phi, prior_phi, post_phi = torch.empty_like(th), torch.empty_like(prior), torch.empty_like(post)
    index = torch.empty(2 * C, dtype=torch.int32, device=cuda)
    count = torch.zeros(1, dtype=torch.int32, device=cuda)
    counters = torch.zeros(4, dtype=torch.int64, device=cuda)
    sub_counters = torch.zeros(3, dtype=torch.int64, device=cuda)
    p = _lib.ptr
    for k in range(steps):
        phi.copy_(th)
        prior_phi.copy_(prior)
        post_phi.copy_(post_s)
        for j in range(subs):
            _lib.call("altar_da_trajectory_propose", p(phi), C, P, p(L), 17, 1, k, j, p(xi), p(cand), None)
            cdata_s = f(gauss_logpdf(cand.cpu().numpy(), mu_s, S_s))
            _lib.call("altar_da_trajectory_step", p(phi), p(prior_phi), p(post_phi), p(cand), p(cprior), p(cdata_s),
                      p(cpost_s), p(mask), 1.0, C, P, 17, 1, k, j, p(index), p(sub_counters), None)
        _lib.call("altar_da_trajectory_list", p(index), C, p(count), None)
        n = int(count.item())
        idx = index[:n].cpu().numpy()
        cdata = f(gauss_logpdf(phi.cpu().numpy()[idx], mu, S))
        _lib.call("altar_da_accept", p(th), p(prior), p(data), p(post), p(post_s), p(phi), p(prior_phi), p(cdata),
                  p(post_phi), p(index), n, 1.0, P, 17, 1, k, None, p(counters), None)
    x = th.cpu().numpy()
    acc, _, unl2, _ = counters.cpu().numpy()
    sacc, srej, sunl = sub_counters.cpu().numpy()
    print("accepted", acc, "exact rejections", unl2, "sub-steps accepted", sacc, "of", sacc + srej + sunl,
          "mean", x.mean(0), "cov", np.cov(x.T))
    assert acc > 0.2 * C * steps and unl2 > 0 and srej == 0 and sacc + sunl == C * steps * subs
    se = np.sqrt(np.diag(S) / C)
    assert np.all(np.abs(x.mean(0) - mu) < 3 * se)
    assert np.allclose(np.cov(x.T), S, rtol=0.05, atol=0.05 * S.max())


def linear_problem(seed=12, N=64, P=3):
    rng = np.random.default_rng(seed)
    G = rng.normal(size=(N, P))
    truth = rng.normal(size=P)
    xy = rng.uniform(0, 2e4, (N, 2))
    cd = norms.exponential_covariance(xy, 0.5, 5e3) + 0.5 * np.eye(N)
    d = G @ truth + np.linalg.cholesky(cd) @ rng.normal(size=N)
    ci = np.linalg.inv(cd)
    S = np.linalg.inv(G.T @ ci @ G + np.eye(P))
    return G, d, cd, S, S @ G.T @ ci @ d


def test_da_linear_gaussian_posterior(cuda):
    print("case: linear-Gaussian DA C=8192 N=64 P=3 steps=20, 6 replicates", flush=True)
    G, d, cd, S, mu = linear_problem()
    P, C, steps = 3, 8192, 20
    # plain Metropolis on the same problem for the efficiency comparison
    job = Job(chains=C, steps=steps, seed=3, tolerance=1e-3)
    model = Linear(default_psets(P), G, d, cd, mode="direct", job=job).initialize()
    plain = Annealer(Metropolis(steps=steps), COV(), seed=3).initialize(job)
    plain.posterior(model)
    plain_acc = sum(h["accepted"] for h in plain.history)
    plain_evals = C * steps * len(plain.history)

    thetas = []
    for seed in range(3, 9):
        job2 = Job(chains=C, steps=steps, seed=seed, tolerance=1e-3)
        model2 = Linear(default_psets(P), G, d, cd, mode="direct", job=job2).initialize()
        sur = Surrogate(model2, K=32, capacity=C * steps * 2, max_batch=C, seed=1)
        da = DelayedAcceptance(sur, train_beta=0.05, max_rmse=1e9, steps=steps,
                               fit_kwargs=dict(epochs=30, batch=4096, lr=2e-3))
        ann = Annealer(da, COV(), seed=seed).initialize(job2)
        ann.posterior(model2)
        thetas.append(ann.worker.step.theta.cpu().numpy())
        modes = [m["mode"] for m in da.mode_history]
        assert sur.trained and "da" in modes
        if seed == 3:
            da_acc = sum(h["accepted"] for h in ann.history)
            print("surrogate field rmse", sur.validation_rmse, "modes", modes)
            print("plain: accepted", plain_acc, "exact evals", plain_evals,
                  "| DA: accepted", da_acc, "exact evals", da.exact_evals, "screened", da.screened)
            assert da.exact_evals / max(da_acc, 1) < plain_evals / max(plain_acc, 1)
    m = replicate_moments(thetas, mu, np.sqrt(np.diag(S)))
    print("DA posterior:", summary(m))
    assert_moments(m)


@pytest.mark.parametrize("fused", [True, False])
def test_trajectory_linear_gaussian_posterior(cuda, fused):
    print(f"case: linear-Gaussian surrogate trajectories C=8192 N=64 P=3 steps=10 sub-steps=6 fused={fused}, "
          "6 replicates", flush=True)
    G, d, cd, S, mu = linear_problem()
    P, C, steps = 3, 8192, 10
    thetas = []
    for seed in range(4, 10):
        job = Job(chains=C, steps=steps, seed=seed, tolerance=1e-3)
        model = Linear(default_psets(P), G, d, cd, mode="direct", job=job).initialize()
        sur = Surrogate(model, K=32, capacity=C * steps * 2, max_batch=C, seed=1)
        da = DelayedAcceptance(sur, train_beta=0.05, max_rmse=1e9, steps=steps, substeps=6, fused=fused,
                               fit_kwargs=dict(epochs=30, batch=4096, lr=2e-3))
        assert da.fused_supported(model, P) == fused
        ann = Annealer(da, COV(), seed=seed).initialize(job)
        ann.posterior(model)
        thetas.append(ann.worker.step.theta.cpu().numpy())
        modes = [m["mode"] for m in da.mode_history]
        assert sur.trained and "da" in modes and da.substep_counts[0] > 0
        if seed == 4:
            first = (ann, da, model)
            print("surrogate field rmse", sur.validation_rmse, "modes", modes, "exact evals", da.exact_evals,
                  "sub-steps (accepted, rejected, unlikely)", da.substep_counts)
    m = replicate_moments(thetas, mu, np.sqrt(np.diag(S)))
    print("trajectory DA posterior:", summary(m))
    assert_moments(m)
    # what the network gains: squared displacement (in posterior standard
    # deviations) per exact evaluation, trajectories against plain Metropolis
    # over 3 steps from the final state of the first replicate
    ann, da, model = first
    ann.model = model
    final = ann.worker.step
    sdev = final.theta.std(0)

    def jump(walk):
        st = final.clone()
        before = st.theta.clone()
        walk(st)
        return float((((st.theta - before) / sdev) ** 2).sum(1).mean())
    plain = Metropolis(steps=3)
    plain.scaling, plain.sigma_chol = da.scaling, da.sigma_chol
    ann.sampler = plain
    j_plain = jump(lambda st: plain.walkChains(ann, st)) / (3 * C)
    ann.sampler = da
    da.steps = 3
    ev0 = da.exact_evals
    j_traj = jump(lambda st: da._walk_trajectory(ann, st))
    j_traj /= max(da.exact_evals - ev0, 1)
    print("squared jump per exact evaluation: plain", j_plain, "trajectories", j_traj)
    assert j_traj > 2.0 * j_plain


@pytest.mark.parametrize("layers,C", [(4, 1000), (3, 1000), (5, 1000), (4, 37)])
def test_fused_trajectory_matches_separate_calls(cuda, layers, C):
    """One launch per trajectory (altar_da_trajectory_fused) against the
    separate calls, CDM source (free-surface check), two exact steps of 7
    sub-steps with wide sub-step proposals: states, likelihoods, end points
    and counters agree bit for bit.  Four layers take the kernel's
    register-resident path, three and five its general layer loop; 1,000
    chains end in a partial 64-chain tile, 37 fill less than one."""
    print(f"case: fused vs separate trajectories, CDM {C} chains x 512 LOS, 2 steps x 7 sub-steps, "
          f"{layers} layers", flush=True)
    from altar_amd.cases import cdm_case
    from altar_amd.cdm import CDM
    case = cdm_case(N=512)
    P = 10
    job = Job(chains=C, steps=1, seed=5)
    model = CDM(case["psets"], case["x"], case["y"], case["los"], case["d"], case["oid"], case["cd"], nu=0.25,
                job=job).initialize()
    ann = Annealer(Metropolis(), COV(), seed=5).initialize(job)
    ann.model = model
    w = ann.worker
    w.start(ann)
    w.cool(ann)
    step = w.step
    ann.sampler.prepareSamplingPDF(ann, step)
    sur = Surrogate(model, K=64, layers=layers, capacity=1024, max_batch=C, seed=2)
    dev = step.theta.device
    rng = np.random.default_rng(3)
    sur.e_d = torch.from_numpy(rng.normal(0.0, 1.0, 64)).to(dev)
    sur.s_d = torch.from_numpy(rng.uniform(0.05, 0.2, 64)).to(dev)
    sur.cperp = 3.0
    sur.trained = True
    runs = []
    for fused in (False, True):
        da = DelayedAcceptance(sur, steps=2, substeps=7, substep_scale=3.0, fused=fused)
        assert da.fused_supported(model, P) == fused
        da.sigma_chol = ann.sampler.sigma_chol
        st = step.clone()
        stats = da._walk_trajectory(ann, st)
        torch.cuda.synchronize()
        runs.append((stats, st, list(da.substep_counts), {k: v.clone() for k, v in da._dws.items()}))
    (s0, st0, c0, w0), (s1, st1, c1, w1) = runs
    print("exact-stage stats", s0, s1, "sub-steps (accepted, rejected, unlikely)", c0, c1)
    assert s0 == s1 and c0 == c1
    assert c0[0] > 0 and s0[0] > 0
    # every outcome occurs at 1,000 chains (at 37 the rare ones may not)
    assert (c0[1] > 0 and c0[2] > 0) if C >= 1000 else c0[1] + c0[2] > 0
    for name in ("theta", "prior", "data", "posterior"):
        assert torch.equal(getattr(st0, name), getattr(st1, name)), name
    for name in ("phi", "prior_phi", "post_phi", "post_s"):
        assert torch.equal(w0[name], w1[name]), name
    assert torch.equal(w0["index"][C:], w1["index"][C:])


def test_fused_trajectory_matches_separate_calls_linear_p32(cuda):
    """The fused kernel at its largest parameter count (P = 32, a full
    network-input tile) with Gaussian priors (the prior terms the CDM case's
    uniform priors do not reach) on a linear model: states, likelihoods,
    end points and counters equal the separate calls bit for bit."""
    print("case: fused vs separate trajectories, linear N=256 P=32 C=700, 2 steps x 7 sub-steps", flush=True)
    rng = np.random.default_rng(21)
    N, P, C = 256, 32, 700
    G = rng.normal(size=(N, P))
    xy = rng.uniform(0, 2e4, (N, 2))
    cd = norms.exponential_covariance(xy, 0.5, 5e3) + 0.5 * np.eye(N)
    d = G @ rng.normal(size=P) + np.linalg.cholesky(cd) @ rng.normal(size=N)
    job = Job(chains=C, steps=1, seed=9)
    model = Linear(default_psets(P), G, d, cd, mode="direct", job=job).initialize()
    ann = Annealer(Metropolis(), COV(), seed=9).initialize(job)
    ann.model = model
    w = ann.worker
    w.start(ann)
    w.cool(ann)
    step = w.step
    ann.sampler.prepareSamplingPDF(ann, step)
    sur = Surrogate(model, K=64, capacity=1024, max_batch=C, seed=4)
    dev = step.theta.device
    sur.e_d = torch.from_numpy(rng.normal(0.0, 1.0, 64)).to(dev)
    sur.s_d = torch.from_numpy(rng.uniform(0.05, 0.2, 64)).to(dev)
    sur.cperp = 2.0
    sur.trained = True
    runs = []
    for fused in (False, True):
        da = DelayedAcceptance(sur, steps=2, substeps=7, substep_scale=1.0, fused=fused)
        assert da.fused_supported(model, P) == fused
        da.sigma_chol = ann.sampler.sigma_chol
        st = step.clone()
        stats = da._walk_trajectory(ann, st)
        torch.cuda.synchronize()
        runs.append((stats, st, list(da.substep_counts), {k: v.clone() for k, v in da._dws.items()}))
    (s0, st0, c0, w0), (s1, st1, c1, w1) = runs
    print("exact-stage stats", s0, s1, "sub-steps (accepted, rejected, unlikely)", c0, c1)
    assert s0 == s1 and c0 == c1
    assert c0[0] > 0 and c0[2] > 0 and s0[0] > 0
    for name in ("theta", "prior", "data", "posterior"):
        assert torch.equal(getattr(st0, name), getattr(st1, name)), name
    for name in ("phi", "prior_phi", "post_phi", "post_s"):
        assert torch.equal(w0[name], w1[name]), name
    assert torch.equal(w0["index"][C:], w1["index"][C:])


def test_substep_scale_tuning(cuda):
    """Pilot trajectories tune the sub-step scale toward the target sub-step
    acceptance and leave the chains where they were."""
    print("case: sub-step scale tuning, linear-Gaussian C=4096 P=3", flush=True)
    G, d, cd, S, mu = linear_problem()
    P, C = 3, 4096
    job = Job(chains=C, steps=5, seed=6, tolerance=1e-3)
    model = Linear(default_psets(P), G, d, cd, mode="direct", job=job).initialize()
    sur = Surrogate(model, K=32, capacity=C * 10, max_batch=C, seed=1)
    da = DelayedAcceptance(sur, train_beta=0.05, max_rmse=1e9, steps=5, substeps=8,
                           fit_kwargs=dict(epochs=20, batch=2048, lr=2e-3))
    ann = Annealer(da, COV(), seed=6).initialize(job)
    ann.model = model
    w = ann.worker
    w.start(ann)
    while not sur.trained:
        w.cool(ann)
        w.resample(ann, w.walk(ann))
    w.cool(ann)
    step = w.step
    da.prepareSamplingPDF(ann, step)
    before = step.theta.clone()
    rates = {}
    for target in (0.2, 0.5):
        da.substep_scale, da.substep_target = 4.0, target
        scale = da.tune_substep_scale(ann, step, da._workspace(C, P, step.theta.device), rounds=8)
        # the rate at the tuned scale, from the counters of one more pilot
        cnt0 = list(da.substep_counts)
        da.substep_target = None
        da.steps = 1
        st = step.clone()
        da._walk_trajectory(ann, st)
        c = [b - a for a, b in zip(cnt0, da.substep_counts)]
        rates[target] = (scale, c[0] / max(sum(c), 1))
    print("target -> (scale, sub-step acceptance):", rates)
    assert torch.equal(step.theta, before)
    assert rates[0.2][0] > rates[0.5][0]
    for target, (_, rate) in rates.items():
        assert abs(rate - target) < 0.1


def test_da_on_fp64_mogi_with_diagonal_cd(cuda):
    """Delayed acceptance on the FP64 Mogi model with its diagonal Cd (the
    1,000-point case): the fields are collected from FP32 evaluations beside
    the chains' FP64 llk, the surrogate trains and validates, and the anneal
    reaches beta = 1 with delayed-acceptance stages."""
    from altar_amd.cases import mogi_case
    from altar_amd.mogi import Mogi, default_psets as mogi_psets
    print("case: DA on the FP64 Mogi model, diagonal Cd, C=1024 N=1000", flush=True)
    case = mogi_case()
    C, steps = 1024, 5
    job = Job(chains=C, steps=steps, seed=17, tolerance=1e-3)
    model = Mogi(mogi_psets(), case["x"], case["y"], case["los"], case["d"], case["oid"], case["cd"],
                 precision="fp64", job=job).initialize()
    sur = Surrogate(model, K=32, capacity=C * steps * 4, max_batch=C, seed=1)
    da = DelayedAcceptance(sur, train_beta=0.05, max_rmse=1e9, steps=steps,
                           fit_kwargs=dict(epochs=30, batch=1024, lr=2e-3))
    ann = Annealer(da, COV(), seed=17).initialize(job)
    ann.posterior(model)
    modes = [m["mode"] for m in da.mode_history]
    print("modes", modes, "validation rmse", sur.validation_rmse, "stages", len(ann.history))
    assert sur.trained and "da" in modes and ann.worker.beta >= 1.0 - job.tolerance
    for h in ann.history:
        assert h["accepted"] + h["rejected"] + h["unlikely"] == C * steps
