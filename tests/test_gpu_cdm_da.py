"""Delayed acceptance on the CDM case at its full size: 5,000 LOS points on
the 100 x 50 lattice, full exponential Cd, 16,384 chains, and the
10 -> 256 -> 256 -> 256 -> 64 network (GELU, BF16 weights, FP32 accumulation)
predicting 64 PCA coefficients of the whitened field.

The annealer collects the training fields itself: 13 Metropolis steps per
stage while beta < 0.1 store the whitened residual field of every exact
evaluation, 196,608 fields (12 steps of 16,384 chains fill the 200,000-field
store's multiple of 128 chains).  Then:

* the fitted surrogate validates below the 2 mm field-RMSE bound on fresh
  exact fields, and the next stage runs delayed acceptance;
* with the bound set below the validated RMSE the stage falls back to plain
  Metropolis, which collects that stage's exact fields for a refit;
* one two-stage step at full size: the screen, the compaction and the exact
  stage's decisions equal a numpy restatement of the Christen-Fox rule on the
  oracle's Philox uniforms (Metropolis.py:189-223 extended by the screen);
* surrogate trajectories with the trained network: the fused one-launch
  kernel and the separate sub-step calls give the same states, likelihoods,
  end points and counters, bit for bit.
"""
import numpy as np
import pytest
import torch

from altar_amd import Job, _lib
from altar_amd.bayesian import Annealer, COV
from altar_amd.bayesian.da import DelayedAcceptance
from altar_amd.cases import cdm_case
from altar_amd.cdm import CDM
from altar_amd.surrogate import Surrogate

pytestmark = pytest.mark.gpu

C, N, STAGE_STEPS = 16384, 5000, 13


@pytest.fixture(scope="module")
def trained(cuda):
    print(f"setup: CDM {C} chains x {N} LOS, anneal with field collection while beta < 0.1, fit the surrogate",
          flush=True)
    case = cdm_case(N=N)
    job = Job(chains=C, steps=STAGE_STEPS, seed=780)
    model = CDM(case["psets"], case["x"], case["y"], case["los"], case["d"], case["oid"], case["cd"], nu=0.25,
                job=job).initialize()
    sur = Surrogate(model, K=64, capacity=200_000, max_batch=C, seed=3)
    da = DelayedAcceptance(sur, train_beta=0.1, max_rmse=2e-3, steps=STAGE_STEPS, validate_chains=1024,
                           fit_kwargs=dict(epochs=100, batch=4096, lr=2e-3))
    ann = Annealer(da, COV(), seed=job.seed).initialize(job)
    ann.model = model
    w = ann.worker
    w.start(ann)
    collected_betas = []
    while not sur.trained and w.beta + job.tolerance < 1:
        w.cool(ann)
        if da._collecting(w.step):
            collected_betas.append(w.beta)
        stats = w.walk(ann)
        w.resample(ann, stats)
    print("fields", sur.count, "collected at beta", [round(b, 4) for b in collected_betas],
          "validation rmse", sur.validation_rmse, "modes", [m["mode"] for m in da.mode_history], flush=True)
    return dict(model=model, sur=sur, da=da, ann=ann, collected=collected_betas)


def test_surrogate_trains_on_annealer_fields_and_validates(trained):
    sur, da, ann = trained["sur"], trained["da"], trained["ann"]
    assert sur.trained
    assert sur.count >= 190_000 and sur.count % C == 0
    assert trained["collected"] and max(trained["collected"]) < 0.1
    # the next stage validates the surrogate on fresh exact fields of its chains
    w = ann.worker
    w.cool(ann)
    step = w.step
    da.prepareSamplingPDF(ann, step)
    ok = da._validate(ann, step)
    print(f"beta {step.beta:.4f}: validation RMSE {sur.validation_rmse * 1e3:.3f} mm (bound 2 mm)")
    assert ok and sur.validation_rmse < 2e-3
    n_da = da.exact_evals
    da.steps = 2
    acc, rej, unl = da.walkChains(ann, step)
    assert da.mode_history[-1]["mode"] == "da"
    assert acc > 0 and da.exact_evals - n_da < 2 * C        # the screen spares exact evaluations


def test_fallback_when_the_bound_is_below_the_validated_rmse(trained):
    sur, da, ann = trained["sur"], trained["da"], trained["ann"]
    step = ann.worker.step
    da.prepareSamplingPDF(ann, step)
    keep = da.max_rmse
    da.max_rmse = 0.5 * sur.validation_rmse
    ev0, count0 = da.exact_evals, sur.count
    da.steps = 2
    try:
        acc, rej, unl = da.walkChains(ann, step)
    finally:
        da.max_rmse = keep
    print("forced fallback: mode", da.mode_history[-1], "stored fields", sur.count, "(was", count0, ")")
    assert da.mode_history[-1]["mode"] == "plain"
    assert acc + rej + unl == 2 * C and da.exact_evals - ev0 == 2 * C
    # the plain stage replaced the store by its own exact fields for a refit
    assert da._refit and sur.count == 2 * C


def test_two_stage_decisions_match_restatement(trained, oracle):
    sur, da, ann, model = trained["sur"], trained["da"], trained["ann"], trained["model"]
    step = ann.worker.step.clone()
    da.prepareSamplingPDF(ann, step)
    beta = float(step.beta)
    seed, stage = ann.seed, ann.worker.iteration
    theta0 = step.theta.cpu().numpy()
    prior0, post0 = step.prior.cpu().numpy(), step.posterior.cpu().numpy()
    llk_s = torch.empty(C, dtype=torch.float64, device=step.theta.device)
    sur.data_llk(step.theta, llk_s)
    post_s0 = prior0 + beta * llk_s.cpu().numpy()
    da.steps = 1
    acc, rej, unl = da._walk_da(ann, step)
    ws = da._dws
    cand = ws["cand"].cpu().numpy()
    mask = ws["mask"].cpu().numpy()
    cprior, cdata_s = ws["cprior"].cpu().numpy(), ws["cdata_s"].cpu().numpy()
    n = int(ws["count"].item())
    index = ws["index"][:n].cpu().numpy()
    cdata = ws["cdata"][:n].cpu().numpy()
    # restatement: the screen on its own stream, then the corrected ratio
    cpost_s = cprior + beta * cdata_s
    u1 = oracle.uniforms(seed, stage, 0, _lib.TAG_SCREEN, C)
    passed = (mask == 0) & (np.log(u1) <= cpost_s - post_s0)
    idx = np.nonzero(passed)[0]
    assert np.array_equal(index, idx)
    u2 = oracle.uniforms(seed, stage, 0, _lib.TAG_ACCEPT, C)
    cpost = cprior[idx] + beta * cdata
    ok = np.log(u2[idx]) <= (cpost - post0[idx]) - (cpost_s[idx] - post_s0[idx])
    accepted = idx[ok]
    th = step.theta.cpu().numpy()
    moved = np.zeros(C, bool)
    moved[accepted] = True
    print(f"two-stage step at full size: {mask.sum()} rejected by the model, {n} passed the screen, "
          f"{len(accepted)} accepted")
    assert acc == len(accepted) and rej == int((mask != 0).sum())
    assert np.array_equal(th[moved], cand[moved]) and np.array_equal(th[~moved], theta0[~moved])
    assert np.array_equal(step.data.cpu().numpy()[accepted], cdata[ok])
    # the exact stage's llk of a few screened-in candidates against the oracle
    from altar_amd import norms
    pick = idx[:: max(1, len(idx) // 6)][:6]
    res = oracle.cdm_residuals(cand[pick], (0, 2, 3, 4, 7, -1), model.x, model.y, model.los, model.d,
                               np.zeros(N, np.int32), 0.25)
    ref = oracle.l2_llk(res, oracle.exact_whitener(model.cd), norms.normalization(model.cd))
    got = cdata[np.searchsorted(idx, pick)]
    assert np.max(np.abs(got - ref) / np.abs(ref)) < 4e-6


def test_fused_trajectories_match_separate_calls_with_trained_network(trained):
    sur, ann, model = trained["sur"], trained["ann"], trained["model"]
    step = ann.worker.step
    runs = []
    for fused in (False, True):
        da = DelayedAcceptance(sur, steps=2, substeps=16, substep_scale=1.6, fused=fused)
        assert da.fused_supported(model, 10) == fused
        da.scaling, da.sigma_chol = trained["da"].scaling, None
        da.prepareSamplingPDF(ann, step)
        st = step.clone()
        stats = da._walk_trajectory(ann, st)
        torch.cuda.synchronize()
        runs.append((stats, st, list(da.substep_counts), {k: v.clone() for k, v in da._dws.items()}))
    (s0, st0, c0, w0), (s1, st1, c1, w1) = runs
    print("trained network, 2 exact steps x 16 sub-steps: exact-stage stats", s0, "sub-steps", c0)
    assert s0 == s1 and c0 == c1 and c0[0] > 0 and s0[0] > 0
    for name in ("theta", "prior", "data", "posterior"):
        assert torch.equal(getattr(st0, name), getattr(st1, name)), name
    for name in ("phi", "prior_phi", "post_phi", "post_s"):
        assert torch.equal(w0[name], w1[name]), name
    assert torch.equal(w0["index"][C:], w1["index"][C:])
