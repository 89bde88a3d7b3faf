"""CATMIP on the CPU with the oracle's FP64 pieces (test infrastructure only).

Follows altar/altar/bayesian/Annealer.py:78-135 (stage loop),
COV.py:65-251 (dbeta grid or Brent, weighted covariance, resampling) and
Metropolis.py:93-281 (propose, verify, accept, scaling), drawing from the
same Philox streams as the GPU path: resampling (stage, 0, tag 2), proposal
normals (stage, step, tag 0), acceptance dice (stage, step, tag 1), initial
sample (0, attempt, tag 3).  Resampled rows are in original-sample order.
"""
import numpy as np

from . import oracle as o


def anneal(data_llk, kinds, lo, hi, prep_lo, prep_hi, C, steps, seed, tolerance=1e-3,
           verify_model=None, scaling=0.1, max_stages=200, solver="grid"):
    """Returns a dict with the final state and the per-stage history."""
    kinds = np.asarray(kinds, np.int32)
    P = len(kinds)
    theta = o.prep_sample(seed, 0, kinds, prep_lo, prep_hi, C)
    attempt = 0
    while verify_model is not None:
        bad = verify_model(theta).astype(bool)
        if not bad.any():
            break
        attempt += 1
        fresh = o.prep_sample(seed, attempt, kinds, prep_lo, prep_hi, C)
        theta[bad] = fresh[bad]
    prior, _ = o.prior(theta, kinds, lo, hi)
    data = data_llk(theta)
    beta = 0.0
    post = prior + beta * data
    history = []
    stage = 0
    while beta + tolerance < 1 and stage < max_stages:
        stage += 1
        new_beta, cov, w = o.dbeta_grid(data, beta) if solver == "grid" else o.dbeta_brent(data, beta)
        mean, sigma = o.weighted_covariance(theta, w)
        count = o.multiplicities(w, o.uniforms(seed, stage, 0, 2, C))
        src = np.repeat(np.arange(C), count)
        theta, prior, data = theta[src].copy(), prior[src].copy(), data[src].copy()
        beta = new_beta
        post = prior + beta * data
        L = np.linalg.cholesky(sigma * scaling ** 2)
        acc = rej = unl = 0
        for k in range(steps):
            xi = o.normals(seed, stage, k, 0, C, P)
            cand = o.mh_propose(theta, L, xi)
            _, mask = o.prior(cand, kinds, lo, hi)
            if verify_model is not None:
                mask = mask + verify_model(cand)
            cand[mask != 0] = theta[mask != 0]
            cprior, _ = o.prior(cand, kinds, lo, hi)
            cdata = data_llk(cand)
            cpost = cprior + beta * cdata
            dice = o.uniforms(seed, stage, k, 1, C)
            a, dec = o.mh_accept(theta, prior, data, post, cand, cprior, cdata, cpost, mask, dice)
            acc += a
            rej += int((dec == 1).sum())
            unl += int((dec == 2).sum())
        scaling = o.mh_adjust_scaling(acc, rej, unl)
        history.append(dict(stage=stage, beta=beta, cov=cov, accepted=acc, rejected=rej, unlikely=unl,
                            scaling=scaling))
    return dict(theta=theta, prior=prior, data=data, posterior=post, beta=beta, history=history)
