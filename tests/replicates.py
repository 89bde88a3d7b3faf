"""Posterior moments judged with standard errors measured from replicate runs.

A posterior test runs R independent replicates (different seeds) of the same
sampler and compares the pooled mean and standard deviation of each parameter
with a closed form.  The Monte-Carlo standard error comes from the scatter of
the replicate means, not from an assumed effective sample size:

* for each parameter j, the between-replicate variance of the replicate
  means, divided by the posterior variance sd_j^2, estimates 1 / ESS_j of one
  replicate;
* for a Gaussian target and a proposal covariance proportional to the
  sample covariance (CATMIP's Metropolis), every standardized direction mixes
  alike, so 1 / ESS is pooled over the parameters: (R - 1) P degrees of
  freedom instead of R - 1;
* the pooled mean then has standard error sd_j sqrt(1 / (ESS R)), and the
  pooled standard deviation sd_j sqrt(v_sd / R) with v_sd the pooled
  standardized variance of the replicate standard deviations.
"""
import numpy as np


def replicate_moments(thetas, mu, sd):
    """thetas: R final samples (C x P each); mu, sd: the target's means and
    standard deviations.  Returns a dict with the z-scores of the pooled
    means and standard deviations, and the measured ESS of one replicate."""
    R = len(thetas)
    assert R >= 3, "at least three replicates"
    means = np.stack([t.mean(0) for t in thetas])
    sds = np.stack([t.std(0, ddof=1) for t in thetas])
    v_mean = float(np.mean(means.var(0, ddof=1) / sd ** 2))
    v_sd = float(np.mean(sds.var(0, ddof=1) / sd ** 2))
    z_mean = (means.mean(0) - mu) / (sd * np.sqrt(v_mean / R))
    z_sd = (sds.mean(0) - sd) / (sd * np.sqrt(v_sd / R))
    return dict(R=R, C=thetas[0].shape[0], ess=1.0 / v_mean, z_mean=z_mean, z_sd=z_sd,
                sd_ratio=sds.mean(0) / sd, mean_err_sd=(means.mean(0) - mu) / sd)


def summary(m):
    return (f"R={m['R']} C={m['C']} ESS per replicate {m['ess']:.0f} | max |z| mean "
            f"{np.max(np.abs(m['z_mean'])):.2f}, sd {np.max(np.abs(m['z_sd'])):.2f} | chi2 mean "
            f"{np.sum(m['z_mean'] ** 2):.1f}, sd {np.sum(m['z_sd'] ** 2):.1f} over {len(m['z_mean'])} | "
            f"max |mean err| {np.max(np.abs(m['mean_err_sd'])):.4f} sd, sd ratio "
            f"[{m['sd_ratio'].min():.4f}, {m['sd_ratio'].max():.4f}]")


def assert_moments(m, zmax=3.0):
    """Every pooled mean and standard deviation within zmax standard errors.
    With many parameters the largest of P standard normals grows (about 2.9
    at P = 288), so for P > 20 the bound is max(zmax, 4) together with a
    chi-square bound on the sum of squares (its 99.9% point)."""
    P = len(m["z_mean"])
    bound = zmax if P <= 20 else max(zmax, 4.0)
    assert np.all(np.abs(m["z_mean"]) <= bound), summary(m)
    assert np.all(np.abs(m["z_sd"]) <= bound), summary(m)
    if P > 20:
        from scipy import stats
        crit = stats.chi2.ppf(0.999, P)
        assert np.sum(m["z_mean"] ** 2) <= crit and np.sum(m["z_sd"] ** 2) <= crit, summary(m)
