"""Error of every precision against the FP64 oracle at the benchmark sizes.

* CDM, 5,000 LOS, full exponential Cd: the FP32 forward (FP64 sides below
  |sin beta| = 0.4) and the data llk with FP32, split-BF16 + refinement and
  plain BF16 whitening, over 64 prior draws and 32 sources near the truth.
* Linear, G 4,000 x 400, full Cd (pre-whitened GEMM): FP32, BF16 +
  refinement and plain BF16.

Bounds, 3-5 times the largest error measured on an MI355X: CDM
displacements 5e-5 of the field max (measured 1.3e-5); llk relative 4e-6
(FP32, measured 1.0e-6), 1.2e-3 (BF16 + refinement, 3.5e-4); linear llk
relative 3e-7 (FP32, 8.2e-8), 4e-6 (BF16 + refinement, 1.1e-6).  Plain BF16
is reported, not bounded.  The measured figures
are printed as one JSON line starting with PRECISION.
"""
import json

import numpy as np
import pytest
import torch

from altar_amd import Job, cases
from altar_amd.cdm import CDM
from altar_amd.linear import Linear

pytestmark = pytest.mark.gpu
LAYOUT = (0, 2, 3, 4, 7, -1)


def stats(x):
    x = np.asarray(x)
    return {"median": float(np.median(x)), "max": float(np.max(x))}


def cdm_thetas(rng, oracle, truth):
    rows = []
    while len(rows) < 64:
        p = [rng.uniform(-10e3, 10e3), rng.uniform(-10e3, 10e3), rng.uniform(1e3, 6e3),
             rng.uniform(1e-3, 1.0), rng.uniform(100, 3000), rng.uniform(100, 3000), rng.uniform(100, 3000),
             rng.uniform(-90, 90), rng.uniform(-90, 90), rng.uniform(-90, 90)]
        if not oracle.cdm_vertices(p[0], p[1], p[2], p[4], p[5], p[6], p[7], p[8], p[9])[1]:
            rows.append(p)
    scale = np.array([100, 100, 100, 0.01, 50, 50, 50, 2, 2, 2])
    while len(rows) < 96:
        p = truth + rng.normal(size=10) * scale
        if not oracle.cdm_vertices(p[0], p[1], p[2], p[4], p[5], p[6], p[7], p[8], p[9])[1]:
            rows.append(list(p))
    return np.array(rows)


def test_precision_table(cuda, oracle):
    print("case: precision table CDM 96 x 5000, linear 128 x 4000 x 400", flush=True)
    rng = np.random.default_rng(77)
    out = {}
    # ---- CDM
    fwd = lambda th, x, y, los: oracle.cdm_residuals(th, LAYOUT, x, y, los, np.zeros(len(x)),
                                                      np.zeros(len(x), np.int32), 0.25)
    case = cases.cdm_case(N=5000, forward=fwd)
    theta = cdm_thetas(rng, oracle, case["truth"])
    C = len(theta)
    pred = oracle.cdm_residuals(theta, LAYOUT, case["x"], case["y"], case["los"], np.zeros(5000),
                                case["oid"], 0.25)
    res = pred - case["d"][None, :]
    W = oracle.exact_whitener(case["cd"])
    norm = oracle.normalization(case["cd"])
    ref = oracle.l2_llk(res, W, norm)
    model = CDM(case["psets"], case["x"], case["y"], case["los"], case["d"], case["oid"], case["cd"],
                job=Job(chains=C)).initialize()
    t = torch.from_numpy(theta).to(cuda)
    for thr in (0.4, 0.3, 0.2, 0.1):
        model.source.fp64_sides(thr)
        got = model.forwardModelBatched(t, "fp32").cpu().numpy()
        fe = np.abs(got - pred).max(axis=1) / np.abs(pred).max(axis=1)
        out[f"cdm_forward_fp32_normwise_fp64sides{thr}"] = stats(fe)
        llk = torch.empty(C, dtype=torch.float64, device=cuda)
        model.source.data_llk(t, llk, "fp32")
        d = np.abs(llk.cpu().numpy() - ref)
        out[f"cdm_llk_fp32_fp64sides{thr}"] = {"abs": stats(d), "abs_near_truth": stats(d[64:])}
        assert fe.max() < 5e-5
    model.source.fp64_sides(model.fp64_sides)
    for prec, tol in (("fp32", 4e-6), ("bf16r", 1.2e-3), ("bf16", None)):
        llk = torch.empty(C, dtype=torch.float64, device=cuda)
        model.source.data_llk(t, llk, prec)
        d = np.abs(llk.cpu().numpy() - ref)
        out[f"cdm_llk_{prec}"] = {"abs": stats(d), "rel": stats(d / np.abs(ref)),
                                  "abs_near_truth": stats(d[64:])}
        if tol is not None:
            assert (d / np.abs(ref)).max() < tol, prec
    # ---- linear
    lc = cases.linear_case()
    Cl = 128
    # half far from the data (0.3 m per parameter), half inside the posterior
    # bulk (5 mm: the posterior sd is about 8 mm)
    th = lc["truth"] + rng.normal(0, 1, (Cl, 400)) * np.repeat([0.3, 0.005], Cl // 2)[:, None]
    lref = oracle.l2_llk(oracle.linear_residuals(th, lc["G"], lc["d"]), oracle.exact_whitener(lc["cd"]),
                         oracle.normalization(lc["cd"]))
    tl = torch.from_numpy(th).to(cuda)
    for prec, tol in (("fp32", 3e-7), ("bf16r", 4e-6), ("bf16", None)):
        m = Linear(lc["psets"], lc["G"], lc["d"], lc["cd"], precision=prec, job=Job(chains=Cl)).initialize()
        llk = torch.empty(Cl, dtype=torch.float64, device=cuda)
        m.source.data_llk(tl, llk, prec)
        d = np.abs(llk.cpu().numpy() - lref)
        out[f"linear_llk_{prec}"] = {"abs": stats(d), "rel": stats(d / np.abs(lref)),
                                     "abs_near_truth": stats(d[Cl // 2:])}
        if tol is not None:
            assert (d / np.abs(lref)).max() < tol, prec
    print("PRECISION " + json.dumps(out), flush=True)
