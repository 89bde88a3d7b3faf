"""GPU parity of the forward models and the data log-likelihood against the
FP64 oracle (oracle/oracle.c), on identical theta and observations.

Each bar is 3-5 times the largest error measured on an MI355X (stated per
test as "measured"):
* Mogi FP64 path: residuals to 5e-15 of the field's max, llk to 2e-13 relative.
* Mogi FP32 path: residuals to 3e-6 of the field's max; llk to 3e-5 relative.
* CDM (FP32, FP64 near-vertical sides): displacement error <= 1e-4 of the
  field's max per sample (normwise), golden vectors <= 1.2e-4; llk relative
  error <= 6e-6 (FP32 whitening), 6e-5 (BF16 + refinement).
* CDM with every side in FP64: the reference's own bar, 1e-6 relative per
  component at its MATLAB points (models/cdm/tests/libcdm.py:83-86).
* Whitening, full Cd (Mogi): FP64 2e-13, FP32 MFMA 3e-5, BF16 + refinement
  5e-4 relative; plain BF16 is reported, not bounded.
"""
import numpy as np
import pytest
import torch

from altar_amd import _lib, norms
from altar_amd.cases import mogi_case, unit_los
from altar_amd.source import Source

pytestmark = pytest.mark.gpu


def mogi_theta(rng, C, nsets=1):
    t = np.column_stack([rng.uniform(-10e3, 10e3, C), rng.uniform(-10e3, 10e3, C),
                         rng.uniform(500, 8000, C), rng.uniform(5, 7, C)] +
                        [rng.uniform(-0.1, 0.1, C) for _ in range(nsets)])
    return t


def cdm_theta(rng, C, oracle, offsets=0):
    rows = []
    while len(rows) < C:
        p = [rng.uniform(-10e3, 10e3), rng.uniform(-10e3, 10e3), rng.uniform(1e3, 6e3),
             rng.uniform(1e-3, 1.0), rng.uniform(100, 3000), rng.uniform(100, 3000), rng.uniform(100, 3000),
             rng.uniform(-90, 90), rng.uniform(-90, 90), rng.uniform(-90, 90)]
        p += [rng.uniform(-0.05, 0.05) for _ in range(offsets)]
        V, bad = oracle.cdm_vertices(p[0], p[1], p[2], p[4], p[5], p[6], p[7], p[8], p[9])
        if not bad:
            rows.append(p)
    return np.array(rows)


def make_mogi(case, P):
    N = len(case["x"])
    src = Source(_lib.MOGI, P, N, 0.25)
    src.layout([0, 2, 3, 4, 1])
    src.observations(case["x"], case["y"], case["los"], case["d"], case["oid"])
    return src


def test_mogi_residuals_fp64_fp32(cuda, oracle):
    print("case: mogi residuals C=256 N=1000", flush=True)
    case = mogi_case()
    rng = np.random.default_rng(1)
    theta = mogi_theta(rng, 256)
    ref = oracle.mogi_residuals(theta, (0, 2, 3, 4), case["x"], case["y"], case["los"], case["d"],
                                case["oid"], 0.25)
    src = make_mogi(case, 5)
    t = torch.from_numpy(theta).to(cuda)
    r64 = src.residuals(t, "fp64").cpu().numpy()
    r32 = src.residuals(t, "fp32").cpu().numpy()
    scale = np.abs(ref).max(axis=1, keepdims=True)
    e64, e32 = np.max(np.abs(r64 - ref) / scale), np.max(np.abs(r32 - ref) / scale)
    print("mogi residual err of the field max: fp64", e64, "fp32", e32)
    assert e64 < 5e-15     # measured 1.1e-15
    assert e32 < 3e-6      # measured 6.5e-7


def test_mogi_llk_diag(cuda, oracle):
    print("case: mogi llk diag C=1024 N=1000", flush=True)
    case = mogi_case()
    rng = np.random.default_rng(2)
    theta = mogi_theta(rng, 1024)
    theta[:8] = case["truth"] + rng.normal(0, [10, 10, 10, 0.001, 1e-4], (8, 5))
    res = oracle.mogi_residuals(theta, (0, 2, 3, 4), case["x"], case["y"], case["los"], case["d"],
                                case["oid"], 0.25)
    W = 1.0 / np.sqrt(case["cd"])
    norm = norms.normalization(case["cd"])
    ref = oracle.l2_llk(res, W, norm)
    src = make_mogi(case, 5)
    src.whitening("diag", W, norm)
    t = torch.from_numpy(theta).to(cuda)
    # measured: fp64 4.6e-14, fp32 8.4e-6
    for prec, tol in (("fp64", 2e-13), ("fp32", 3e-5)):
        out = torch.empty(1024, dtype=torch.float64, device=cuda)
        src.data_llk(t, out, prec)
        got = out.cpu().numpy()
        rel = np.abs(got - ref) / np.abs(ref)
        print(prec, "max rel llk err", rel.max())
        assert rel.max() < tol, prec


# N = 300, 1,000, 1,100 and 1,120 leave 3 + 0, 4 + 3, 4 + 1 and 4 + 2 of the
# last column tile's 16-column FP64 MFMA tiles with valid columns (the rest
# are skipped); C = 333 and 200 leave part of the last 128-row tile unused
@pytest.mark.parametrize("N,C", [(300, 512), (1000, 512), (1100, 333), (1120, 200)])
def test_mogi_llk_full_cd(cuda, oracle, N, C):
    print(f"case: mogi llk full Cd C={C} N={N}", flush=True)
    case = mogi_case(N=N)
    rng = np.random.default_rng(3)
    xy = np.stack([case["x"], case["y"]], 1)
    cd = norms.exponential_covariance(xy, 25e-6, 5e3)
    case["d"] = case["d"] + np.linalg.cholesky(cd) @ rng.standard_normal(N)
    theta = mogi_theta(rng, C)
    theta[:16] = case["truth"] + rng.normal(0, [50, 50, 50, 0.01, 1e-3], (16, 5))
    res = oracle.mogi_residuals(theta, (0, 2, 3, 4), case["x"], case["y"], case["los"], case["d"],
                                case["oid"], 0.25)
    norm = norms.normalization(cd)
    src = make_mogi(case, 5)
    t = torch.from_numpy(theta).to(cuda)
    for mode in ("exact", "reference"):
        kind, W = norms.whitener(cd, mode)
        ref_w = oracle.exact_whitener(cd) if mode == "exact" else oracle.covariance_inverse_cholesky(cd)
        ref = oracle.l2_llk(res, ref_w, norm)
        if mode == "exact":
            # the exact whitener gives the textbook quadratic form
            q = np.einsum("ci,ci->c", res, np.linalg.solve(cd, res.T).T)
            assert np.allclose(ref, norm - q / 2, rtol=1e-9)
        src.whitening(kind, W, norm)
        # measured (both modes, N = 300 and 1,000): fp64 4.7e-14, fp32 7.8e-6, bf16r 1.5e-4
        for prec, tol in (("fp64", 2e-13), ("fp32", 3e-5), ("bf16r", 5e-4), ("bf16", None)):
            out = torch.empty(C, dtype=torch.float64, device=cuda)
            src.data_llk(t, out, prec)
            got = out.cpu().numpy()
            rel = np.abs(got - ref) / np.abs(ref)
            print(mode, prec, "max rel llk err", rel.max(), "near-truth", rel[:16].max())
            if tol is not None:
                assert rel.max() < tol, (mode, prec)


def test_fp64_whitening_at_full_size(cuda, oracle):
    """The FP64 whitening GEMM at the headline size (16,384 chains x 5,000
    observations, full exponential Cd; 5,000 = 39 x 128 + 8 leaves one valid
    16-column MFMA tile in the last column tile): the Mogi FP64 llk of every
    512th chain against the oracle's residuals whitened by numpy, and every
    chain against the FP32 path."""
    print("case: mogi llk FP64 whitening C=16384 N=5000", flush=True)
    N, C = 5000, 16384
    case = mogi_case(N=N)
    rng = np.random.default_rng(21)
    xy = np.stack([case["x"], case["y"]], 1)
    cd = norms.exponential_covariance(xy, 25e-6, 5e3)
    case["d"] = case["d"] + np.linalg.cholesky(cd) @ rng.standard_normal(N)
    theta = mogi_theta(rng, C)
    theta[:16] = case["truth"] + rng.normal(0, [50, 50, 50, 0.01, 1e-3], (16, 5))
    norm = norms.normalization(cd)
    kind, W = norms.whitener(cd)
    src = make_mogi(case, 5)
    src.whitening(kind, W, norm)
    t = torch.from_numpy(theta).to(cuda)
    out = torch.empty(C, dtype=torch.float64, device=cuda)
    src.data_llk(t, out, "fp64")
    got = out.cpu().numpy()
    pick = np.concatenate([np.arange(16), np.arange(512, C, 512), [C - 1]])
    res = oracle.mogi_residuals(theta[pick], (0, 2, 3, 4), case["x"], case["y"], case["los"], case["d"],
                                case["oid"], 0.25)
    ref = oracle.l2_llk(res, oracle.exact_whitener(cd), norm)
    rel = np.abs(got[pick] - ref) / np.abs(ref)
    out32 = torch.empty(C, dtype=torch.float64, device=cuda)
    src.data_llk(t, out32, "fp32")
    rel32 = np.abs(out32.cpu().numpy() - got) / np.abs(got)
    print("fp64 whitening at full size: max rel llk err", rel.max(), "near truth", rel[:16].max(),
          "| fp32 path against it", rel32.max())
    # measured: 5.1e-15 (FP64), 1.5e-5 (the FP32 path against it)
    assert np.all(np.isfinite(got))
    assert rel.max() < 3e-14
    assert rel32.max() < 6e-5


def test_cdm_forward_vs_oracle(cuda, oracle):
    print("case: cdm forward C=64 N=2000", flush=True)
    rng = np.random.default_rng(4)
    N = 2000
    x = rng.uniform(-15e3, 15e3, N)
    y = rng.uniform(-15e3, 15e3, N)
    los = unit_los(N)
    theta = cdm_theta(rng, 64, oracle)
    layout = (0, 2, 3, 4, 7, -1)
    ref = oracle.cdm_residuals(theta, layout, x, y, los, np.zeros(N), np.zeros(N, np.int32), 0.25)
    src = Source(_lib.CDM, 10, N, 0.25)
    src.layout([0, 2, 3, 4, 7, -1, 0])
    src.observations(x, y, los, np.zeros(N), np.zeros(N, np.int32))
    got = src.forward(torch.from_numpy(theta).to(cuda), "fp32").cpu().numpy()
    err = np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)
    print("cdm normwise err: median", np.median(err), "max", err.max())
    assert err.max() < 1e-4     # measured 2.4e-5


def test_cdm_golden_vectors(cuda):
    """The reference's MATLAB points and the libcdm.py vectors (tests/golden)."""
    import os
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "cdm_vectors.npz"))
    worst = 0.0
    for k, p in enumerate(g["params"]):
        X0, Y0, depth, ax, ay, az, ox, oy, oz, op, nu = p
        X = g["X"] * g["point_scale"][k]
        Y = g["Y"] * g["point_scale"][k]
        M = len(X)
        for comp, los in (("ue", [1, 0, 0]), ("un", [0, 1, 0]), ("uv", [0, 0, 1])):
            src = Source(_lib.CDM, 10, M, nu)
            src.layout([0, 2, 3, 4, 7, -1, 0])
            src.observations(X, Y, np.tile(los, (M, 1)), np.zeros(M), np.zeros(M, np.int32))
            th = torch.tensor([[X0, Y0, depth, op, ax, ay, az, ox, oy, oz]], dtype=torch.float64, device=cuda)
            got = src.forward(th, "fp32").cpu().numpy()[0]
            ref = g[comp][k]
            scale = max(np.abs(g["ue"][k]).max(), np.abs(g["un"][k]).max(), np.abs(g["uv"][k]).max())
            e = np.abs(got - ref).max() / scale
            worst = max(worst, e)
            assert e < 1.2e-4, (k, comp, e)     # measured worst 3.5e-5
    print("golden worst normwise", worst)


def test_cdm_every_side_fp64_matlab_points(cuda):
    """Every side in FP64 (threshold 2, with the FP64 atan2): the reference's
    MATLAB points (models/cdm/tests/libcdm.py:27-46, parameters :103-117) to
    its own bar, 1e-6 relative per component (:83-86); the FP32 path's
    component errors are printed beside it."""
    from test_oracle_pins import KAT_DE, KAT_DN, KAT_DV, KAT_PARAMS, KAT_X, KAT_Y
    p = KAT_PARAMS
    M = len(KAT_X)
    th = torch.tensor([[p["X0"], p["Y0"], p["depth"], p["opening"], p["ax"], p["ay"], p["az"], p["omegaX"],
                        p["omegaY"], p["omegaZ"]]], dtype=torch.float64, device=cuda)
    worst = {}
    for thr in (2.0, 0.2):
        w = 0.0
        for los, want in (([1, 0, 0], KAT_DE), ([0, 1, 0], KAT_DN), ([0, 0, 1], KAT_DV)):
            src = Source(_lib.CDM, 10, M, p["nu"])
            src.layout([0, 2, 3, 4, 7, -1, 0])
            src.fp64_sides(thr)
            src.observations(KAT_X, KAT_Y, np.tile(los, (M, 1)), np.zeros(M), np.zeros(M, np.int32))
            got = src.forward(th, "fp32").cpu().numpy()[0]
            rel = np.abs((got - want) / want)
            w = max(w, rel.max())
            if thr == 2.0:
                assert rel.max() < 1e-6, (los, rel.max())
        worst[thr] = w
    print("MATLAB points, max relative error per component: every side FP64", worst[2.0],
          "| default threshold 0.2", worst[0.2])


def test_cdm_verify_matches_vertices(cuda, oracle):
    rng = np.random.default_rng(5)
    C = 2000
    theta = np.column_stack([rng.uniform(-1e4, 1e4, C), rng.uniform(-1e4, 1e4, C), rng.uniform(1e3, 6e3, C),
                             rng.uniform(1e-3, 1, C), rng.uniform(100, 3000, (C, 3)),
                             rng.uniform(-90, 90, (C, 3))])
    want = np.array([oracle.cdm_vertices(t[0], t[1], t[2], t[4], t[5], t[6], t[7], t[8], t[9])[1]
                     for t in theta], dtype=np.int32)
    src = Source(_lib.CDM, 10, 4, 0.25)
    src.layout([0, 2, 3, 4, 7, -1, 0])
    mask = torch.zeros(C, dtype=torch.int32, device=cuda)
    src.verify(torch.from_numpy(theta).to(cuda), mask)
    assert np.array_equal(mask.cpu().numpy(), want)
    assert 0 < want.sum() < C


def test_cdm_llk_full_cd(cuda, oracle):
    print("case: cdm llk full Cd C=128 N=1200", flush=True)
    rng = np.random.default_rng(6)
    N = 1200
    x = rng.uniform(-15e3, 15e3, N)
    y = rng.uniform(-15e3, 15e3, N)
    los = unit_los(N)
    truth = np.array([[1000.0, -500.0, 4000.0, 0.1, 1000.0, 1200.0, 800.0, 5.0, -8.0, 30.0]])
    cd = norms.exponential_covariance(np.stack([x, y], 1), 64e-6, 5e3)
    u = oracle.cdm_residuals(truth, (0, 2, 3, 4, 7, -1), x, y, los, np.zeros(N), np.zeros(N, np.int32), 0.25)[0]
    d = u + np.linalg.cholesky(cd) @ rng.standard_normal(N)
    theta = cdm_theta(rng, 128, oracle)
    theta[:32] = truth + rng.normal(0, [20, 20, 20, 1e-3, 10, 10, 10, 0.2, 0.2, 0.2], (32, 10))
    res = oracle.cdm_residuals(theta, (0, 2, 3, 4, 7, -1), x, y, los, d, np.zeros(N, np.int32), 0.25)
    norm = norms.normalization(cd)
    kind, W = norms.whitener(cd)
    ref = oracle.l2_llk(res, oracle.exact_whitener(cd), norm)
    src = Source(_lib.CDM, 10, N, 0.25)
    src.layout([0, 2, 3, 4, 7, -1, 0])
    src.observations(x, y, los, d, np.zeros(N, np.int32))
    src.whitening(kind, W, norm)
    t = torch.from_numpy(theta).to(cuda)
    # measured: fp32 1.8e-6, bf16r 1.6e-5
    for prec, tol in (("fp32", 6e-6), ("bf16r", 6e-5), ("bf16", None)):
        out = torch.empty(128, dtype=torch.float64, device=cuda)
        src.data_llk(t, out, prec)
        got = out.cpu().numpy()
        rel = np.abs(got - ref) / np.abs(ref)
        absn = np.abs(got - ref)[:32]
        print(prec, "max rel llk err", rel.max(), "near-truth abs err", absn.max())
        if tol is not None:
            assert rel.max() < tol, prec
        # a ragged call (77 chains, the whitening tile's rows past 77 still hold
        # the previous call's residuals): each chain's llk is its own, bit for bit
        out77 = torch.empty(77, dtype=torch.float64, device=cuda)
        src.data_llk(t[:77].contiguous(), out77, prec)
        assert np.array_equal(out77.cpu().numpy(), got[:77]), prec


def test_fields_write_only_the_callers_rows(cuda, oracle):
    """altar_source_data_llk_fields after a larger call on the same handle:
    the whitened residuals fill exactly C rows (rows past C keep a sentinel)
    and equal W r of the oracle; the llk matches the plain entry point."""
    from altar_amd import _lib as L
    rng = np.random.default_rng(44)
    case = mogi_case(N=300)
    cd = norms.exponential_covariance(np.stack([case["x"], case["y"]], 1), 25e-6, 5e3)
    src = make_mogi(case, 5)
    kind, W = norms.whitener(cd)
    norm = norms.normalization(cd)
    src.whitening(kind, W, norm)
    big = torch.from_numpy(mogi_theta(rng, 1024)).to(cuda)
    out = torch.empty(1024, dtype=torch.float64, device=cuda)
    src.data_llk(big, out, "fp32")                 # capacity 1024 chains
    C, Npad, guard = 256, 384, 512
    theta = mogi_theta(rng, C)
    t = torch.from_numpy(theta).to(cuda)
    buf = torch.full((C + guard, Npad), 12345.0, dtype=torch.float32, device=cuda)
    llk = torch.empty(C, dtype=torch.float64, device=cuda)
    L.call("altar_source_data_llk_fields", src.handle, L.ptr(t), C, L.ptr(llk), L.FP32, L.ptr(buf), Npad, None)
    b = buf.cpu().numpy()
    assert np.all(b[C:] == 12345.0)
    res = oracle.mogi_residuals(theta, (0, 2, 3, 4), case["x"], case["y"], case["los"], case["d"], case["oid"], 0.25)
    wr = res @ oracle.exact_whitener(cd).T
    assert np.max(np.abs(b[:C, :300] - wr)) < 1e-4 * np.max(np.abs(wr))
    assert np.all(b[:C, 300:] == 0.0)
    ref = torch.empty(C, dtype=torch.float64, device=cuda)
    src.data_llk(t, ref, "fp32")
    # both are FP32 squared norms summed per tile: equal to FP32 rounding of
    # the norm (normalization - llk), not of the llk, which can be near zero
    a, r = llk.cpu().numpy(), ref.cpu().numpy()
    assert np.all(np.abs(a - r) <= 1e-6 * (abs(norm) + np.abs(r)))


def test_cdm_fp64_side_policy(cuda, oracle):
    """The FP64-side threshold: 0 puts every side in FP32, 2 every side in
    FP64 with the FP64 atan2 (<= 1e-10 normwise); the census counts the
    sides of each kind."""
    print("case: cdm fp64 side policy C=48 N=1500", flush=True)
    rng = np.random.default_rng(14)
    N = 1500
    x = rng.uniform(-15e3, 15e3, N)
    y = rng.uniform(-15e3, 15e3, N)
    los = unit_los(N)
    theta = cdm_theta(rng, 48, oracle)
    layout = (0, 2, 3, 4, 7, -1)
    ref = oracle.cdm_residuals(theta, layout, x, y, los, np.zeros(N), np.zeros(N, np.int32), 0.25)
    src = Source(_lib.CDM, 10, N, 0.25)
    src.layout([0, 2, 3, 4, 7, -1, 0])
    src.observations(x, y, los, np.zeros(N), np.zeros(N, np.int32))
    t = torch.from_numpy(theta).to(cuda)
    errs, census = {}, {}
    for thr in (0.0, 0.4, 2.0):
        src.fp64_sides(thr)
        got = src.forward(t, "fp32").cpu().numpy()
        errs[thr] = (np.abs(got - ref).max(axis=1) / np.abs(ref).max(axis=1)).max()
        census[thr] = src.side_census(48)
    print("normwise err by threshold", errs, "census", census)
    assert census[0.0][1] == 0 and census[2.0][0] == 0
    assert census[0.0][3] == 0 and sum(census[0.0][:3]) == 12 * 48
    for thr in (0.4, 2.0):
        assert sum(census[thr][:3]) == 12 * 48 and census[thr][2] == census[0.0][2]
    assert errs[2.0] < 1e-10    # measured 2.4e-11
    assert errs[0.4] < 2e-5     # measured 3.5e-6
    with pytest.raises(RuntimeError):
        src.fp64_sides(-1.0)


def test_fields_diagonal_and_prewhitened(cuda, oracle):
    """altar_source_data_llk_fields with a diagonal whitener (Mogi) and with a
    pre-whitened linear source: the fields equal w r and W (G theta - d) of the
    oracle (zero in the padding columns), and the llk equals the plain entry
    point's to FP32 rounding of the norm."""
    from altar_amd import _lib as L
    from altar_amd import Job
    from altar_amd.linear import Linear, default_psets
    rng = np.random.default_rng(45)
    case = mogi_case(N=300)
    src = make_mogi(case, 5)
    norm = norms.normalization(case["cd"])
    src.whitening("diag", 1.0 / np.sqrt(case["cd"]), norm)
    C, Npad = 256, 384
    theta = mogi_theta(rng, C)
    t = torch.from_numpy(theta).to(cuda)
    buf = torch.full((C, Npad), 7.0, dtype=torch.float32, device=cuda)
    llk = torch.empty(C, dtype=torch.float64, device=cuda)
    L.call("altar_source_data_llk_fields", src.handle, L.ptr(t), C, L.ptr(llk), L.FP32, L.ptr(buf), Npad, None)
    res = oracle.mogi_residuals(theta, (0, 2, 3, 4), case["x"], case["y"], case["los"], case["d"], case["oid"], 0.25)
    want = res / np.sqrt(case["cd"])[None, :]
    b = buf.cpu().numpy()
    err = np.max(np.abs(b[:, :300] - want)) / np.max(np.abs(want))
    print("mogi diagonal fields: max error of the field max", err)
    assert err < 1e-5 and np.all(b[:, 300:] == 0.0)
    ref = torch.empty(C, dtype=torch.float64, device=cuda)
    src.data_llk(t, ref, "fp32")
    a, r = llk.cpu().numpy(), ref.cpu().numpy()
    assert np.all(np.abs(a - r) <= 1e-6 * (abs(norm) + np.abs(r)))
    # pre-whitened linear source
    N, P = 300, 20
    G = rng.normal(size=(N, P)) * 1e-2
    cd = norms.exponential_covariance(rng.uniform(0, 3e4, (N, 2)), 25e-6, 5e3)
    d = G @ rng.uniform(0, 2, P) + np.linalg.cholesky(cd) @ rng.normal(size=N)
    m = Linear(default_psets(P), G, d, cd, job=Job(chains=C)).initialize()
    th = rng.normal(1.0, 0.5, (C, P))
    t = torch.from_numpy(th).to(cuda)
    buf.fill_(7.0)
    L.call("altar_source_data_llk_fields", m.source.handle, L.ptr(t), C, L.ptr(llk), L.FP32, L.ptr(buf), Npad, None)
    want = oracle.linear_residuals(th, G, d) @ oracle.exact_whitener(cd).T
    b = buf.cpu().numpy()
    err = np.max(np.abs(b[:, :N] - want)) / np.max(np.abs(want))
    print("pre-whitened linear fields: max error of the field max", err)
    assert err < 1e-5 and np.all(b[:, N:] == 0.0)
    m.source.data_llk(t, ref, "fp32")
    a, r = llk.cpu().numpy(), ref.cpu().numpy()
    assert np.all(np.abs(a - r) <= 1e-6 * (abs(m.normalization) + np.abs(r)))


def test_cdm_llk_fp64_reference_precision(cuda, oracle):
    """The reference's precision end to end: every CDM side in FP64 (FP64
    atan2), the residual formed in FP64, |W r|^2 by the FP64 MFMA whitening
    GEMM with FP64 sums, against the oracle's FP64 llk.  With FP32 sides the FP64
    llk is refused (the sides would not be reference precision)."""
    print("case: cdm llk FP64 reference precision C=64 N=1200", flush=True)
    rng = np.random.default_rng(16)
    N, C = 1200, 64
    x = rng.uniform(-15e3, 15e3, N)
    y = rng.uniform(-15e3, 15e3, N)
    los = unit_los(N)
    truth = np.array([[1000.0, -500.0, 4000.0, 0.1, 1000.0, 1200.0, 800.0, 5.0, -8.0, 30.0]])
    cd = norms.exponential_covariance(np.stack([x, y], 1), 64e-6, 5e3)
    u = oracle.cdm_residuals(truth, (0, 2, 3, 4, 7, -1), x, y, los, np.zeros(N), np.zeros(N, np.int32), 0.25)[0]
    d = u + np.linalg.cholesky(cd) @ rng.standard_normal(N)
    theta = cdm_theta(rng, C, oracle)
    theta[:16] = truth + rng.normal(0, [20, 20, 20, 1e-3, 10, 10, 10, 0.2, 0.2, 0.2], (16, 10))
    res = oracle.cdm_residuals(theta, (0, 2, 3, 4, 7, -1), x, y, los, d, np.zeros(N, np.int32), 0.25)
    norm = norms.normalization(cd)
    kind, W = norms.whitener(cd)
    ref = oracle.l2_llk(res, oracle.exact_whitener(cd), norm)
    src = Source(_lib.CDM, 10, N, 0.25)
    src.layout([0, 2, 3, 4, 7, -1, 0])
    src.observations(x, y, los, d, np.zeros(N, np.int32))
    src.whitening(kind, W, norm)
    t = torch.from_numpy(theta).to(cuda)
    out = torch.empty(C, dtype=torch.float64, device=cuda)
    with pytest.raises(RuntimeError):
        src.data_llk(t, out, "fp64")
    src.fp64_sides(2.0)
    src.data_llk(t, out, "fp64")
    got = out.cpu().numpy()
    rel = np.abs(got - ref) / np.abs(ref)
    print("fp64 reference precision: max rel llk err", rel.max(), "abs max", np.abs(got - ref).max(),
          "near truth", np.abs(got - ref)[:16].max())
    assert rel.max() < 3e-11     # measured 6.3e-12
