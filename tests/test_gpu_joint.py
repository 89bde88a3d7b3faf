"""Joint Mogi + CDM model on three datasets with a block-diagonal Cd,
against the FP64 oracle's Mogi and CDM restatements summed per observation
(oracle.mogi_residuals + oracle.cdm_residuals) and whitened per dataset.

Bars, 3-5 times the largest error measured on an MI355X: displacements 1e-5
of the field max per sample (FP32 with FP64 near-singular CDM sides); llk
relative 2e-5 (FP32 whitening) and 2.5e-4 (BF16 + refinement); the block
whitener equals one lower-triangular whitener over the whole (padded)
observation vector built from the same blocks."""
import numpy as np
import pytest
import torch

from altar_amd import Job, cases, norms
from altar_amd.joint import Joint

pytestmark = pytest.mark.gpu
MOGI_LAYOUT = (0, 2, 3, 14)
CDM_LAYOUT = (4, 6, 7, 8, 11, -1)


def oracle_forward(oracle):
    def fwd(th, x, y, los, oid):
        z = np.zeros(len(x))
        return (oracle.mogi_residuals(th, MOGI_LAYOUT, x, y, los, z, oid, 0.25) +
                oracle.cdm_residuals(th, CDM_LAYOUT, x, y, los, z, oid, 0.25))
    return fwd


def joint_thetas(rng, oracle, truth, C):
    rows = []
    scale = np.array([300, 300, 200, 0.05, 100, 100, 100, 0.01, 50, 50, 50, 2, 2, 2, 0.002, 0.002, 0.002])
    while len(rows) < C:
        p = truth + rng.normal(size=17) * scale * (1 if len(rows) % 2 else 5)
        if not oracle.cdm_vertices(p[4], p[5], p[6], p[8], p[9], p[10], p[11], p[12], p[13])[1]:
            rows.append(p)
    return np.array(rows)


def test_joint_forward_and_llk(cuda, oracle):
    print("case: joint 2 x 900 InSAR + 40 GNSS, C=256", flush=True)
    rng = np.random.default_rng(21)
    fwd = oracle_forward(oracle)
    case = cases.joint_case(n_insar=900, n_gnss=40, forward=fwd)
    ds = case["datasets"]
    C = 256
    theta = joint_thetas(rng, oracle, case["truth"], C)
    model = Joint(case["psets"], ds, job=Job(chains=C)).initialize()
    t = torch.from_numpy(theta).to(cuda)
    # forward at the true observations
    x = np.concatenate([d.x for d in ds]); y = np.concatenate([d.y for d in ds])
    los = np.concatenate([d.los for d in ds])
    oid = np.concatenate([np.full(len(d.x), k, np.int32) for k, d in enumerate(ds)])
    pred = fwd(theta, x, y, los, oid)
    got = model.forwardModelBatched(t, "fp32").cpu().numpy()
    err = np.abs(got - pred).max(axis=1) / np.abs(pred).max(axis=1)
    print("joint forward normwise err: median", np.median(err), "max", err.max())
    assert err.max() < 1e-5     # measured 2.8e-6
    # oracle llk: per dataset, exact whitener of its own Cd
    dd = np.concatenate([d.d for d in ds])
    res = pred - dd[None, :]
    ref = np.zeros(C)
    at = 0
    for d in ds:
        n = len(d.x)
        W = oracle.exact_whitener(np.diag(d.cd) if np.ndim(d.cd) == 1 else d.cd)
        ref += oracle.l2_llk(res[:, at:at + n], W, oracle.normalization(
            np.diag(d.cd) if np.ndim(d.cd) == 1 else d.cd))
        at += n
    # measured: fp32 4.8e-6, bf16r 6.1e-5
    for prec, tol in (("fp32", 2e-5), ("bf16r", 2.5e-4)):
        llk = torch.empty(C, dtype=torch.float64, device=cuda)
        model.source.data_llk(t, llk, prec)
        rel = np.abs(llk.cpu().numpy() - ref) / np.abs(ref)
        print("joint", prec, "max rel llk err", rel.max())
        assert rel.max() < tol, prec
    # the same blocks as one lower-triangular whitener over the padded vector
    Wfull = np.zeros((model.observations, model.observations))
    for off, d in zip(model.block_offsets, ds):
        kind, W = norms.whitener(d.cd)
        W = np.diag(W) if kind == "diag" else W
        Wfull[off:off + len(d.x), off:off + len(d.x)] = W
    model.source.whitening("lower", Wfull, model.normalization)
    llk2 = torch.empty(C, dtype=torch.float64, device=cuda)
    model.source.data_llk(t, llk2, "fp32")
    rel = np.abs(llk2.cpu().numpy() - ref) / np.abs(ref)
    print("joint, the blocks as one lower-triangular whitener: max rel llk err", rel.max())
    assert rel.max() < 2e-5
    # the free-surface check covers the CDM part
    bad = theta[:4].copy()
    bad[:, 6] = 50.0      # CDM centre 50 m deep with km-sized axes
    mask = torch.zeros(4, dtype=torch.int32, device=cuda)
    model.source.verify(torch.from_numpy(bad).to(cuda), mask)
    assert mask.cpu().numpy().tolist() == [1, 1, 1, 1]


def test_joint_fields_and_surrogate(cuda, oracle):
    """Whitened fields of the joint source with its block whitener (each
    dataset's columns hold W_b r_b, padding columns zero) against the oracle,
    and a surrogate fitted on them screening a short delayed-acceptance walk
    (every move accounted for: accepted + rejected + unlikely = C x steps)."""
    print("case: joint fields + surrogate DA, 2 x 900 InSAR + 40 GNSS, C=1024", flush=True)
    from altar_amd import _lib
    from altar_amd.bayesian import Annealer, COV
    from altar_amd.bayesian.da import DelayedAcceptance
    from altar_amd.surrogate import Surrogate
    rng = np.random.default_rng(22)
    fwd = oracle_forward(oracle)
    case = cases.joint_case(n_insar=900, n_gnss=40, forward=fwd)
    ds = case["datasets"]
    C = 256
    theta = joint_thetas(rng, oracle, case["truth"], C)
    model = Joint(case["psets"], ds, job=Job(chains=C)).initialize()
    t = torch.from_numpy(theta).to(cuda)
    Np = (model.observations + 127) // 128 * 128
    wr = torch.full((C, Np), 7.0, dtype=torch.float32, device=cuda)
    llk = torch.empty(C, dtype=torch.float64, device=cuda)
    _lib.call("altar_source_data_llk_fields", model.source.handle, _lib.ptr(t), C, _lib.ptr(llk), _lib.FP32,
              _lib.ptr(wr), Np, None)
    got = wr.cpu().numpy().astype(np.float64)
    x = np.concatenate([d.x for d in ds]); y = np.concatenate([d.y for d in ds])
    los = np.concatenate([d.los for d in ds])
    oid = np.concatenate([np.full(len(d.x), k, np.int32) for k, d in enumerate(ds)])
    res = fwd(theta, x, y, los, oid) - np.concatenate([d.d for d in ds])[None, :]
    at = 0
    real = np.zeros(Np, bool)
    for off, d in zip(model.block_offsets, ds):
        n = len(d.x)
        cd = np.diag(d.cd) if np.ndim(d.cd) == 1 else d.cd
        want = res[:, at:at + n] @ oracle.exact_whitener(cd).T
        assert np.max(np.abs(got[:, off:off + n] - want)) < 1e-4 * np.max(np.abs(want))
        real[off:off + n] = True
        at += n
    assert np.all(got[:, ~real] == 0.0)
    # a short anneal with the surrogate screen
    Cd = 1024
    job = Job(chains=Cd, steps=3, seed=31)
    model = Joint(case["psets"], ds, job=job).initialize()
    sur = Surrogate(model, K=32, capacity=6144, max_batch=Cd, seed=1)
    da = DelayedAcceptance(sur, train_beta=0.05, max_rmse=1.0, steps=3,
                           fit_kwargs=dict(epochs=10, batch=1024, lr=2e-3, pca_sample=4096))
    ann = Annealer(da, COV(), seed=job.seed).initialize(job)
    ann.model = model
    w = ann.worker
    w.start(ann)
    used = []
    for _ in range(6):
        w.cool(ann)
        acc, rej, unl = w.walk(ann)
        assert acc + rej + unl == Cd * da.steps
        used.append(da.mode_history[-1]["mode"])
        w.resample(ann, (acc, rej, unl))
        if w.beta + job.tolerance >= 1:
            break
    print("joint DA modes", used, "validation rmse", sur.validation_rmse)
    assert sur.trained and "da" in used
    # the device field RMSE (GEMMs through each dataset's chol(Cd_b)) against
    # the host form on the current chains
    st = w.step
    th = st.theta[:512].contiguous()
    Np2 = sur.Npad
    wr2 = torch.empty((512, Np2), dtype=torch.float32, device=cuda)
    llk2 = torch.empty(512, dtype=torch.float64, device=cuda)
    _lib.call("altar_source_data_llk_fields", model.source.handle, _lib.ptr(th), 512, _lib.ptr(llk2), _lib.FP32,
              _lib.ptr(wr2), Np2, None)
    dev_rmse = sur.field_rmse_device(th, wr2)
    host_rmse = sur.field_rmse(sur.predict_wr(th), wr2[:, :sur.N].double().cpu().numpy())
    print("field rmse device", dev_rmse, "host", host_rmse)
    assert abs(dev_rmse / host_rmse - 1) < 1e-3


def test_joint_fused_trajectory_matches_separate_calls(cuda, oracle):
    """Surrogate trajectories of the joint source (17 parameters, the CDM part
    checked against the free surface through the joint layout): the one-launch
    kernel and the separate calls give the same states, likelihoods and
    counters bit for bit."""
    print("case: joint fused vs separate trajectories, 2 x 400 InSAR + 30 GNSS, C=700, 2 x 5 sub-steps", flush=True)
    from altar_amd.bayesian import Annealer, COV, Metropolis
    from altar_amd.bayesian.da import DelayedAcceptance
    from altar_amd.surrogate import Surrogate
    case = cases.joint_case(n_insar=400, n_gnss=30, forward=oracle_forward(oracle))
    C = 700
    job = Job(chains=C, steps=1, seed=8)
    model = Joint(case["psets"], case["datasets"], job=job).initialize()
    ann = Annealer(Metropolis(), COV(), seed=8).initialize(job)
    ann.model = model
    w = ann.worker
    w.start(ann)
    w.cool(ann)
    step = w.step
    ann.sampler.prepareSamplingPDF(ann, step)
    P = step.parameters
    sur = Surrogate(model, K=32, capacity=1024, max_batch=C, seed=2)
    rng = np.random.default_rng(5)
    sur.e_d = torch.from_numpy(rng.normal(0.0, 1.0, 32)).to(cuda)
    sur.s_d = torch.from_numpy(rng.uniform(0.05, 0.2, 32)).to(cuda)
    sur.cperp = 2.0
    sur.trained = True
    runs = []
    for fused in (False, True):
        da = DelayedAcceptance(sur, steps=2, substeps=5, substep_scale=2.5, fused=fused)
        assert da.fused_supported(model, P) == fused
        da.sigma_chol = ann.sampler.sigma_chol
        st = step.clone()
        stats = da._walk_trajectory(ann, st)
        torch.cuda.synchronize()
        runs.append((stats, st, list(da.substep_counts), {k: v.clone() for k, v in da._dws.items()}))
    (s0, st0, c0, w0), (s1, st1, c1, w1) = runs
    print("exact-stage stats", s0, s1, "sub-steps", c0, c1)
    assert s0 == s1 and c0 == c1 and c0[0] > 0 and c0[1] > 0
    for name in ("theta", "prior", "data", "posterior"):
        assert torch.equal(getattr(st0, name), getattr(st1, name)), name
    for name in ("phi", "prior_phi", "post_phi", "post_s"):
        assert torch.equal(w0[name], w1[name]), name


@pytest.fixture(scope="module")
def joint_full(cuda, oracle):
    """The joint case at its full size: 2 x 20,000 InSAR points with full
    exponential Cd blocks and 200 x 3 GNSS components (40,600 observations),
    data from the oracle's forward models."""
    print("setup: joint case 2 x 20,000 InSAR + 600 GNSS (host Cholesky of two 20,000^2 blocks)", flush=True)
    fwd = oracle_forward(oracle)
    case = cases.joint_case(forward=fwd)
    return case, fwd


def joint_oracle_llk(case, fwd, theta):
    """Per-dataset FP64 llk by triangular solves with chol(Cd_b): no explicit
    inverse of a 20,000 x 20,000 block."""
    from scipy.linalg import solve_triangular
    ds = case["datasets"]
    x = np.concatenate([d.x for d in ds]); y = np.concatenate([d.y for d in ds])
    los = np.concatenate([d.los for d in ds])
    oid = np.concatenate([np.full(len(d.x), k, np.int32) for k, d in enumerate(ds)])
    res = fwd(theta, x, y, los, oid) - np.concatenate([d.d for d in ds])[None, :]
    llk = np.zeros(len(theta))
    at = 0
    for d in ds:
        n = len(d.x)
        r = res[:, at:at + n]
        if np.ndim(d.cd) == 1:
            q = np.sum(r * r / d.cd[None, :], axis=1)
            norm = norms.normalization(d.cd)
        else:
            z = solve_triangular(d.chol, r.T, lower=True, check_finite=False)
            q = np.sum(z * z, axis=0)
            norm = norms.normalization(d.cd, chol=d.chol)
        llk += norm - 0.5 * q
        at += n
    return llk


def test_joint_full_size_llk_vs_oracle(cuda, oracle, joint_full):
    """FP32 and BF16 + refinement llk at the full joint size against the
    oracle on 16 chains (half near the truth, half far)."""
    case, fwd = joint_full
    rng = np.random.default_rng(23)
    C = 16
    print(f"case: joint full size, {C} chains vs the oracle", flush=True)
    theta = joint_thetas(rng, oracle, case["truth"], C)
    ref = joint_oracle_llk(case, fwd, theta)
    model = Joint(case["psets"], case["datasets"], job=Job(chains=C)).initialize()
    assert model.data_count == 40600
    t = torch.from_numpy(theta).to(cuda)
    # measured: fp32 1.2e-6, bf16r 4.2e-6
    for prec, tol in (("fp32", 5e-6), ("bf16r", 1.5e-5)):
        llk = torch.empty(C, dtype=torch.float64, device=cuda)
        model.source.data_llk(t, llk, prec)
        got = llk.cpu().numpy()
        rel = np.abs(got - ref) / np.abs(ref)
        print("joint full size", prec, "llk relative error: median", np.median(rel), "max", rel.max(),
              "| abs max", np.abs(got - ref).max())
        assert rel.max() < tol, prec


def test_joint_full_size_65536_chains(cuda, oracle, joint_full):
    """65,536 chains at the full joint size: every llk finite, a second
    evaluation bit-identical, and a slice of chains spread over the whole
    range (first, last, around the 128-row tile and workgroup boundaries)
    equal to the oracle's per-block llk."""
    case, fwd = joint_full
    rng = np.random.default_rng(24)
    C = 65536
    print(f"case: joint full size, {C} chains: finite, replay, slice vs the oracle", flush=True)
    truth = case["truth"]
    scale = np.array([300, 300, 200, 0.05, 100, 100, 100, 0.01, 50, 50, 50, 2, 2, 2, 0.002, 0.002, 0.002])
    theta = truth + rng.normal(size=(C, 17)) * scale
    theta[:, 6] = np.maximum(theta[:, 6], 3500.0)      # keep every CDM below the free surface
    bad = [c for c in range(0, C, 997)
           if oracle.cdm_vertices(*theta[c, [4, 5, 6, 8, 9, 10, 11, 12, 13]])[1]]
    assert not bad
    model = Joint(case["psets"], case["datasets"], job=Job(chains=C)).initialize()
    t = torch.from_numpy(theta).to(cuda)
    a = torch.empty(C, dtype=torch.float64, device=cuda)
    b = torch.empty(C, dtype=torch.float64, device=cuda)
    model.source.data_llk(t, a, "fp32")
    model.source.data_llk(t, b, "fp32")
    la, lb = a.cpu().numpy(), b.cpu().numpy()
    assert np.all(np.isfinite(la))
    assert np.array_equal(la, lb)
    pick = np.array([0, 1, 127, 128, 255, 256, 32767, 32768, 65407, 65408, 65534, 65535])
    ref = joint_oracle_llk(case, fwd, theta[pick])
    rel = np.abs(la[pick] - ref) / np.abs(ref)
    print("joint 65,536 chains: llk range", la.min(), la.max(), "| slice vs oracle max rel", rel.max())
    assert rel.max() < 2.5e-8      # measured 5.2e-9 (chains near the truth)
