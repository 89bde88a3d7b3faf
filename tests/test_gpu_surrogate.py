"""Surrogate network kernels against the FP64 oracle (oracle/mlp.py).

* forward: relative error of the network output <= 2e-2 (BF16 operands)
* gradients: per layer, relative Frobenius error <= 5e-2 and cosine >= 0.995
  against the chain-rule gradients of the same (BF16-rounded) weights
* training: Adam on a smooth target reduces the loss by 10x; a native epoch
  equals its batches run one by one, bit for bit
* delayed acceptance: the screen list, the stage-2 rule and the counters
  match a numpy restatement of Christen and Fox (2005) on the same uniforms
"""
import numpy as np
import pytest
import torch

from altar_amd import _lib
from altar_amd.surrogate import MLP

pytestmark = pytest.mark.gpu


def bf16(a):
    return torch.from_numpy(np.asarray(a, np.float32)).bfloat16().float().numpy().astype(np.float64)


def setup_net(seed=0, nin=10, nhid=256, nout=64):
    net = MLP(nin, nhid, 4, nout, max_batch=1024).init(seed)
    lo, hi = -np.ones(nin) * 5, np.ones(nin) * 5
    net.input_scaling(lo, hi)
    layers = [tuple(a.astype(np.float64) for a in net.layer(l)) for l in range(4)]
    return net, layers, lo, hi


@pytest.mark.parametrize("nhid", [256, 320])
def test_mlp_forward_and_gradients(cuda, nhid):
    from oracle import mlp as om
    net, layers, lo, hi = setup_net(nhid=nhid)
    rng = np.random.default_rng(0)
    B = 200   # not a multiple of the 64-row tile
    theta = rng.uniform(-5, 5, (B, 10))
    target = rng.normal(size=(B, 64))
    w = rng.random(64)
    w /= w.sum()
    t = torch.from_numpy(theta).to(cuda)
    z = net.forward(t).cpu().numpy().astype(np.float64)
    x = bf16(om.scale_inputs(theta, lo, hi))
    ref_layers = [(bf16(W), b) for W, b in layers]
    y, _ = om.forward(ref_layers, x)
    assert np.linalg.norm(z - y) / np.linalg.norm(y) < 2e-2
    tgt = torch.from_numpy(target.astype(np.float32)).to(cuda)
    wd = torch.from_numpy(w.astype(np.float32)).to(cuda)
    loss = torch.zeros(1, dtype=torch.float32, device=cuda)
    net.backward(t, tgt, wd, loss)
    ref_loss, grads = om.loss_and_grads(ref_layers, x, target, w)
    assert abs(loss.item() - ref_loss) / ref_loss < 2e-2
    for l in range(4):
        gW, gb = (a.astype(np.float64) for a in net.layer(l, grads=True))
        rW, rb = grads[l]
        rel = np.linalg.norm(gW - rW) / np.linalg.norm(rW)
        cos = np.sum(gW * rW) / (np.linalg.norm(gW) * np.linalg.norm(rW))
        print(f"layer {l}: dW rel {rel:.3e} cos {cos:.5f}; db rel {np.linalg.norm(gb - rb) / np.linalg.norm(rb):.3e}")
        assert rel < 5e-2 and cos > 0.995
        assert np.linalg.norm(gb - rb) / np.linalg.norm(rb) < 5e-2


def test_mlp_learns_smooth_map(cuda):
    net, _, lo, hi = setup_net(1, nin=4, nout=64)
    rng = np.random.default_rng(2)
    B = 1024
    A = rng.normal(size=(4, 64))

    def f(th):
        return np.tanh(th / 5.0 @ A)

    w = torch.full((64,), 1.0 / 64, dtype=torch.float32, device=cuda)
    loss = torch.zeros(1, dtype=torch.float32, device=cuda)
    losses = []
    for it in range(300):
        th = rng.uniform(-5, 5, (B, 4))
        t = torch.from_numpy(th).to(cuda)
        tgt = torch.from_numpy(f(th).astype(np.float32)).to(cuda)
        net.backward(t, tgt, w, loss)
        net.adam(1e-3)
        losses.append(loss.item())
    print("loss first/last", losses[0], losses[-1])
    assert losses[-1] < losses[0] / 10


def test_train_epoch_matches_batch_calls(cuda):
    """One native epoch (batch gather, gradients, Adam with the cosine rate)
    gives the weights of the same batches run through altar_mlp_backward and
    altar_mlp_adam one by one, bit for bit; its loss is the last batch's."""
    import math
    rng = np.random.default_rng(4)
    N, batch, nb, lr, lr_final, total = 3000, 640, 4, 2e-3, 1e-4, 12
    theta = torch.from_numpy(rng.uniform(-5, 5, (N, 10))).to(cuda)
    A = rng.normal(size=(10, 64))
    target = torch.from_numpy(np.tanh(theta.cpu().numpy() / 5.0 @ A).astype(np.float32)).to(cuda)
    w = torch.from_numpy((rng.random(64) / 32).astype(np.float32)).to(cuda)
    order = rng.permutation(N)[: nb * batch].astype(np.int32)
    a, _, lo, hi = setup_net(3)
    b, _, _, _ = setup_net(3)
    for net in (a, b):   # a few steps first, so that Adam's moments are not zero
        for it in range(2):
            net.backward(theta[it * 100:(it + 1) * 100 + 37], target[it * 100:(it + 1) * 100 + 37], w)
            net.adam(1e-3)
    loss_a = torch.zeros(1, dtype=torch.float32, device=cuda)
    order_d = torch.from_numpy(order).to(cuda)
    _lib.call("altar_mlp_train_epoch", a.handle, _lib.ptr(theta), _lib.ptr(target), 64, _lib.ptr(w),
              _lib.ptr(order_d), nb, batch, 4, total, lr, lr_final, 0.9, 0.999, 1e-8, _lib.ptr(loss_a), _lib.stream())
    loss_b = torch.zeros(1, dtype=torch.float32, device=cuda)
    for i in range(nb):
        sel = torch.from_numpy(order[i * batch:(i + 1) * batch].astype(np.int64)).to(cuda)
        b.backward(theta[sel].contiguous(), target[sel].contiguous(), w, loss_b)
        # the epoch takes both rates as FP32 and forms its schedule in FP64
        hi_, lo_ = float(np.float32(lr)), float(np.float32(lr_final))
        frac = (4 + i) / (total - 1)
        b.adam(lo_ + 0.5 * (hi_ - lo_) * (1.0 + math.cos(math.pi * frac)))
    torch.cuda.synchronize()
    for l in range(4):
        for x, y in zip(a.layer(l), b.layer(l)):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), f"layer {l} differs"
    print("epoch loss", loss_a.item(), "last batch loss", loss_b.item())
    assert abs(loss_a.item() - loss_b.item()) <= 1e-6 * abs(loss_b.item())


def test_delayed_acceptance_rule(cuda, oracle):
    rng = np.random.default_rng(3)
    C, P = 3000, 4
    beta = 0.6
    theta = rng.normal(size=(C, P))
    prior, data = rng.normal(size=C), rng.normal(size=C) * 5
    post = prior + beta * data
    post_s = post + rng.normal(0, 0.5, C)
    cand = rng.normal(size=(C, P))
    cprior, cdata_s = rng.normal(size=C), rng.normal(size=C) * 5
    cdata = cdata_s + rng.normal(0, 0.5, C)
    mask = (rng.random(C) < 0.1).astype(np.int32)
    seed, stage, step = 9, 4, 2
    # numpy restatement
    cpost_s = cprior + beta * cdata_s
    u1 = oracle.uniforms(seed, stage, step, _lib.TAG_SCREEN, C)
    passed = (mask == 0) & (np.log(u1) <= cpost_s - post_s)
    idx = np.nonzero(passed)[0]
    u2 = oracle.uniforms(seed, stage, step, _lib.TAG_ACCEPT, C)
    cpost = cprior + beta * cdata
    acc = np.zeros(C, bool)
    acc[idx] = np.log(u2[idx]) <= (cpost[idx] - post[idx]) - (cpost_s[idx] - post_s[idx])
    keep = []

    def D(a):   # device copy kept alive until the kernels have run
        t = torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
        keep.append(t)
        return t

    th_d, pr_d, da_d, po_d, ps_d = D(theta.copy()), D(prior.copy()), D(data.copy()), D(post.copy()), D(post_s.copy())
    cpost_s_d = torch.empty(C, dtype=torch.float64, device=cuda)
    index = torch.empty(2 * C, dtype=torch.int32, device=cuda)
    count = torch.zeros(1, dtype=torch.int32, device=cuda)
    counters = torch.zeros(4, dtype=torch.int64, device=cuda)
    _lib.call("altar_da_screen", _lib.ptr(D(cprior)), _lib.ptr(D(cdata_s)), _lib.ptr(ps_d), _lib.ptr(D(mask)), beta,
              C, seed, stage, step, _lib.ptr(cpost_s_d), _lib.ptr(index), _lib.ptr(count), _lib.ptr(counters), None)
    n = int(count.item())
    assert n == len(idx)
    assert np.array_equal(index[:n].cpu().numpy(), idx)
    packed = torch.empty((n, P), dtype=torch.float64, device=cuda)
    cd = D(cand)
    _lib.call("altar_da_pack", _lib.ptr(cd), _lib.ptr(index), n, P, _lib.ptr(packed), None)
    assert np.array_equal(packed.cpu().numpy(), cand[idx])
    _lib.call("altar_da_accept", _lib.ptr(th_d), _lib.ptr(pr_d), _lib.ptr(da_d), _lib.ptr(po_d), _lib.ptr(ps_d),
              _lib.ptr(cd), _lib.ptr(D(cprior)), _lib.ptr(D(cdata[idx])), _lib.ptr(cpost_s_d), _lib.ptr(index), n,
              beta, P, seed, stage, step, None, _lib.ptr(counters), None)
    th = th_d.cpu().numpy()
    assert np.array_equal(th[acc], cand[acc]) and np.array_equal(th[~acc], theta[~acc])
    assert np.array_equal(po_d.cpu().numpy()[acc], cpost[acc])
    assert np.array_equal(ps_d.cpu().numpy()[acc], cpost_s[acc])
    cnt = counters.cpu().numpy()
    assert cnt[0] == acc.sum() and cnt[1] == mask.sum()
    assert cnt[3] == ((mask == 0) & ~passed).sum() and cnt[2] == len(idx) - acc.sum()


def test_trajectory_kernels_match_restatement(cuda, oracle):
    """Sub-step proposal, surrogate Metropolis sub-step, moved list: bit-exact
    against a numpy restatement on the oracle's Philox streams."""
    print("case: trajectory kernels C=3001 P=7", flush=True)
    rng = np.random.default_rng(44)
    C, P, beta = 3001, 7, 0.37
    seed, stage, step = 21, 5, 3
    phi0 = rng.normal(size=(C, P))
    A = rng.normal(size=(P, P)) * 0.1
    L = np.linalg.cholesky(A @ A.T + 0.05 * np.eye(P))
    prior_phi0, post_phi0 = rng.normal(size=C), rng.normal(size=C) * 3
    keep = []

    def D(a):
        t = torch.from_numpy(np.ascontiguousarray(a)).to(cuda)
        keep.append(t)
        return t

    phi_d, pp_d, po_d = D(phi0.copy()), D(prior_phi0.copy()), D(post_phi0.copy())
    xi_d, cand_d = torch.empty_like(phi_d), torch.empty_like(phi_d)
    index = torch.empty(2 * C, dtype=torch.int32, device=cuda)
    count = torch.zeros(1, dtype=torch.int32, device=cuda)
    counters = torch.zeros(3, dtype=torch.int64, device=cuda)
    cpost_s_d = torch.empty(C, dtype=torch.float64, device=cuda)
    phi, prior_phi, post_phi = phi0.copy(), prior_phi0.copy(), post_phi0.copy()
    moved = np.zeros(C, bool)
    n_acc = n_rej = n_unl = 0
    for sub in range(3):
        step2 = step * _lib.TRAJ_MAX_SUBSTEPS + sub
        xi = oracle.normals(seed, stage, step2, _lib.TAG_TRAJ_PROPOSAL, C, P)
        _lib.call("altar_da_trajectory_propose", _lib.ptr(phi_d), C, P, _lib.ptr(D(L)), seed, stage, step, sub,
                  _lib.ptr(xi_d), _lib.ptr(cand_d), None)
        # normals: device log/sincos against the host's (last bits); the
        # displacement from the same normals is bit for bit
        xi_dev = xi_d.cpu().numpy()
        assert np.max(np.abs(xi_dev - xi)) < 1e-12
        cand = oracle.mh_propose(phi, L, xi_dev)
        assert np.array_equal(cand_d.cpu().numpy(), cand)
        cprior, cdata_s = rng.normal(size=C), rng.normal(size=C) * 3
        mask = (rng.random(C) < 0.1).astype(np.int32)
        cpost_s = cprior + beta * cdata_s
        u = oracle.uniforms(seed, stage, step2, _lib.TAG_TRAJ_ACCEPT, C)
        acc = (mask == 0) & (np.log(u) <= cpost_s - post_phi)
        _lib.call("altar_da_trajectory_step", _lib.ptr(phi_d), _lib.ptr(pp_d), _lib.ptr(po_d), _lib.ptr(cand_d),
                  _lib.ptr(D(cprior)), _lib.ptr(D(cdata_s)), _lib.ptr(cpost_s_d), _lib.ptr(D(mask)), beta, C, P,
                  seed, stage, step, sub, _lib.ptr(index), _lib.ptr(counters), None)
        phi[acc], prior_phi[acc], post_phi[acc] = cand[acc], cprior[acc], cpost_s[acc]
        moved |= acc
        n_acc += acc.sum()
        n_rej += mask.sum()
        n_unl += ((mask == 0) & ~acc).sum()
        assert np.array_equal(cpost_s_d.cpu().numpy(), cpost_s)
        assert np.array_equal(phi_d.cpu().numpy(), phi)
        assert np.array_equal(pp_d.cpu().numpy(), prior_phi) and np.array_equal(po_d.cpu().numpy(), post_phi)
    _lib.call("altar_da_trajectory_list", _lib.ptr(index), C, _lib.ptr(count), None)
    n = int(count.item())
    assert np.array_equal(index[:n].cpu().numpy(), np.nonzero(moved)[0])
    assert counters.cpu().tolist() == [n_acc, n_rej, n_unl]


def test_device_pca_matches_host_routine(cuda):
    """device_pca (products on the device, FP32 MFMA) against whitened_pca
    (host FP64) on rank-20 fields plus noise and a mean offset: the same
    mean, singular values and subspace."""
    from altar_amd.surrogate import device_pca, whitened_pca
    print("case: device PCA 1024 rows x 700 columns (pitch 768), K=16", flush=True)
    rng = np.random.default_rng(12)
    M, N, ld, K = 1024, 700, 768, 16
    U = rng.normal(size=(M, 20))
    Vb = rng.normal(size=(20, N)) * np.linspace(3.0, 0.3, 20)[:, None]
    F = U @ Vb + 0.01 * rng.normal(size=(M, N)) + rng.normal(size=N)
    rows = np.sort(rng.choice(M + 256, M, replace=False))
    store = torch.zeros((M + 256, ld), dtype=torch.float32, device=cuda)
    store[torch.from_numpy(rows).to(cuda), :N] = torch.from_numpy(F.astype(np.float32)).to(cuda)
    mean_d, V_d, sv_d = device_pca(store, ld, N, rows, K, seed=4)
    mean_h, V_h, sv_h = whitened_pca(F.astype(np.float32), K, seed=4)
    print("mean err", np.max(np.abs(mean_d - mean_h)), "sv rel err", np.max(np.abs(sv_d / sv_h - 1)))
    assert np.max(np.abs(mean_d - mean_h)) < 1e-5 * np.max(np.abs(F))
    assert np.allclose(sv_d, sv_h, rtol=1e-4)
    cosines = np.linalg.svd(V_d.T @ V_h, compute_uv=False)
    assert cosines.min() > 1 - 1e-4
