"""Learned nprobe on the GPU: the bf16 MFMA product, the MLP's forward and
backward against the CPU oracle (oracle/mlp_oracle.py), Adam, training, and
the learned search (its per-query nprobe reproduced exactly by the explicit
form, and the work it saves at equal recall)."""
import numpy as np
import pytest

import mlp_oracle as M

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _bf(x):
    t = _torch()
    return t.from_numpy(np.ascontiguousarray(x, np.float32)).cuda().to(t.bfloat16)


@pytest.mark.parametrize("M_,N_,K_,trans", [(64, 64, 32, False), (100, 70, 45, False), (1, 256, 1000, True),
                                            (513, 1, 256, False), (256, 64, 1024, True), (37, 300, 1, False)])
def test_gemm_matches_oracle(M_, N_, K_, trans):
    """C = act(A B^T + bias) gated, with row-major and transposed (strided) operands."""
    t = _torch()
    from katoann import learned
    print("gemm", M_, N_, K_, trans, flush=True)
    rng = np.random.default_rng(M_ * 7 + N_)
    A = rng.standard_normal((M_, K_)).astype(np.float32)
    Bm = rng.standard_normal((N_, K_)).astype(np.float32)
    bias = rng.standard_normal(N_).astype(np.float32)
    gate = (rng.random((M_, N_)) > 0.3).astype(np.float32)
    if trans:  # operands stored k-major: A(m,k) = At[k][m]
        At, Bt = _bf(A.T.copy()), _bf(Bm.T.copy())
        sa, sb = (1, M_), (1, N_)
    else:
        At, Bt = _bf(A), _bf(Bm)
        sa, sb = (K_, 1), (K_, 1)
    for relu, use_gate in ((False, False), (True, True)):
        C32 = t.full((M_, N_), np.nan, device="cuda")
        C16 = t.empty((M_, N_), dtype=t.bfloat16, device="cuda")
        g = t.from_numpy(gate).cuda() if use_gate else None
        learned.gemm(M_, N_, K_, At, sa[0], sa[1], Bt, sb[0], sb[1], bias=t.from_numpy(bias).cuda(), gate=g,
                     ldg=N_, relu=relu, C32=C32, C16=C16, ldc=N_)
        t.cuda.synchronize()
        ref = M.bf16(A) @ M.bf16(Bm).T + bias
        if relu:
            ref = np.maximum(ref, 0.0)
        if use_gate:
            ref = ref * (gate > 0)
        got = C32.cpu().numpy().astype(np.float64)
        scale = np.abs(M.bf16(A)) @ np.abs(M.bf16(Bm)).T + np.abs(bias)
        assert np.all(np.abs(got - ref) <= 1e-5 * scale + 1e-6)
        assert np.array_equal(C16.float().cpu().numpy(), M.bf16(got).astype(np.float32))


def _model_from(p, F, H):
    t = _torch()
    from katoann import learned
    m = learned.NprobeMLP(F, H)
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        m.w[k].copy_(t.from_numpy(np.asarray(p[k], np.float32).reshape(m.w[k].shape)))
        if k in m.w16:
            learned.to_bf16(m.w[k], m.w16[k])
    return m


def test_forward_backward_match_oracle():
    t = _torch()
    F, H, B = 64, 256, 512
    rng = np.random.default_rng(5)
    p = M.init(F, H, 5)
    for k in ("b1", "b2", "b3"):
        p[k] = rng.standard_normal(p[k].shape) * 0.1
    X = rng.standard_normal((B, F)).astype(np.float32)
    y = rng.standard_normal(B).astype(np.float32)
    m = _model_from(p, F, H)
    z, cache = m.forward(_bf(X))
    loss, gz = m.loss_grad(z, t.from_numpy(y).cuda())
    m.backward(cache, gz)
    t.cuda.synchronize()
    zo, co = M.forward(p, X, bf=True)
    lo, go = M.loss_grad(zo, y.astype(np.float64), bf=True)
    grads = M.backward(co, go, bf=True)
    zg = z.cpu().numpy()
    assert np.max(np.abs(zg - zo)) <= 1e-3 * np.max(np.abs(zo))
    assert abs(float(loss.item()) - lo) <= 1e-3 * lo
    for k in ("W1", "b1", "W2", "b2", "W3", "b3"):
        gg = m.g[k].cpu().numpy().astype(np.float64).reshape(grads[k].shape)
        err = np.linalg.norm(gg - grads[k]) / max(np.linalg.norm(grads[k]), 1e-12)
        print(k, err)
        assert err < 2e-3, k


@pytest.mark.parametrize("tau", [None, 0.9])
def test_loss_kernels_match_oracle(tau):
    t = _torch()
    from katoann import learned
    rng = np.random.default_rng(9)
    n = 3000
    z = rng.standard_normal(n).astype(np.float32)
    y = rng.standard_normal(n).astype(np.float32)
    y[:5] = z[:5]  # u = 0 takes the u >= 0 side
    m = learned.NprobeMLP(8, 16)
    loss, g = m.loss_grad(t.from_numpy(z).cuda(), t.from_numpy(y).cuda(), tau=tau)
    t.cuda.synchronize()
    lo, go = M.loss_grad(z.astype(np.float64), y.astype(np.float64), bf=True, tau=tau)
    assert abs(float(loss.item()) - lo) <= 1e-5 * max(abs(lo), 1e-6)
    assert np.allclose(g.cpu().numpy(), go, rtol=1e-6, atol=0)


def test_adam_matches_numpy():
    t = _torch()
    from katoann import learned
    m = learned.NprobeMLP(8, 16, seed=2)
    w0 = {k: v.cpu().numpy().astype(np.float64) for k, v in m.w.items()}
    rng = np.random.default_rng(1)
    gs = [{k: rng.standard_normal(v.shape) for k, v in w0.items()} for _ in range(3)]
    mm = {k: np.zeros_like(v) for k, v in w0.items()}
    vv = {k: np.zeros_like(v) for k, v in w0.items()}
    w = {k: v.copy() for k, v in w0.items()}
    for step, g in enumerate(gs, 1):
        for k in m.w:
            m.g[k].copy_(t.from_numpy(g[k].astype(np.float32)))
        m.adam(lr=1e-2)
        for k in w:
            gk = g[k].astype(np.float32).astype(np.float64)
            mm[k] = 0.9 * mm[k] + 0.1 * gk
            vv[k] = 0.999 * vv[k] + 0.001 * gk * gk
            w[k] -= 1e-2 * (mm[k] / (1 - 0.9 ** step)) / (np.sqrt(vv[k] / (1 - 0.999 ** step)) + 1e-8)
    t.cuda.synchronize()
    for k in w:
        assert np.allclose(m.w[k].cpu().numpy(), w[k], rtol=1e-4, atol=1e-6), k
        if k in m.w16:
            assert np.array_equal(m.w16[k].float().cpu().numpy(), M.bf16(m.w[k].cpu().numpy()).astype(np.float32))


def test_training_reduces_loss():
    t = _torch()
    from katoann import learned
    rng = np.random.default_rng(0)
    X = rng.standard_normal((8192, 64)).astype(np.float32)
    y = (np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2]).astype(np.float32)
    m = learned.NprobeMLP(64, 256, seed=0)
    m.y_mean, m.y_std = 0.0, 1.0
    labels = t.from_numpy(np.exp2(y).astype(np.float32)).cuda()  # fit() regresses log2(label)
    hist = m.fit(_bf(X), labels, epochs=6, batch=512, lr=2e-3)
    print(hist)
    assert hist[-1] < 0.25 * hist[0]


@pytest.fixture(scope="module")
def spectrum_index():
    """100k x 128 i^-1-spectrum vectors, nlist 512, built and loaded on the GPU."""
    t = _torch()
    from katoann import builder, datasets
    base, q = datasets.spectrum_torch(100000, 2000, 128, alpha=1.0, seed=7, device="cuda")
    ix, info = builder.load_torch(base, nlist=512, train_size=100000, kmeans_iters=8, seed=1)
    train_q = datasets.spectrum_queries_torch(24000, 128, alpha=1.0, seed=7, draw_seed=11, device="cuda")
    gt = datasets.ground_truth_torch(base, q, 10)
    gt_train = datasets.ground_truth_torch(base, train_q, 10)
    return ix, info, q, gt, train_q, t.from_numpy(gt_train.astype(np.int64)).cuda()


def _search(ix, q, k, np_, S, model=None):
    t = _torch()
    nq = q.shape[0]
    ids = t.empty((nq, k), dtype=t.int32, device="cuda")
    dists = t.empty((nq, k), dtype=t.float32, device="cuda")
    used = t.empty(nq, dtype=t.int32, device="cuda")
    ix.search_device(q.data_ptr(), nq, ids.data_ptr(), dists.data_ptr(), k=k, nprobe=np_, survivors=S,
                     model=model, npq_out_ptr=used.data_ptr())
    t.cuda.synchronize()
    return ids.cpu().numpy().view(np.uint32), dists.cpu().numpy(), used.cpu().numpy().astype(np.int64)


def _recall(ids, gt):
    return np.mean([len(set(ids[i]) & set(gt[i])) / 10.0 for i in range(ids.shape[0])])


def test_learned_search_gain(spectrum_index):
    """Trained on the GPU, the learned per-query nprobe reaches the recall of a
    high fixed nprobe (about 0.99 recall@10) with far fewer probed lists; its
    result equals the explicit per-query form."""
    t = _torch()
    from katoann import learned
    ix, info, q, gt, train_q, gt_train = spectrum_index
    model, labels, hist = learned.train(ix, train_q, info["list_of_id"], gt_train, epochs=12, seed=0, log=print)
    assert hist[-1] < hist[0]
    S, nmax, np_fixed = 64, 512, 192
    val, test = q[:1000].contiguous(), q[1000:].contiguous()
    gt_val, gt_test = gt[:1000], gt[1000:]
    # the recall of fixed nprobe 192 (about 0.99) on each half
    ids, _, _ = _search(ix, val, 10, np_fixed, S)
    target_val = _recall(ids, gt_val)
    ids, _, used_f = _search(ix, test, 10, np_fixed, S)
    rec_f, work_f = _recall(ids, gt_test), used_f.sum()
    # learned: the smallest margin reaching that recall on the validation half
    margin = None
    for mg in np.arange(-0.5, 4.01, 0.125):
        ids, _, used = _search(ix, val, 10, nmax, S, model.struct(float(mg), nmin=1))
        if _recall(ids, gt_val) >= target_val:
            margin = float(mg)
            break
    assert margin is not None
    ids_l, dists_l, used_l = _search(ix, test, 10, nmax, S, model.struct(margin, nmin=1))
    rec_l, work_l = _recall(ids_l, gt_test), used_l.sum()
    print("fixed nprobe %d: recall %.4f probes %d | learned margin %.3f: recall %.4f probes %d" % (
        np_fixed, rec_f, work_f, margin, rec_l, work_l))
    assert rec_l >= rec_f - 0.005
    assert work_l < 0.75 * work_f
    # the learned search equals the explicit per-query nprobe search
    ids_e, dists_e, _ = ix.search_ex(test.cpu().numpy(), k=10, nprobe=nmax, survivors=S, normalize=False,
                                     nprobe_per_query=used_l.astype(np.uint32))
    assert np.array_equal(ids_e, ids_l) and np.array_equal(dists_e.view(np.uint32), dists_l.view(np.uint32))
