"""Learned nprobe: an MLP reads each query's smallest sorted coarse distances
and predicts how many lists the query needs (no reference counterpart; the
reference searches with one nprobe per call, include/katoann/index.h:56-62).

    model = NprobeMLP(nfeat=64, hidden=256)
    data = make_training_set(ix, train_queries, list_of_id, gt_ids, nprobe_max)
    model.fit(data)                               # forward/backward on the GPU
    margin = calibrate(ix, model, val_queries, val_gt, target=0.99, ...)
    ix.search_device(..., nprobe=nprobe_max, model=model.struct(margin, nmin))

Every product runs in the library's own bf16 MFMA kernels (katoann_mlp_*),
with fp32 accumulation, activations, gradients and Adam state; torch only
holds the device buffers. The label of a training query is the smallest
nprobe whose first lists cover the lists of its exact top-10, and the network
regresses log2 of it.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import NprobeModel, _check, lib


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def gemm(M, N, K, A, sam, sak, B, sbn, sbk, bias=None, gate=None, ldg=0, relu=False, C32=None, C16=None, ldc=0):
    """C[m][n] = act(sum_k A(m,k) B(n,k) + bias[n]) on the GPU (bf16 operands, fp32 sums)."""
    _check(lib().katoann_mlp_gemm(M, N, K, _p(A), sam, sak, _p(B), sbn, sbk, _p(bias), _p(gate), ldg,
                                  int(relu), _p(C32), _p(C16), ldc, _stream()))


def to_bf16(x32, x16):
    _check(lib().katoann_mlp_to_bf16(_p(x32), _p(x16), x32.numel(), _stream()))


class NprobeMLP:
    """nfeat -> hidden (ReLU) -> hidden (ReLU) -> 1, on the GPU."""

    def __init__(self, nfeat: int = 64, hidden: int = 256, seed: int = 0, device="cuda"):
        import torch
        self.F, self.H, self.dev = nfeat, hidden, torch.device(device)
        g = torch.Generator(device="cpu").manual_seed(seed)
        F, H = nfeat, hidden
        w = {"W1": torch.randn(H, F, generator=g) * np.sqrt(2.0 / F), "b1": torch.zeros(H),
             "W2": torch.randn(H, H, generator=g) * np.sqrt(2.0 / H), "b2": torch.zeros(H),
             "W3": torch.randn(1, H, generator=g) * np.sqrt(1.0 / H), "b3": torch.zeros(1)}
        self.w = {k: v.float().to(self.dev).contiguous() for k, v in w.items()}
        self.w16 = {k: torch.empty(v.shape, dtype=torch.bfloat16, device=self.dev) for k, v in self.w.items()
                    if k.startswith("W")}
        for k in self.w16:
            to_bf16(self.w[k], self.w16[k])
        self.m = {k: torch.zeros_like(v) for k, v in self.w.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.w.items()}
        self.g = {k: torch.zeros_like(v) for k, v in self.w.items()}
        self.step = 0
        self.feat_mean = torch.zeros(F, device=self.dev)
        self.feat_inv_std = torch.ones(F, device=self.dev)
        self.y_mean, self.y_std = 0.0, 1.0
        self._bufs = {}

    def _buf(self, name, shape, dtype):
        import torch
        b = self._bufs.get(name)
        if b is None or b.numel() < int(np.prod(shape)) or b.dtype != dtype:
            b = torch.empty(int(np.prod(shape)), dtype=dtype, device=self.dev)
            self._bufs[name] = b
        return b[:int(np.prod(shape))].view(*shape)

    # ---- forward / backward (all products in the library's kernels)
    def forward(self, X16):
        """X16: B x F bf16 (standardised features). Returns z (B fp32) and the cache."""
        import torch
        B, F, H = X16.shape[0], self.F, self.H
        H1 = self._buf("H1", (B, H), torch.float32)
        H1b = self._buf("H1b", (B, H), torch.bfloat16)
        H2 = self._buf("H2", (B, H), torch.float32)
        H2b = self._buf("H2b", (B, H), torch.bfloat16)
        z = self._buf("z", (B, 1), torch.float32)
        gemm(B, H, F, X16, F, 1, self.w16["W1"], F, 1, bias=self.w["b1"], relu=True, C32=H1, C16=H1b, ldc=H)
        gemm(B, H, H, H1b, H, 1, self.w16["W2"], H, 1, bias=self.w["b2"], relu=True, C32=H2, C16=H2b, ldc=H)
        gemm(B, 1, H, H2b, H, 1, self.w16["W3"], H, 1, bias=self.w["b3"], C32=z, ldc=1)
        return z.view(B), dict(X16=X16, H1=H1, H1b=H1b, H2=H2, H2b=H2b)

    def backward(self, cache, gz):
        """Gradients into self.g given dL/dz (B fp32)."""
        import torch
        X16, H1, H1b, H2, H2b = (cache[k] for k in ("X16", "H1", "H1b", "H2", "H2b"))
        B, F, H = X16.shape[0], self.F, self.H
        gb = self._buf("gb", (B,), torch.bfloat16)
        to_bf16(gz, gb)
        G2 = self._buf("G2", (B, H), torch.float32)
        G2b = self._buf("G2b", (B, H), torch.bfloat16)
        G1 = self._buf("G1", (B, H), torch.float32)
        G1b = self._buf("G1b", (B, H), torch.bfloat16)
        s = _stream()
        # dW3 = g^T H2b, db3 = sum g
        gemm(1, H, B, gb, 0, 1, H2b, 1, H, C32=self.g["W3"], ldc=H)
        _check(lib().katoann_mlp_colsum(_p(gz), B, 1, _p(self.g["b3"]), s))
        # G2 = (g W3) gated by H2 > 0
        gemm(B, H, 1, gb, 1, 0, self.w16["W3"], 1, 0, gate=H2, ldg=H, C32=G2, C16=G2b, ldc=H)
        gemm(H, H, B, G2b, 1, H, H1b, 1, H, C32=self.g["W2"], ldc=H)
        _check(lib().katoann_mlp_colsum(_p(G2), B, H, _p(self.g["b2"]), s))
        # G1 = (G2b W2) gated by H1 > 0
        gemm(B, H, H, G2b, H, 1, self.w16["W2"], 1, H, gate=H1, ldg=H, C32=G1, C16=G1b, ldc=H)
        gemm(H, F, B, G1b, 1, H, X16, 1, F, C32=self.g["W1"], ldc=F)
        _check(lib().katoann_mlp_colsum(_p(G1), B, H, _p(self.g["b1"]), s))

    def loss_grad(self, z, y, tau=None):
        """Mean squared error, or with tau the pinball loss of the tau-quantile
        (device scalar), and dL/dz."""
        import torch
        gz = self._buf("gz", z.shape, torch.float32)
        loss = self._buf("loss", (1,), torch.float32)
        loss.zero_()
        if tau is None:
            _check(lib().katoann_mlp_mse(_p(z), _p(y), z.numel(), _p(gz), _p(loss), _stream()))
        else:
            _check(lib().katoann_mlp_pinball(_p(z), _p(y), z.numel(), tau, _p(gz), _p(loss), _stream()))
        return loss, gz

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.step += 1
        s = _stream()
        for k, w in self.w.items():
            _check(lib().katoann_mlp_adam(_p(w), _p(self.g[k]), _p(self.m[k]), _p(self.v[k]), w.numel(), lr, b1,
                                          b2, eps, self.step, _p(self.w16.get(k)), s))

    # ---- features and training
    def features(self, probe_dists, nprobe: int):
        """probe_dists: B x nprobe fp32 device tensor (ascending) -> B x F bf16 standardised features."""
        import torch
        B = probe_dists.shape[0]
        X16 = torch.empty((B, self.F), dtype=torch.bfloat16, device=self.dev)
        _check(lib().katoann_mlp_features(_p(probe_dists), B, nprobe, self.F, _p(self.feat_mean),
                                          _p(self.feat_inv_std), None, _p(X16), _stream()))
        return X16

    def set_scales(self, raw_feats, labels):
        """Standardisation from training data: raw_feats B x F (x_0 = d_0, x_i = d_i - d_0),
        labels: covering nprobe per query."""
        import torch
        self.feat_mean = raw_feats.mean(0).float().contiguous()
        self.feat_inv_std = (1.0 / raw_feats.std(0).clamp_min(1e-6)).float().contiguous()
        y = torch.log2(labels.double())
        self.y_mean, self.y_std = float(y.mean()), float(max(y.std().item(), 1e-3))

    def fit(self, X16, labels, epochs=30, batch=1024, lr=2e-3, seed=0, log=None, tau=None):
        """Adam on the standardised log2 label: mean squared error, or with tau
        the pinball loss (the network then predicts the tau-quantile)."""
        import torch
        y = ((torch.log2(labels.double()) - self.y_mean) / self.y_std).float().contiguous()
        n = X16.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.dev)
            tot = torch.zeros((), device=self.dev)
            nb = 0
            for s0 in range(0, n - batch + 1, batch):
                idx = perm[s0:s0 + batch]
                xb, yb = X16.index_select(0, idx).contiguous(), y.index_select(0, idx).contiguous()
                z, cache = self.forward(xb)
                loss, gz = self.loss_grad(z, yb, tau)
                self.backward(cache, gz)
                self.adam(lr=lr * (0.5 ** (ep // max(1, epochs // 3))))
                tot += loss[0]
                nb += 1
            hist.append(float(tot) / max(nb, 1))
            if log:
                log("nprobe mlp epoch %d loss %.4f" % (ep, hist[-1]))
        return hist

    def predict_log2(self, X16):
        z, _ = self.forward(X16)
        return z * self.y_std + self.y_mean

    def struct(self, margin: float, nmin: int = 1) -> NprobeModel:
        """The model as katoann_nprobe_model for Index.search_device(model=...)."""
        m = NprobeModel()
        m.nfeat, m.hidden = self.F, self.H
        m.feat_mean, m.feat_inv_std = self.feat_mean.data_ptr(), self.feat_inv_std.data_ptr()
        m.w1, m.b1 = self.w16["W1"].data_ptr(), self.w["b1"].data_ptr()
        m.w2, m.b2 = self.w16["W2"].data_ptr(), self.w["b2"].data_ptr()
        m.w3, m.b3 = self.w16["W3"].data_ptr(), self.w["b3"].data_ptr()
        m.y_mean, m.y_std, m.margin, m.nmin = self.y_mean, self.y_std, margin, nmin
        return m


MAX_LABEL = 4096


def coarse(ix, queries, nprobe: int, chunk: int = 4096):
    """Probe lists and sorted coarse distances (device tensors, nq x nprobe) of
    queries (device float32, not normalised here), by the index's own coarse
    pass."""
    import torch
    nq = queries.shape[0]
    probes = torch.empty((nq, nprobe), dtype=torch.int32, device=queries.device)
    pd = torch.empty((nq, nprobe), dtype=torch.float32, device=queries.device)
    for s in range(0, nq, chunk):
        e = min(nq, s + chunk)
        ix.search_device(queries[s:e].data_ptr(), e - s, 0, 0, k=1, nprobe=nprobe, survivors=1,
                         probes_ptr=probes[s:e].data_ptr(), probe_dists_ptr=pd[s:e].data_ptr(), coarse_only=True)
    return probes, pd


def raw_features(pd, F: int):
    """x_0 = d_0, x_i = d_i - d_0 (before standardisation), B x F."""
    x = pd[:, :F].clone()
    x[:, 1:] -= pd[:, :1]
    return x


def coverage_labels(probes, list_of_id, gt_ids, nlist=None):
    """Smallest nprobe covering the lists of each query's exact neighbours:
    probes (nq x m, the first m lists of the coarse order, m = nlist for the
    whole order), list_of_id (N), gt_ids (nq x k). A neighbour in a list past
    the first m counts as at rank m (the label is capped at m)."""
    import torch
    nq, m = probes.shape
    nlist = nlist or m
    pos = torch.full((nq, nlist), m - 1, dtype=probes.dtype, device=probes.device)
    pos.scatter_(1, probes.long(), torch.arange(m, dtype=probes.dtype, device=probes.device).expand(nq, m).contiguous())
    nn_lists = list_of_id[gt_ids.long()].long()
    return torch.gather(pos, 1, nn_lists).max(1).values + 1


def train(ix, train_queries, list_of_id, gt_ids, nfeat=64, hidden=256, epochs=30, seed=0, log=None, tau=None,
          lr=2e-3):
    """Build the training set with the index's coarse pass and fit a model."""
    import torch
    nlist = ix.nlist
    m = min(nlist, MAX_LABEL)  # the engine probes at most 4,096 lists per query
    probes, pd = coarse(ix, train_queries, m)
    labels = coverage_labels(probes, list_of_id, gt_ids, nlist)
    del probes
    model = NprobeMLP(nfeat, hidden, seed=seed, device=train_queries.device)
    model.set_scales(raw_features(pd, nfeat), labels)
    X16 = model.features(pd, m)
    del pd
    torch.cuda.synchronize()
    hist = model.fit(X16, labels, epochs=epochs, seed=seed, log=log, tau=tau, lr=lr)
    return model, labels, hist
