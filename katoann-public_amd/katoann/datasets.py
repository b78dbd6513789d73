"""Seeded synthetic workloads (no network, no downloads).

numpy generators are for CPU-sized data; the ``*_torch`` twins draw the same
laws on the GPU for million-scale workloads. Every generator is a pure
function of its seed on a given machine image.

Workloads:
  mixture(n, d, centres, sigma)   Gaussian mixture, centres ~ N(0, I), isotropic noise,
                                  L2-normalised (100,000 x 128, 256 centres, sigma 0.3).
  spectrum(n, d, alpha)           x = Q (lambda^1/2 * z), lambda_i = i^-alpha, Q a seeded
                                  random orthogonal basis (QR of a Gaussian, sign-fixed),
                                  L2-normalised (1,000,000 x 1,024 with alpha 1;
                                  990,000 x 3,072 with alpha 0.8).
  zipf_mixture(n, d, comps, s)    `comps` components with weights ~ j^-s, means and
                                  within-component noise both Q (lambda^1/2 * z) with
                                  lambda_i = i^-1, noise scale 0.5, L2-normalised
                                  (10,000,000 x 1,024, 16,384 components, s = 1.1).
  decaying_clusters(...)          the reference's end-to-end test law
                                  (tests/test_index.cpp:66-84): 64 centres 3 N(0,1),
                                  noise 0.5, dimension j scaled by exp(-j/40), normalised.
"""
from __future__ import annotations

import numpy as np


def _normalize_np(x: np.ndarray) -> np.ndarray:
    nrm = np.sqrt((x.astype(np.float64) ** 2).sum(1, keepdims=True))
    nrm[nrm == 0] = 1.0
    return (x / nrm).astype(np.float32)


def mixture(n: int, nq: int, d: int = 128, centres: int = 256, sigma: float = 0.3, seed: int = 1):
    rng = np.random.default_rng(seed)
    c = rng.standard_normal((centres, d))
    def draw(m):
        pick = rng.integers(0, centres, m)
        return _normalize_np(c[pick] + sigma * rng.standard_normal((m, d)))
    return draw(n), draw(nq)


def decaying_clusters(n: int, nq: int, d: int = 256, centres: int = 64, seed: int = 11):
    rng = np.random.default_rng(seed)
    c = 3.0 * rng.standard_normal((centres, d))
    scale = np.exp(-np.arange(d) / 40.0)
    def draw(m):
        pick = rng.integers(0, centres, m)
        return _normalize_np((c[pick] + 0.5 * rng.standard_normal((m, d))) * scale)
    return draw(n), draw(nq)


def random_basis_np(d: int, rng) -> np.ndarray:
    q, r = np.linalg.qr(rng.standard_normal((d, d)))
    return q * np.sign(np.diag(r))[None, :]


def spectrum(n: int, nq: int, d: int, alpha: float = 1.0, seed: int = 2):
    rng = np.random.default_rng(seed)
    q = random_basis_np(d, rng)
    lam = np.arange(1, d + 1, dtype=np.float64) ** (-alpha)
    def draw(m):
        return _normalize_np((rng.standard_normal((m, d)) * np.sqrt(lam)) @ q.T)
    return draw(n), draw(nq)


# ---------------------------------------------------------------- torch (GPU)

def _torch_gen(seed: int, device):
    import torch
    return torch.Generator(device=device).manual_seed(int(seed))


def _normalize_torch(x):
    import torch
    nrm = torch.sqrt((x.double() ** 2).sum(1, keepdim=True))
    nrm[nrm == 0] = 1.0
    return (x.double() / nrm).float()


def random_basis_torch(d: int, gen, device):
    import torch
    a = torch.randn((d, d), generator=gen, device=device, dtype=torch.float64)
    q, r = torch.linalg.qr(a)
    return q * torch.sign(torch.diagonal(r))[None, :]


def spectrum_torch(n: int, nq: int, d: int, alpha: float = 1.0, seed: int = 2, device="cuda",
                   chunk: int = 262144):
    """Returns (base, queries) float32 tensors on ``device``."""
    import torch
    g = _torch_gen(seed, device)
    q = random_basis_torch(d, g, device).float()
    lam = torch.arange(1, d + 1, dtype=torch.float64, device=device) ** (-alpha)
    sl = torch.sqrt(lam).float()
    def draw(m):
        out = torch.empty((m, d), dtype=torch.float32, device=device)
        for s in range(0, m, chunk):
            e = min(m, s + chunk)
            z = torch.randn((e - s, d), generator=g, device=device, dtype=torch.float32)
            out[s:e] = _normalize_torch((z * sl) @ q.T)
        return out
    return draw(n), draw(nq)


def spectrum_queries_torch(nq: int, d: int, alpha: float = 1.0, seed: int = 2, draw_seed: int = 101,
                           device="cuda", chunk: int = 262144):
    """More queries of spectrum_torch(..., seed)'s law (same basis and
    spectrum), drawn from their own seeded generator."""
    import torch
    q = random_basis_torch(d, _torch_gen(seed, device), device).float()
    g = _torch_gen(draw_seed, device)
    sl = torch.sqrt(torch.arange(1, d + 1, dtype=torch.float64, device=device) ** (-alpha)).float()
    out = torch.empty((nq, d), dtype=torch.float32, device=device)
    for s in range(0, nq, chunk):
        e = min(nq, s + chunk)
        z = torch.randn((e - s, d), generator=g, device=device, dtype=torch.float32)
        out[s:e] = _normalize_torch((z * sl) @ q.T)
    return out


def zipf_mixture_torch(n: int, nq: int, d: int = 1024, comps: int = 16384, s: float = 1.1,
                       noise: float = 0.5, seed: int = 3, device="cuda", chunk: int = 262144):
    """Component weights ~ j^-s; means and noise share the i^-1 spectrum in one basis."""
    import torch
    g = _torch_gen(seed, device)
    q = random_basis_torch(d, g, device).float()
    sl = torch.sqrt(torch.arange(1, d + 1, dtype=torch.float64, device=device) ** -1.0).float()
    means = (torch.randn((comps, d), generator=g, device=device) * sl) @ q.T
    w = torch.arange(1, comps + 1, dtype=torch.float64, device=device) ** (-s)
    w = w / w.sum()
    def draw(m):
        out = torch.empty((m, d), dtype=torch.float32, device=device)
        for st in range(0, m, chunk):
            e = min(m, st + chunk)
            pick = torch.multinomial(w, e - st, replacement=True, generator=g)
            z = torch.randn((e - st, d), generator=g, device=device, dtype=torch.float32)
            out[st:e] = _normalize_torch(means[pick] + noise * ((z * sl) @ q.T))
        return out
    return draw(n), draw(nq)


def zipf_queries_torch(nq: int, d: int = 1024, comps: int = 16384, s: float = 1.1, noise: float = 0.5,
                       seed: int = 3, draw_seed: int = 101, device="cuda", chunk: int = 262144):
    """More queries of zipf_mixture_torch(..., seed)'s law (same basis, means and
    weights), drawn from their own seeded generator."""
    import torch
    g = _torch_gen(seed, device)
    q = random_basis_torch(d, g, device).float()
    sl = torch.sqrt(torch.arange(1, d + 1, dtype=torch.float64, device=device) ** -1.0).float()
    means = (torch.randn((comps, d), generator=g, device=device) * sl) @ q.T
    w = torch.arange(1, comps + 1, dtype=torch.float64, device=device) ** (-s)
    w = w / w.sum()
    g2 = _torch_gen(draw_seed, device)
    out = torch.empty((nq, d), dtype=torch.float32, device=device)
    for st in range(0, nq, chunk):
        e = min(nq, st + chunk)
        pick = torch.multinomial(w, e - st, replacement=True, generator=g2)
        z = torch.randn((e - st, d), generator=g2, device=device, dtype=torch.float32)
        out[st:e] = _normalize_torch(means[pick] + noise * ((z * sl) @ q.T))
    return out


def ground_truth_np(base: np.ndarray, queries: np.ndarray, k: int = 10, block: int = 256) -> np.ndarray:
    """Exact L2 top-k in float64, ties by id."""
    b = base.astype(np.float64)
    bn = (b ** 2).sum(1)
    out = np.empty((queries.shape[0], k), np.uint32)
    for s in range(0, queries.shape[0], block):
        q = queries[s:s + block].astype(np.float64)
        dd = ((q ** 2).sum(1)[:, None] + bn[None, :] - 2.0 * (q @ b.T))
        part = np.argpartition(dd, k, axis=1)[:, :k + 1] if dd.shape[1] > k + 1 else np.argsort(dd, axis=1)
        for i in range(q.shape[0]):
            cand = part[i]
            vals = ((b[cand] - q[i]) ** 2).sum(1)
            o = np.lexsort((cand, vals))[:k]
            out[s + i] = cand[o]
    return out


def ground_truth_torch(base, queries, k: int = 10, cand: int = 64, qblock: int = 1024,
                       bblock: int = 1 << 20):
    """GPU brute force: fp32 GEMM shortlist of ``cand`` per query, re-ranked exactly in
    float64 (ties by id). Returns a uint32 numpy array (nq, k)."""
    import torch
    nq = queries.shape[0]
    out = np.empty((nq, k), np.uint32)
    bn = (base * base).sum(1)
    for s in range(0, nq, qblock):
        q = queries[s:s + qblock]
        best_d = None
        best_i = None
        for b0 in range(0, base.shape[0], bblock):
            bb = base[b0:b0 + bblock]
            dd = bn[b0:b0 + bblock][None, :] - 2.0 * (q @ bb.T)
            v, i = torch.topk(dd, min(cand, dd.shape[1]), dim=1, largest=False)
            i = i + b0
            if best_d is None:
                best_d, best_i = v, i
            else:
                v2 = torch.cat([best_d, v], 1)
                i2 = torch.cat([best_i, i], 1)
                v3, j = torch.topk(v2, cand, dim=1, largest=False)
                best_d, best_i = v3, torch.gather(i2, 1, j)
        ex = ((base[best_i].double() - q[:, None, :].double()) ** 2).sum(2)
        ex_np = ex.cpu().numpy()
        ids_np = best_i.cpu().numpy()
        for r in range(q.shape[0]):
            o = np.lexsort((ids_np[r], ex_np[r]))[:k]
            out[s + r] = ids_np[r][o]
    return out
