"""Index builders that write the reference's file format.

The reference builds with k-means++ seeding, Elkan k-means on a strided training
sample, nearest-centre assignment and a Jacobi PCA rotation (src/index.cpp:419-533,
src/kmeans.cpp). The search path only reads the files, so parity is judged on the
same files whichever builder wrote them. These builders keep the same structure
(strided sample, seeded initial centres, Lloyd iterations, full assignment with ids
in ascending order inside each list, PCA basis sorted by decreasing variance) and
are written for speed on their hardware:

* ``build_numpy``: CPU, for test-sized data.
* ``build_torch``: on the GPU with torch matrix products, for the million-scale
  synthetic workloads (the k-means assignment is a GEMM + argmin; the PCA is
  ``torch.linalg.eigh`` on the sample covariance).
"""
from __future__ import annotations

import numpy as np

from . import formats


def _strided_sample(n: int, ts: int) -> np.ndarray:
    """Rows i * n / ts (src/index.cpp:433-443)."""
    if ts >= n:
        return np.arange(n)
    stride = (n << 32) // ts
    return np.minimum((np.arange(ts, dtype=np.uint64) * np.uint64(stride)) >> np.uint64(32), n - 1).astype(np.int64)


def _pca_rotation_np(sample: np.ndarray) -> np.ndarray:
    x = sample.astype(np.float64)
    x -= x.mean(axis=0, keepdims=True)
    cov = (x.T @ x) / x.shape[0]
    w, v = np.linalg.eigh(cov)
    order = np.argsort(-w, kind="stable")
    return np.ascontiguousarray(v[:, order].T.astype(np.float32))


def _assign_np(data: np.ndarray, cen: np.ndarray, block: int = 16384) -> np.ndarray:
    out = np.empty(data.shape[0], np.int64)
    cn = (cen.astype(np.float64) ** 2).sum(1)
    for s in range(0, data.shape[0], block):
        x = data[s:s + block].astype(np.float64)
        d = cn[None, :] - 2.0 * (x @ cen.T.astype(np.float64))
        out[s:s + block] = np.argmin(d, axis=1)
    return out


def kmeans_np(train: np.ndarray, nlist: int, iters: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    cen = train[rng.choice(train.shape[0], nlist, replace=False)].astype(np.float64)
    for _ in range(iters):
        a = _assign_np(train, cen)
        sums = np.zeros_like(cen)
        np.add.at(sums, a, train.astype(np.float64))
        cnt = np.bincount(a, minlength=nlist).astype(np.float64)
        nz = cnt > 0
        cen[nz] = sums[nz] / cnt[nz, None]
    return cen.astype(np.float32)


def _lists(assign: np.ndarray, nlist: int):
    order = np.argsort(assign, kind="stable")
    counts = np.bincount(assign, minlength=nlist)
    bounds = np.concatenate([[0], np.cumsum(counts)])
    return [order[bounds[c]:bounds[c + 1]].astype(np.uint32) for c in range(nlist)]


def build_numpy(data: np.ndarray, prefix: str, nlist: int = 2048, train_size: int = 256000,
                kmeans_iters: int = 12, seed: int = 0, normalized: bool = False,
                rotation_sample: int = 100000) -> None:
    data = np.ascontiguousarray(data, dtype=np.float32)
    n, dim = data.shape
    if dim == 0 or dim > 4096:
        raise ValueError("dimension must be 1 to 4096")
    if n == 0 or nlist == 0 or nlist > n:
        raise ValueError("need at least nlist vectors")
    train = data[_strided_sample(n, min(train_size, n))]
    cen = kmeans_np(train, nlist, kmeans_iters, seed)
    lists = _lists(_assign_np(data, cen), nlist)
    target = min(max(rotation_sample, 1024), n)
    stride = max(1, n // target)
    samp = data[np.concatenate(lists)[::stride][:target]]
    rot = _pca_rotation_np(samp) if samp.shape[0] >= 2 and dim >= 2 else None
    formats.write_index(prefix, cen, lists, [data[ids] for ids in lists], n, normalized, rot)


def build_torch(data, prefix: str | None, nlist: int, train_size: int = 256000,
                kmeans_iters: int = 12, seed: int = 0, normalized: bool = False,
                rotation_sample: int = 100000, block: int = 65536):
    """GPU builder. ``data`` is an (n, dim) float32 torch tensor on the GPU.

    Returns dict(centroids, order, sizes, rotation) as torch tensors (order: row
    ids grouped by list, ascending inside each list); writes files when
    ``prefix`` is given."""
    import torch

    n, dim = data.shape
    dev = data.device
    g = torch.Generator(device="cpu").manual_seed(int(seed))
    tidx = torch.from_numpy(_strided_sample(n, min(train_size, n))).to(dev)
    train = data.index_select(0, tidx)
    init = torch.randperm(train.shape[0], generator=g)[:nlist].to(dev)
    cen = train.index_select(0, init).clone()

    def assign(x):
        out = torch.empty(x.shape[0], dtype=torch.int64, device=dev)
        for s in range(0, x.shape[0], block):
            xb = x[s:s + block]
            cn = (cen * cen).sum(1)
            d = cn[None, :] - 2.0 * (xb @ cen.T)
            out[s:s + block] = torch.argmin(d, dim=1)
        return out

    for _ in range(kmeans_iters):
        a = assign(train)
        sums = torch.zeros((nlist, dim), dtype=torch.float64, device=dev)
        sums.index_add_(0, a, train.double())
        cnt = torch.bincount(a, minlength=nlist).double()
        nz = cnt > 0
        cen[nz] = (sums[nz] / cnt[nz, None]).float()
    a = assign(data)
    order = torch.argsort(a, stable=True)
    sizes = torch.bincount(a, minlength=nlist)
    target = min(max(rotation_sample, 1024), n)
    stride = max(1, n // target)
    samp = data.index_select(0, order[::stride][:target]).double()
    samp = samp - samp.mean(0, keepdim=True)
    cov = (samp.T @ samp) / samp.shape[0]
    w, v = torch.linalg.eigh(cov)
    idx = torch.argsort(-w, stable=True)
    rot = v[:, idx].T.contiguous().float()
    out = dict(centroids=cen.contiguous(), order=order, sizes=sizes, rotation=rot)
    if prefix is not None:
        sz = sizes.cpu().numpy()
        bounds = np.concatenate([[0], np.cumsum(sz)])
        o = order.cpu().numpy().astype(np.uint32)
        host = data.cpu().numpy()
        lists = [o[bounds[c]:bounds[c + 1]] for c in range(nlist)]
        formats.write_index(prefix, cen.cpu().numpy(), lists, [host[ids] for ids in lists], n,
                            normalized, rot.cpu().numpy())
    return out


def load_torch(data, nlist: int, train_size: int = 256000, kmeans_iters: int = 10, seed: int = 1,
               normalized: bool = True):
    """Builds an index from a GPU tensor with build_torch and loads it from
    device memory (no files). Returns (Index, info) where info holds the
    build dict plus ``sizes`` (numpy) and ``list_of_id`` (device int32: the
    list of every base id)."""
    import torch
    from . import Index, formats
    b = build_torch(data, None, nlist=nlist, train_size=train_size, kmeans_iters=kmeans_iters, seed=seed,
                    normalized=normalized)
    order = b["order"]
    sizes = b["sizes"].cpu().numpy().astype(np.uint32)
    rows = data.index_select(0, order).contiguous()
    ix = Index()
    ix.load_arrays(data.shape[1], nlist, data.shape[0], normalized, b["centroids"].cpu().numpy(), sizes,
                   order.cpu().numpy().astype(np.uint32), None, b["rotation"].cpu().numpy(),
                   formats.default_prefix_dims(data.shape[1]), vectors_device_ptr=rows.data_ptr())
    torch.cuda.synchronize()
    del rows
    list_of_id = torch.empty(data.shape[0], dtype=torch.int32, device=data.device)
    list_of_id[order] = torch.repeat_interleave(torch.arange(nlist, dtype=torch.int32, device=data.device),
                                                b["sizes"].to(data.device).long())
    info = dict(b, sizes=sizes, list_of_id=list_of_id)
    return ix, info
