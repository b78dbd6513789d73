"""Index builders that write the reference's file format.

The search path only reads the files (or the arrays of ``load_arrays``), so
parity is judged on the same files whichever builder wrote them.

* ``build_torch`` / ``load_torch``: the library's GPU build (katoann_build_device,
  csrc/build.hip) on vectors already in device memory: uniform training sample,
  distinct training rows as initial centres, Lloyd iterations with the fp32 MFMA
  assignment kernel and fp64 means, every vector assigned, ids ascending inside
  each list, PCA rotation from rocSOLVER's eigensolver.
* ``build_numpy``: a CPU builder of the same method for test-sized data where no
  GPU is present (the CPU tests and the golden fixtures).
"""
from __future__ import annotations

import numpy as np

from . import formats


def _sample(n: int, k: int, rng) -> np.ndarray:
    """k distinct rows of n, uniform, ascending."""
    if k >= n:
        return np.arange(n)
    return np.sort(rng.choice(n, k, replace=False))


def _pca_rotation_np(sample: np.ndarray) -> np.ndarray:
    x = sample.astype(np.float64)
    x -= x.mean(axis=0, keepdims=True)
    cov = (x.T @ x) / x.shape[0]
    w, v = np.linalg.eigh(cov)
    rot = v[:, ::-1].T.copy()  # rows by decreasing eigenvalue
    big = np.argmax(np.abs(rot), axis=1)
    rot *= np.where(rot[np.arange(rot.shape[0]), big] < 0, -1.0, 1.0)[:, None]
    return np.ascontiguousarray(rot.astype(np.float32))


def _assign_np(data: np.ndarray, cen: np.ndarray, block: int = 16384) -> np.ndarray:
    out = np.empty(data.shape[0], np.int64)
    cn = (cen.astype(np.float64) ** 2).sum(1)
    for s in range(0, data.shape[0], block):
        x = data[s:s + block].astype(np.float64)
        d = cn[None, :] - 2.0 * (x @ cen.T.astype(np.float64))
        out[s:s + block] = np.argmin(d, axis=1)
    return out


def kmeans_np(train: np.ndarray, nlist: int, iters: int, rng) -> np.ndarray:
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
    rng = np.random.default_rng(seed)
    train = data[_sample(n, max(min(train_size, n), nlist), rng)]
    cen = kmeans_np(train, nlist, kmeans_iters, rng)
    lists = _lists(_assign_np(data, cen), nlist)
    samp = data[_sample(n, min(max(rotation_sample, 1024), n), rng)]
    rot = _pca_rotation_np(samp) if samp.shape[0] >= 2 and dim >= 2 else None
    formats.write_index(prefix, cen, lists, [data[ids] for ids in lists], n, normalized, rot)


def build_torch(data, prefix: str | None, nlist: int, train_size: int = 256000,
                kmeans_iters: int = 12, seed: int = 0, normalized: bool = False,
                rotation_sample: int = 100000):
    """The GPU build (katoann_build_device) of an (n, dim) float32 CUDA tensor.

    Returns dict(centroids, order, sizes, rotation) as CUDA tensors (order: row
    ids grouped by list, ascending inside each list; rotation None when not
    computed); writes the index files when ``prefix`` is given."""
    import torch
    from . import build_device

    data = data.contiguous()
    n, dim = data.shape
    dev = data.device
    cen = torch.empty((nlist, dim), dtype=torch.float32, device=dev)
    order = torch.empty(n, dtype=torch.int32, device=dev)
    sizes = torch.empty(nlist, dtype=torch.int32, device=dev)
    rot = torch.empty((dim, dim), dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    rotated = build_device(data.data_ptr(), n, dim, cen.data_ptr(), order.data_ptr(), sizes.data_ptr(),
                           rot.data_ptr(), nlist, train_size=train_size, kmeans_iters=kmeans_iters, seed=seed,
                           rotation_sample=rotation_sample)
    order = order.long()
    out = dict(centroids=cen, order=order, sizes=sizes.long(), rotation=rot if rotated else None)
    if prefix is not None:
        sz = out["sizes"].cpu().numpy()
        bounds = np.concatenate([[0], np.cumsum(sz)])
        o = order.cpu().numpy().astype(np.uint32)
        host = data.cpu().numpy()
        lists = [o[bounds[c]:bounds[c + 1]] for c in range(nlist)]
        formats.write_index(prefix, cen.cpu().numpy(), lists, [host[ids] for ids in lists], n,
                            normalized, rot.cpu().numpy() if rotated else None)
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
    rot = None if b["rotation"] is None else b["rotation"].cpu().numpy()
    ix.load_arrays(data.shape[1], nlist, data.shape[0], normalized, b["centroids"].cpu().numpy(), sizes,
                   order.cpu().numpy().astype(np.uint32), None, rot,
                   formats.default_prefix_dims(data.shape[1]), vectors_device_ptr=rows.data_ptr())
    torch.cuda.synchronize()
    del rows
    list_of_id = torch.empty(data.shape[0], dtype=torch.int32, device=data.device)
    list_of_id[order] = torch.repeat_interleave(torch.arange(nlist, dtype=torch.int32, device=data.device),
                                                b["sizes"].to(data.device).long())
    info = dict(b, sizes=sizes, list_of_id=list_of_id)
    return ix, info
