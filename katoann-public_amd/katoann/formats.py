"""Index files of a KatoANN index with prefix P (reference include/katoann/index.h:14-20).

    P.meta       5 x u32: dim, nlist, n, metric (0 = L2), normalized   (src/index.cpp:482-487)
    P.centroids  nlist x dim f32                                         (src/index.cpp:488-492)
    P.invlist    per list: u32 size, size x u32 ids, size x dim f32      (src/index.cpp:493-505)
    P.norms      per list: u32 size                                      (src/index.cpp:493-505)
    P.dimc       8 x u32 header (magic "DIMC", version 2, dim, nlist, n, prefix dims,
                 rotated, 0), then the dim x dim f32 rotation, row j = basis j
                                                                         (src/index.cpp:27-28, 535-545)
Also the .fbin / ground-truth formats of the CLI (tools/katoann.cpp:78-100).
"""
from __future__ import annotations

import os
import struct

import numpy as np

DIMC_MAGIC = 0x434D4944
DIMC_VERSION = 2


def default_prefix_dims(dim: int) -> int:
    """min(max(64, roundup(dim / 4, 64)), Dc) (src/index.cpp:512)."""
    dc = (dim + 63) // 64 * 64
    p = ((dim // 4) + 63) // 64 * 64
    return min(max(64, p), dc)


def write_index(prefix: str, centroids: np.ndarray, list_ids: list, list_vecs: list, n: int,
                normalized: bool, rotation: np.ndarray | None, prefix_dims: int | None = None) -> None:
    """Writes the five files; list_ids[c] (u32) and list_vecs[c] (size x dim f32) per list."""
    centroids = np.ascontiguousarray(centroids, dtype=np.float32)
    nlist, dim = centroids.shape
    os.makedirs(os.path.dirname(os.path.abspath(prefix)), exist_ok=True)
    with open(prefix + ".meta", "wb") as f:
        f.write(struct.pack("<5I", dim, nlist, n, 0, 1 if normalized else 0))
    with open(prefix + ".centroids", "wb") as f:
        f.write(centroids.tobytes())
    with open(prefix + ".invlist", "wb") as f, open(prefix + ".norms", "wb") as g:
        for c in range(nlist):
            ids = np.ascontiguousarray(list_ids[c], dtype=np.uint32)
            vecs = np.ascontiguousarray(list_vecs[c], dtype=np.float32).reshape(len(ids), dim)
            f.write(struct.pack("<I", len(ids)))
            f.write(ids.tobytes())
            f.write(vecs.tobytes())
            g.write(struct.pack("<I", len(ids)))
    pd = default_prefix_dims(dim) if prefix_dims is None else prefix_dims
    with open(prefix + ".dimc", "wb") as f:
        f.write(struct.pack("<8I", DIMC_MAGIC, DIMC_VERSION, dim, nlist, n, pd,
                            1 if rotation is not None else 0, 0))
        if rotation is not None:
            f.write(np.ascontiguousarray(rotation, dtype=np.float32).tobytes())


def read_index(prefix: str):
    """Returns dict(dim, nlist, n, normalized, centroids, sizes, ids, vecs, rotation, prefix_dims)
    with ids / vecs concatenated in list order."""
    with open(prefix + ".meta", "rb") as f:
        meta = f.read()
    dim, nlist, n = struct.unpack("<3I", meta[:12])
    normalized = len(meta) >= 20 and struct.unpack("<I", meta[16:20])[0] != 0
    cen = np.fromfile(prefix + ".centroids", dtype=np.float32, count=nlist * dim).reshape(nlist, dim)
    sizes = np.zeros(nlist, np.uint32)
    ids, vecs = [], []
    with open(prefix + ".invlist", "rb") as f:
        for c in range(nlist):
            (sz,) = struct.unpack("<I", f.read(4))
            sizes[c] = sz
            ids.append(np.frombuffer(f.read(4 * sz), dtype=np.uint32))
            vecs.append(np.frombuffer(f.read(4 * sz * dim), dtype=np.float32).reshape(sz, dim))
    rot, pd = None, 0
    if os.path.exists(prefix + ".dimc"):
        with open(prefix + ".dimc", "rb") as f:
            hdr = struct.unpack("<8I", f.read(32))
            pd = hdr[5]
            if hdr[6] == 1:
                rot = np.frombuffer(f.read(4 * dim * dim), dtype=np.float32).reshape(dim, dim)
    return dict(dim=dim, nlist=nlist, n=n, normalized=normalized, centroids=cen, sizes=sizes,
                ids=np.concatenate(ids) if ids else np.zeros(0, np.uint32),
                vecs=np.concatenate(vecs) if vecs else np.zeros((0, dim), np.float32),
                rotation=rot, prefix_dims=pd)


def write_fbin(path: str, x: np.ndarray) -> None:
    x = np.ascontiguousarray(x, dtype=np.float32)
    with open(path, "wb") as f:
        f.write(struct.pack("<2I", x.shape[0], x.shape[1]))
        f.write(x.tobytes())


def read_fbin(path: str) -> np.ndarray:
    with open(path, "rb") as f:
        n, d = struct.unpack("<2I", f.read(8))
        return np.frombuffer(f.read(4 * n * d), dtype=np.float32).reshape(n, d).copy()


def write_gt(path: str, gt: np.ndarray) -> None:
    gt = np.ascontiguousarray(gt, dtype=np.uint32)
    with open(path, "wb") as f:
        f.write(struct.pack("<2I", gt.shape[0], gt.shape[1]))
        f.write(gt.tobytes())
