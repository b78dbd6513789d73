"""Embedding-width golden fixtures that are regenerated, not committed.

An index at d = 1024 or 3072 is tens of MB of files, so these fixtures commit
only their spec (below), the SHA-256 of the files this generator writes, and
the reference's outputs on them (reference_outputs.npz, from
tests/golden/make_golden.py). Every value is exact on any machine:

* vectors: integers v in [-8128, 8128] (a component centre in [-60, 60] plus a
  sum of four uniform integers in [-8, 8], clipped to +-127, times a weight
  64 >> (k // 128), at least 1, per dimension k, so the variance decays
  along the dimensions), stored as v / 8192 (exact in float32);
* centroids: the component centres, weighted the same way;
* lists: each vector in the list of its nearest centroid, by exact int64
  squared distance (ties to the lower list), ids ascending inside a list;
* rotation (.dimc): a dense d x d matrix of integers in [-4095, 4095] / 2^14
  (a sum of three uniform integers), not orthogonal; the reference rotates with
  whatever the file holds, and its fp32 sums round.

Only numpy integer generation and integer arithmetic are used (no BLAS, no
floating-point library functions), so the files are bit-identical wherever
this image runs.
"""
from __future__ import annotations

import hashlib
import os

import numpy as np

# name: (n, nq, d, nlist, seed)
SPECS = {
    "w1024": (8192, 256, 1024, 32, 1024),
    "w3072": (4096, 128, 3072, 16, 3072),
}

FILES = ("meta", "centroids", "invlist", "norms", "dimc")


def _weights(d):
    k = np.arange(d)
    return np.maximum(64 >> np.minimum(k // 128, 6), 1).astype(np.int64)


def generate(name):
    """(vectors int64 n x d, queries int64 nq x d, centres int64 nlist x d,
    lists [ids per list], rotation int64 d x d): the integer forms."""
    n, nq, d, nlist, seed = SPECS[name]
    rng = np.random.default_rng(seed)
    w = _weights(d)
    centres = rng.integers(-60, 61, size=(nlist, d), dtype=np.int64)

    def draw(m):
        pick = rng.integers(0, nlist, size=m, dtype=np.int64)
        noise = rng.integers(-8, 9, size=(4, m, d), dtype=np.int64).sum(0)
        return np.clip(centres[pick] + noise, -127, 127) * w

    x = draw(n)
    q = draw(nq)
    c = centres * w
    # exact nearest centroid: |x|^2 - 2 x.c + |c|^2 in int64 (integer matmul, no BLAS)
    assign = np.empty(n, np.int64)
    cn = (c * c).sum(1)
    for s in range(0, n, 1024):
        xb = x[s:s + 1024]
        d2 = (xb * xb).sum(1)[:, None] - 2 * (xb @ c.T) + cn[None, :]
        assign[s:s + 1024] = d2.argmin(1)
    lists = [np.nonzero(assign == j)[0].astype(np.uint32) for j in range(nlist)]
    rot = rng.integers(-1365, 1366, size=(3, d, d), dtype=np.int64).sum(0)
    return x, q, c, lists, rot


def write(name, prefix):
    """Writes the fixture's index files under prefix and returns the queries
    (float32)."""
    import sys
    root = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    sys.path.insert(0, os.path.join(root, "katoann-public_amd"))
    from katoann import formats
    n, nq, d, nlist, seed = SPECS[name]
    x, q, c, lists, rot = generate(name)
    xf = (x.astype(np.float32) / np.float32(8192)).astype(np.float32)
    cf = (c.astype(np.float32) / np.float32(8192)).astype(np.float32)
    rf = (rot.astype(np.float32) / np.float32(16384)).astype(np.float32)
    formats.write_index(prefix, cf, lists, [xf[ids] for ids in lists], n, False, rf,
                        formats.default_prefix_dims(d))
    return (q.astype(np.float32) / np.float32(8192)).astype(np.float32)


def files_sha(prefix):
    h = hashlib.sha256()
    for ext in FILES:
        with open(prefix + "." + ext, "rb") as f:
            h.update(f.read())
    return h.hexdigest()
