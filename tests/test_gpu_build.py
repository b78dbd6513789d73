"""The GPU index build (csrc/build.hip) against brute force and closed forms.

* assignment (katoann_assign_lists): on integer-valued data every fp32 sum is
  exact, so the nearest centroid must equal numpy's brute-force arg-min of the
  squared L2 distance exactly, ties (duplicated centroids) to the lowest index;
  dims that are and are not multiples of 4 and 32, centroid counts around the
  128-centroid blocks, ragged row counts. On float data: the fp64 distance of
  the chosen centroid is within the fp32 rounding bound of the best one.
* k-means (katoann_kmeans_device): one Lloyd step on integer data equals the
  numpy step exactly (fp64 sums of integers are exact): means of the assigned
  rows, a centroid with no rows unchanged, lists longer than one 1,024-row
  segment; the objective does not increase over iterations.
* rotation (katoann_pca_rotation_device): eigenvalues equal numpy's eigh of the
  fp64 sample covariance, the rows are orthonormal, diagonalise the covariance,
  match numpy's eigenvectors under the sign rule.
* the whole build: host and device entry points give the same files byte for
  byte; each vector sits in the list of its nearest centroid, ids ascending.
"""
import os

import numpy as np
import pytest

import katoann
from katoann import builder, formats

pytestmark = pytest.mark.gpu


def _cuda(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def gpu_assign(x, cen):
    import torch
    dx, dc = _cuda(x.astype(np.float32)), _cuda(cen.astype(np.float32))
    out = torch.empty(x.shape[0], dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    katoann.assign_lists(dx.data_ptr(), x.shape[0], x.shape[1], dc.data_ptr(), cen.shape[0], out.data_ptr())
    return out.cpu().numpy().astype(np.int64)


def brute_int(x, cen):
    x = x.astype(np.int64)
    c = cen.astype(np.int64)
    out = np.empty(x.shape[0], np.int64)
    for s in range(0, x.shape[0], 256):
        d = ((x[s:s + 256, None, :] - c[None]) ** 2).sum(-1)
        out[s:s + 256] = np.argmin(d, axis=1)  # first minimum: lowest index
    return out


@pytest.mark.parametrize("n,d,nlist", [(3000, 64, 37), (2000, 66, 300), (1500, 100, 129), (1001, 1024, 200),
                                       (700, 3072, 130), (129, 5, 1), (300, 32, 256)])
def test_assign_integer_data_exact(n, d, nlist):
    rng = np.random.default_rng(n + d + nlist)
    x = rng.integers(-8, 9, (n, d))
    cen = rng.integers(-8, 9, (nlist, d))
    if nlist > 20:  # exact ties: duplicated centroids, and rows equal to a centroid
        cen[nlist - 3] = cen[4]
        cen[nlist // 2] = cen[11]
        x[:50] = cen[rng.integers(0, nlist, 50)]
    got = gpu_assign(x, cen)
    want = brute_int(x, cen)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, (bad[:10], got[bad[:10]], want[bad[:10]])


def test_assign_float_data_within_rounding():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((20000, 128)).astype(np.float32)
    cen = rng.standard_normal((1000, 128)).astype(np.float32)
    got = gpu_assign(x, cen)
    xd, cd = x.astype(np.float64), cen.astype(np.float64)
    worst = 0.0
    n_eq = 0
    for s in range(0, x.shape[0], 1000):
        D = (xd[s:s + 1000] ** 2).sum(1)[:, None] + (cd ** 2).sum(1)[None] - 2 * xd[s:s + 1000] @ cd.T
        best = D.min(1)
        chosen = D[np.arange(D.shape[0]), got[s:s + 1000]]
        # fp32 error of |c|^2 - 2 x.c: a few ulps of the magnitudes summed
        bound = 1e-5 * ((xd[s:s + 1000] ** 2).sum(1) + (cd ** 2).sum(1).max())
        worst = max(worst, float(((chosen - best) / bound).max()))
        n_eq += int((got[s:s + 1000] == D.argmin(1)).sum())
    assert worst <= 1.0
    assert n_eq >= 0.999 * x.shape[0]


def lloyd_step_np(train, cen):
    a = brute_int(train, cen)
    new = cen.astype(np.float64).copy()
    for c in range(cen.shape[0]):
        rows = train[a == c]
        if len(rows):
            new[c] = rows.astype(np.float64).sum(0) / len(rows)
    return new.astype(np.float32), a


@pytest.mark.parametrize("n,d,nlist", [(5000, 64, 40), (5000, 100, 4), (3001, 1024, 17)])
def test_kmeans_step_exact(n, d, nlist):
    import torch
    rng = np.random.default_rng(n * 7 + d)
    train = rng.integers(-8, 9, (n, d))
    cen = rng.integers(-8, 9, (nlist, d)).astype(np.float32)
    cen[1] = 1000.0  # far from every row: keeps no row and must not move
    want, a = lloyd_step_np(train, cen)
    assert (a == 1).sum() == 0
    if nlist == 4:
        assert np.bincount(a, minlength=4).max() > 1024  # lists of several segments
    dt, dc = _cuda(train.astype(np.float32)), _cuda(cen)
    torch.cuda.synchronize()
    katoann.kmeans_device(dt.data_ptr(), n, d, dc.data_ptr(), nlist, 1)
    got = dc.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_kmeans_objective_decreases():
    import torch
    rng = np.random.default_rng(11)
    centres = rng.standard_normal((64, 48)) * 4
    train = (centres[rng.integers(0, 64, 20000)] + rng.standard_normal((20000, 48))).astype(np.float32)
    init = train[rng.choice(20000, 64, replace=False)].copy()
    dt = _cuda(train)
    prev = np.inf
    for iters in (0, 1, 2, 4, 8):
        dc = _cuda(init)
        torch.cuda.synchronize()
        katoann.kmeans_device(dt.data_ptr(), 20000, 48, dc.data_ptr(), 64, iters)
        cen = dc.cpu().numpy().astype(np.float64)
        D = (train.astype(np.float64) ** 2).sum(1)[:, None] + (cen ** 2).sum(1)[None] - 2 * train @ cen.T
        obj = float(D.min(1).sum())
        assert obj <= prev * (1 + 1e-6), (iters, obj, prev)
        prev = obj


def test_pca_rotation_vs_eigh():
    import torch
    rng = np.random.default_rng(5)
    d, m = 64, 5000
    q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    scale = np.linspace(3.0, 0.2, d)  # distinct variances
    x = ((rng.standard_normal((m, d)) * scale) @ q.T + 0.5).astype(np.float32)
    rot = torch.empty((d, d), dtype=torch.float32, device="cuda")
    dx = _cuda(x)
    torch.cuda.synchronize()
    rotated, eig = katoann.pca_rotation_device(dx.data_ptr(), m, d, rot.data_ptr())
    assert rotated
    R = rot.cpu().numpy().astype(np.float64)
    xc = x.astype(np.float64) - x.astype(np.float64).mean(0)
    cov = xc.T @ xc / m
    w, v = np.linalg.eigh(cov)
    assert np.allclose(eig, w[::-1], rtol=1e-9, atol=1e-12 * w.max())
    assert np.abs(R @ R.T - np.eye(d)).max() <= 1e-5
    off = R @ cov @ R.T
    assert np.abs(off - np.diag(np.diag(off))).max() <= 1e-5 * w.max()
    big = np.argmax(np.abs(R), axis=1)
    assert (R[np.arange(d), big] > 0).all()
    want = v[:, ::-1].T
    want = want * np.where(want[np.arange(d), np.argmax(np.abs(want), 1)] < 0, -1.0, 1.0)[:, None]
    assert np.abs(R - want).max() <= 1e-4


def test_pca_rotation_needs_two_rows():
    import torch
    x = _cuda(np.ones((1, 8), np.float32))
    rot = torch.zeros((8, 8), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    rotated, _ = katoann.pca_rotation_device(x.data_ptr(), 1, 8, rot.data_ptr())
    assert not rotated


def _files(prefix):
    return {ext: open(prefix + ext, "rb").read() for ext in (".meta", ".centroids", ".invlist", ".norms", ".dimc")}


@pytest.mark.parametrize("n,d,nlist", [(6000, 96, 50), (2500, 66, 300)])
def test_build_host_and_device_files_identical(tmp_path, n, d, nlist):
    import torch
    from katoann import datasets
    base, _ = datasets.decaying_clusters(n, 10, d=d, seed=n)
    p1, p2 = str(tmp_path / "host"), str(tmp_path / "dev")
    katoann.build(base, p1, nlist=nlist, train_size=4000, kmeans_iters=5, seed=9, normalized=True,
                  rotation_sample=3000)
    builder.build_torch(torch.from_numpy(base).cuda(), p2, nlist=nlist, train_size=4000, kmeans_iters=5, seed=9,
                        normalized=True, rotation_sample=3000)
    f1, f2 = _files(p1), _files(p2)
    for ext in f1:
        assert f1[ext] == f2[ext], ext
    raw = formats.read_index(p1)
    assert raw["rotation"] is not None and raw["prefix_dims"] == formats.default_prefix_dims(d)
    # every vector once, ids ascending inside each list, in the list of its nearest centroid
    ids = raw["ids"].astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(n))
    bounds = np.concatenate([[0], np.cumsum(raw["sizes"].astype(np.int64))])
    lists = np.repeat(np.arange(nlist), raw["sizes"].astype(np.int64))
    for c in range(nlist):
        assert np.all(np.diff(ids[bounds[c]:bounds[c + 1]]) > 0)
    a = gpu_assign(base, raw["centroids"])
    assert np.array_equal(a[ids], lists)


def test_build_seed_changes_partition(tmp_path):
    from katoann import datasets
    base, _ = datasets.decaying_clusters(4000, 10, d=32, seed=2)
    out = []
    for seed in (1, 1, 2):
        p = str(tmp_path / ("s%d_%d" % (seed, len(out))))
        katoann.build(base, p, nlist=20, seed=seed, kmeans_iters=3)
        out.append(open(p + ".centroids", "rb").read())
    assert out[0] == out[1] and out[0] != out[2]
