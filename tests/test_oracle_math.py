"""The oracle's arithmetic pinned by independent restatements: closed forms,
brute force in numpy, and the reference's kernel test formula
(tests/test_kernels.cpp:25-108)."""
import heapq

import numpy as np
import pytest

import oracle as O
from katoann import datasets, formats


def lround_np(t):
    """round half away from zero on float64, written without libm."""
    a = np.abs(t)
    f = np.floor(a)
    r = f + ((a - f) >= 0.5)
    return np.sign(t) * r


def draw_bytes(rng, shape):
    """0 w.p. 1/4, 255 w.p. 1/4, uniform otherwise (tests/test_kernels.cpp:26-30)."""
    c = rng.integers(0, 4, shape)
    u = rng.integers(0, 256, shape)
    return np.where(c == 0, 0, np.where(c == 1, 255, u)).astype(np.uint8)


@pytest.mark.parametrize("dc", [64, 128, 1024])
@pytest.mark.parametrize("m,n", [(1, 1), (5, 17), (16, 70)])
def test_kernel_formula(dc, m, n):
    rng = np.random.default_rng(dc * 100 + m * 10 + n)
    qh, ql, xh, xl = (draw_bytes(rng, s) for s in ((m, dc), (m, dc), (n, dc), (n, dc)))
    qc = qh.astype(np.int64) * 256 + ql
    xc = xh.astype(np.int64) * 256 + xl
    qn, xn = (qc * qc).sum(1), (xc * xc).sum(1)
    ssq = 1.0 / (65535.0 * 65535.0)
    ref = ((qn[:, None] + xn[None, :] - 2 * (qc @ xc.T)).astype(np.float64) * ssq).astype(np.float32)
    got = O.sq16_l2(qh, ql, xh, xl, qn, xn, ssq)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    for pd in [p for p in (64, 128, 256, 512) if p < dc]:
        qp = qh[:, :pd].astype(np.int64)
        xp = xh[:, :pd].astype(np.int64)
        refh = (((qp * qp).sum(1)[:, None] + (xp * xp).sum(1)[None, :] - 2 * (qp @ xp.T)).astype(np.float64)
                * ssq).astype(np.float32)
        goth = O.sq16_high_l2(np.ascontiguousarray(qh[:, :pd]), xh, (qp * qp).sum(1), (xp * xp).sum(1), ssq, pd)
        assert np.array_equal(goth.view(np.uint32), refh.view(np.uint32))


@pytest.mark.parametrize("d", [3, 7, 66, 100, 128])
def test_rotation_close_to_fp64(d):
    rng = np.random.default_rng(d)
    x = rng.standard_normal((50, d)).astype(np.float32)
    q, _ = np.linalg.qr(rng.standard_normal((d, d)))
    r = q.T.astype(np.float32)
    got = O.rotate_rows(x, r).astype(np.float64)
    exact = x.astype(np.float64) @ r.astype(np.float64).T
    bound = 2 * d * np.finfo(np.float32).eps * (np.abs(x).astype(np.float64) @ np.abs(r).T.astype(np.float64))
    assert np.all(np.abs(got - exact) <= bound + 1e-30)


def test_quantisation_closed_form(tmp_path):
    base, q = datasets.mixture(2000, 64, d=64, centres=8, seed=3)
    from katoann import builder
    prefix = str(tmp_path / "ix")
    builder.build_numpy(base, prefix, nlist=8, kmeans_iters=3, seed=3)
    ix = O.OracleIndex.from_files(prefix)
    lo, inv, ssq = ix.scale()
    raw = formats.read_index(prefix)
    rot = O.rotate_rows(raw["vecs"], raw["rotation"])
    assert np.float32(lo) == rot.min()
    rng_ = np.float64(rot.max()) - np.float64(np.float32(lo))
    assert inv == 65535.0 / rng_ and ssq == (rng_ / 65535.0) ** 2
    t = (rot.astype(np.float64) - np.float64(lo)) * inv
    want = np.clip(lround_np(t), 0, 65535).astype(np.uint16)
    got = ix.export()["codes"][:, :64]
    assert np.array_equal(got, want)
    qr, qc = ix.queries(q)
    tq = (qr.astype(np.float64) - np.float64(lo)) * inv
    assert np.array_equal(qc[:, :64], np.clip(lround_np(tq), 0, 65535).astype(np.uint16))


def brute_search(codes, ccodes, ids, sizes, qc, P, ssq, k, nprobe, S):
    """Independent numpy restatement of the search under the canonical ties."""
    f = lambda l2: (l2.astype(np.float64) * ssq).astype(np.float32)
    c64, cc64, q64 = codes.astype(np.int64), ccodes.astype(np.int64), qc.astype(np.int64)
    off = np.concatenate([[0], np.cumsum(sizes)])
    out_i = np.full((len(qc), k), 0xFFFFFFFF, np.uint32)
    out_d = np.full((len(qc), k), np.finfo(np.float32).max, np.float32)
    for i, q in enumerate(q64):
        cd = f(((q - cc64) ** 2).sum(1))
        order = np.lexsort((np.arange(len(cd)), cd))[:nprobe]
        rows = np.concatenate([np.arange(off[c], off[c + 1]) for c in order]).astype(np.int64)
        if rows.size == 0:
            continue
        hq, hx = q[:P] >> 8, c64[rows, :P] >> 8
        r = f(((hq - hx) ** 2).sum(1))
        sel = np.lexsort((ids[rows], r))[:S]
        surv = rows[sel]
        ex = f(((q - c64[surv]) ** 2).sum(1))
        heap = []
        for t in range(len(surv) - 1, -1, -1):       # descending (rank, id)
            e, idv = float(ex[t]), int(ids[surv[t]])
            if len(heap) < k:
                heapq.heappush(heap, (-e, -idv))
            elif e < -heap[0][0]:
                heapq.heapreplace(heap, (-e, -idv))
        fin = sorted((-a, -b) for a, b in heap)
        for j, (e, idv) in enumerate(fin):
            out_i[i, j] = idv
            out_d[i, j] = e
    return out_i, out_d


@pytest.mark.parametrize("k,nprobe,S", [(10, 2, 32), (5, 3, 5), (20, 8, 40), (4, 8, 100)])
def test_search_vs_bruteforce(tmp_path, k, nprobe, S):
    base, q = datasets.mixture(1200, 40, d=128, centres=10, seed=9)
    from katoann import builder
    prefix = str(tmp_path / "ix")
    builder.build_numpy(base, prefix, nlist=8, kmeans_iters=3, seed=9)
    ix = O.OracleIndex.from_files(prefix)
    e = ix.export()
    raw = formats.read_index(prefix)
    _, qc = ix.queries(q)
    _, _, ssq = ix.scale()
    bi, bd = brute_search(e["codes"], e["ccodes"], raw["ids"], raw["sizes"], qc, 64, ssq, k, nprobe, S)
    o = ix.search(q, k=k, nprobe=nprobe, survivors=S)
    assert np.array_equal(o["ids"], bi)
    assert np.array_equal(o["dists"].view(np.uint32), bd.view(np.uint32))


def normalize_closed_form(x, acc=np.float64, divide=False):
    """The front ends' row normalisation (python/katoann_module.cpp:97-110,
    tools/katoann.cpp:102-111): s = sum over k, in k order, of double(x_k)^2;
    rows with s == 0 stay as they are; otherwise x_k *= float(1 / sqrt(s)) in
    float. A product of two floats is exact in double, so a fused or unfused
    accumulation gives the same s; np.cumsum adds in k order. acc and divide
    plant mistakes: a float accumulator, or x_k / float(sqrt(s))."""
    x = np.asarray(x, np.float32)
    s = np.cumsum(x.astype(acc) * x.astype(acc), axis=1, dtype=acc)[:, -1].astype(np.float64)
    out = x.copy()
    nz = s > 0
    if divide:
        out[nz] = x[nz] / np.sqrt(s[nz]).astype(np.float32)[:, None]
    else:
        inv = (1.0 / np.sqrt(s[nz])).astype(np.float32)
        out[nz] = x[nz] * inv[:, None]
    return out


def test_normalize_rows_closed_form():
    """oracle_normalize_rows equals the reference's closed form bit for bit on
    rows of every scale (zero rows, denormal entries, large entries, d = 1 to
    3,072); a float accumulator or a division by the norm each differ from it
    on some rows, so the check separates them."""
    rng = np.random.default_rng(97)
    for d in (1, 7, 128, 1024, 3072):
        x = rng.standard_normal((400, d)).astype(np.float32)
        x[0] = 0.0
        x[1] *= np.float32(1e-40)  # denormal entries
        x[2] *= np.float32(1e18)
        x[3, : d // 2] = 0.0
        x[4:100] *= rng.uniform(1e-6, 1e6, (96, 1)).astype(np.float32)
        got = O.normalize_rows(x)
        want = normalize_closed_form(x)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), d
        assert np.array_equal(got[0], x[0])
        if d >= 128:
            assert not np.array_equal(normalize_closed_form(x, acc=np.float32).view(np.uint32), got.view(np.uint32))
            assert not np.array_equal(normalize_closed_form(x, divide=True).view(np.uint32), got.view(np.uint32))
