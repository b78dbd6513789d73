"""Ranking::kFull on the GPU (src/index.cpp:1130-1219): the exact 16-bit
distance of every candidate of the probed lists, the k nearest by
(distance, id); the per-query fp32 coarse search (LoadOptions::batched_coarse
= false, src/index.cpp:832-844); and the reference's equivalences between the
rankings, the selections and the coarse searches (tests/test_index.cpp:117-160).

Rule: bit-exact ids and distances against the oracle on every query without a
coarse tie at the nprobe-th list or a tie at the k-th exact distance (oracle
flags 1 and 4); the reference's same() on the others.
"""
import numpy as np
import pytest

import common
import katoann
import oracle as O

pytestmark = pytest.mark.gpu


def load(prefix, **kw):
    ix = katoann.Index()
    ix.load(prefix, **kw)
    return ix


@pytest.fixture(scope="module")
def decay_index(tmp_path_factory):
    """The law and shape of the reference's index test (tests/test_index.cpp:65-108):
    64 centres, variance decaying along the dimensions, d 256, 40,000 points,
    300 queries, nlist 64; brute-force top-10 by f64 L2 (ties by id)."""
    from katoann import builder, datasets
    base, q = datasets.decaying_clusters(40000, 300, d=256, seed=11)
    prefix = str(tmp_path_factory.mktemp("decay") / "idx")
    builder.build_numpy(base, prefix, nlist=64, kmeans_iters=8, seed=1, normalized=True, rotation_sample=20000)
    gt = np.empty((len(q), 10), np.int64)
    for i in range(len(q)):
        x = ((base.astype(np.float64) - q[i]) ** 2).sum(1)
        gt[i] = np.lexsort((np.arange(len(x)), x))[:10]
    return prefix, q, gt


@pytest.fixture(scope="module")
def wide_index(tmp_path_factory):
    """d 1024 (sixteen 64-byte slices per plane), power-law spectrum."""
    from katoann import builder, datasets
    base, q = datasets.spectrum(12000, 200, 1024, alpha=1.0, seed=5)
    prefix = str(tmp_path_factory.mktemp("widefull") / "wide")
    builder.build_numpy(base, prefix, nlist=48, kmeans_iters=4, seed=5, normalized=True)
    return prefix, q


def recall(ids, gt):
    return float(np.mean([len(set(ids[i].tolist()) & set(gt[i].tolist())) / gt.shape[1] for i in range(len(gt))]))


def check_full(ix, ox, q, k, nprobe):
    ids, dists, ex = ix.search_ex(q, k=k, nprobe=nprobe, ranking="full", normalize=False,
                                  want=("probes", "survivors"))
    o = ox.search(q, k=k, nprobe=nprobe, ranking="full")
    tie = (o["flags"] & common.tie_mask("full", "threshold")) != 0
    ok = (o["flags"] & 1) == 0
    assert np.array_equal(ex["probes"][ok], o["probes"][ok])
    n_ok, n_tie = common.compare(ids, dists, o["ids"], o["dists"], tie)
    # the survivors output holds the k keys themselves, ascending
    keys = (dists.view(np.uint32).astype(np.uint64) << np.uint64(32)) | ids.astype(np.uint64)
    assert np.array_equal(ex["survivors"], keys)
    return n_ok, n_tie


@pytest.mark.parametrize("k,nprobe", [(10, 1), (10, 16), (32, 64), (100, 8), (1, 4), (10, 64)])
def test_full_vs_oracle(decay_index, k, nprobe):
    prefix, q, _ = decay_index
    print("full decay", k, nprobe, flush=True)
    n_ok, _ = check_full(load(prefix), O.OracleIndex.from_files(prefix), q, k, nprobe)
    assert n_ok > 0


@pytest.mark.parametrize("k,nprobe", [(10, 4), (10, 48), (64, 12)])
def test_full_vs_oracle_d1024(wide_index, k, nprobe):
    prefix, q = wide_index
    print("full wide", k, nprobe, flush=True)
    n_ok, _ = check_full(load(prefix), O.OracleIndex.from_files(prefix), q, k, nprobe)
    assert n_ok > 0


def test_full_bands_and_pool_room_are_invisible(decay_index):
    """Band factors, seed bands and pool overflow reruns change nothing."""
    prefix, q, _ = decay_index
    ix = load(prefix)
    ref = None
    for shift, seed, room in ((3, 0, 0), (1, 0, 0), (5, 64, 0), (2, 1000, 0), (12, 0, 0), (3, 0, 16), (1, 32, 8)):
        ix.set_band_factor(shift)
        ix.set_seed_rows(seed)
        ix.set_pool_allowance(room)
        print("full bands", shift, seed, room, flush=True)
        i, d, _ = ix.search_ex(q, k=10, nprobe=32, ranking="full", normalize=False)
        if ref is None:
            ref = (i, d)
        assert np.array_equal(i, ref[0]) and np.array_equal(d.view(np.uint32), ref[1].view(np.uint32))


def test_full_ragged_batches(decay_index):
    """Each query's result is the same alone, in a ragged batch and in the whole batch."""
    prefix, q, _ = decay_index
    ix = load(prefix)
    all_i, all_d, _ = ix.search_ex(q, k=10, nprobe=16, ranking="full", normalize=False)
    for lo, hi in ((0, 1), (5, 22), (100, 300)):
        i, d, _ = ix.search_ex(q[lo:hi], k=10, nprobe=16, ranking="full", normalize=False)
        assert np.array_equal(i, all_i[lo:hi]) and np.array_equal(d, all_d[lo:hi])


def test_reference_equivalences(decay_index):
    """tests/test_index.cpp:117-152 on the GPU: the full scan of every list
    finds the brute-force neighbours; prefix and high-byte ranking with a wide
    survivor set (1,024 of 40,000 here; the reference uses 4,000, above this
    engine's survivor limit) equal the full scan; local equals threshold
    selection; prefix ranking at nprobe 16 keeps the full scan's recall."""
    prefix, q, gt = decay_index
    ix = load(prefix)
    assert ix.rotated and ix.prefix_dims == 64
    fi, fd, _ = ix.search_ex(q, k=10, nprobe=64, ranking="full", normalize=False)
    assert recall(fi, gt) >= 0.99
    for r in ("high", "prefix"):
        wi, wd, _ = ix.search_ex(q, k=10, nprobe=64, ranking=r, survivors=1024, normalize=False)
        assert common.same(fi, fd, wi, wd), r
    ti, td, _ = ix.search_ex(q, k=10, nprobe=16, normalize=False)
    li, ld, _ = ix.search_ex(q, k=10, nprobe=16, selection="local", normalize=False)
    assert common.same(ti, td, li, ld)
    f16, _, _ = ix.search_ex(q, k=10, nprobe=16, ranking="full", normalize=False)
    r_full, r_prefix = recall(f16, gt), recall(ti, gt)
    assert r_full >= 0.9 and r_prefix >= r_full - 0.02
    flat = load(prefix, batched_coarse=False)
    assert not flat.batched_coarse and ix.batched_coarse
    fl, _, _ = flat.search_ex(q, k=10, nprobe=16, normalize=False)
    assert recall(fl, gt) >= r_prefix - 0.02


@pytest.mark.parametrize("name", common.flat_golden_names())
def test_flat_order_matches_reference(name):
    """The per-query fp32 coarse search (batched_coarse = false) on the GPU
    orders every list as the reference's coarse_flat_ does, and its distances
    are the oracle's bit for bit."""
    import os
    from katoann import formats
    d = os.path.join(common.GOLDEN, name)
    ref = np.load(os.path.join(d, "reference_outputs.npz"))["flat_order"]
    q = formats.read_fbin(os.path.join(d, "queries.fbin"))
    prefix = os.path.join(d, "idx")
    ix = load(prefix, batched_coarse=False)
    nl = ix.nlist
    print("flat order", name, nl, flush=True)
    _, _, ex = ix.search_ex(q, k=10, nprobe=nl, ranking="high", survivors=10, normalize=False,
                            want=("probes", "probe_dists"))
    assert np.array_equal(ex["probes"], ref)
    D = O.OracleIndex.from_files(prefix).flat_dists(q)
    assert np.array_equal(ex["probe_dists"].view(np.uint32),
                          np.take_along_axis(D, ref.astype(np.int64), 1).view(np.uint32))


@pytest.mark.parametrize("k,nprobe,ranking", [(10, 8, "prefix"), (10, 16, "full"), (32, 4, "high")])
def test_flat_search_vs_oracle(decay_index, k, nprobe, ranking):
    """Searches after the per-query coarse search equal the oracle's."""
    prefix, q, _ = decay_index
    ix = load(prefix, batched_coarse=False)
    ox = O.OracleIndex.from_files(prefix).set_flat()
    ids, dists, ex = ix.search_ex(q, k=k, nprobe=nprobe, ranking=ranking, survivors=32, normalize=False,
                                  want=("probes",))
    o = ox.search(q, k=k, nprobe=nprobe, ranking=ranking, survivors=32)
    assert np.array_equal(ex["probes"], o["probes"])
    common.compare(ids, dists, o["ids"], o["dists"], (o["flags"] & common.tie_mask(ranking, "threshold")) != 0)


def check_bf16(ix, ox, q, k, nprobe):
    ids, dists, ex = ix.search_ex(q, k=k, nprobe=nprobe, ranking="full", normalize=False, want=("probes",))
    o = ox.search(q, k=k, nprobe=nprobe, ranking="full")
    assert np.array_equal(ex["probes"], o["probes"])
    return common.compare(ids, dists, o["ids"], o["dists"], (o["flags"] & common.tie_mask("full", "threshold")) != 0)


@pytest.mark.parametrize("k,nprobe", [(10, 1), (10, 16), (32, 64), (100, 8)])
def test_bf16_vs_oracle(decay_index, k, nprobe):
    """bf16 precision (src/index.cpp:752-775, 1221-1293): per-query coarse
    search, then every candidate's distance in the reference's bf16
    arithmetic, the k nearest by (distance, id)."""
    prefix, q, _ = decay_index
    print("bf16 decay", k, nprobe, flush=True)
    n_ok, _ = check_bf16(load(prefix, precision="bf16"), O.OracleIndex.from_files(prefix).set_bf16(), q, k, nprobe)
    assert n_ok > 0


def test_bf16_d1024_and_api(wide_index):
    prefix, q = wide_index
    ix = load(prefix, precision="bf16")
    assert not ix.rotated and not ix.batched_coarse and ix.prefix_dims == 0
    n_ok, _ = check_bf16(ix, O.OracleIndex.from_files(prefix).set_bf16(), q, 10, 12)
    assert n_ok > 0
    for ranking in ("prefix", "high"):
        with pytest.raises(ValueError, match="bf16 precision supports only Ranking::kFull"):
            ix.search(q[:4], k=10, nprobe=4, ranking=ranking)


def test_streamed_load_in_small_chunks(decay_index):
    """The .invlist vectors stream to the GPU through pinned staging buffers;
    with 64 KiB buffers (hundreds of chunks, lists split across them) the
    loaded codes and the searches equal those of a load in one chunk."""
    prefix, q, _ = decay_index
    a = load(prefix)
    try:
        katoann.set_load_chunk(64 << 10)
        b = load(prefix)
        c = load(prefix, precision="bf16")
    finally:
        katoann.set_load_chunk(64 << 20)
    ca, cca, _ = a.export_codes()
    cb, ccb, _ = b.export_codes()
    assert np.array_equal(ca, cb) and np.array_equal(cca, ccb)
    ia, da, _ = a.search_ex(q, k=10, nprobe=16, normalize=False)
    ib, db, _ = b.search_ex(q, k=10, nprobe=16, normalize=False)
    assert np.array_equal(ia, ib) and np.array_equal(da, db)
    n_ok, _ = check_bf16(c, O.OracleIndex.from_files(prefix).set_bf16(), q, 10, 8)
    assert n_ok > 0


@pytest.mark.parametrize("mode", ["full", "flat", "bf16"])
def test_per_query_nprobe_in_every_mode(decay_index, mode):
    """Per-query nprobe (the learned-nprobe input) through full ranking, the
    per-query coarse search and bf16 precision, against the oracle."""
    prefix, q, _ = decay_index
    rng = np.random.default_rng(3)
    npq = rng.integers(1, 33, len(q)).astype(np.uint32)
    if mode == "bf16":
        ix, ox, ranking = load(prefix, precision="bf16"), O.OracleIndex.from_files(prefix).set_bf16(), "full"
    elif mode == "flat":
        ix, ox, ranking = load(prefix, batched_coarse=False), O.OracleIndex.from_files(prefix).set_flat(), "prefix"
    else:
        ix, ox, ranking = load(prefix), O.OracleIndex.from_files(prefix), "full"
    print("per-query nprobe", mode, flush=True)
    ids, dists, _ = ix.search_ex(q, k=10, nprobe=32, ranking=ranking, survivors=64, normalize=False,
                                 nprobe_per_query=npq)
    o = ox.search(q, k=10, nprobe=32, ranking=ranking, survivors=64, nprobe_per_query=npq)
    n_ok, _ = common.compare(ids, dists, o["ids"], o["dists"], (o["flags"] & common.tie_mask(ranking, "threshold")) != 0)
    assert n_ok > 0
