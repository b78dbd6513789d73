"""The GPU search against the reference (golden fixtures) and the oracle.

Rule: bit-exact ids and distances on every query without a tie at a selection
boundary (a coarse tie at the nprobe-th list, or, for threshold selection, a
ranking tie at the S-th survivor), the reference's own same() on the others
(tests/test_index.cpp:43-52). Ties are detected by the oracle.
"""
import numpy as np
import pytest

import common
import katoann
import oracle as O
from katoann import formats

pytestmark = pytest.mark.gpu


def load(prefix, **kw):
    ix = katoann.Index()
    ix.load(prefix, **kw)
    return ix


@pytest.mark.parametrize("name", common.golden_names())
def test_load_matches_reference(name):
    prefix, q, ref, _ = common.golden(name)
    print("golden load", name, flush=True)
    ix = load(prefix)
    lo, inv, ssq = ix.scale()
    assert np.float32(lo) == ref["lo"] and inv == ref["inv_step"] and ssq == ref["step_sq"]
    codes, ccodes, _ = ix.export_codes()
    import hashlib
    assert hashlib.sha256(codes.tobytes()).hexdigest() == str(ref["codes_sha"])
    assert hashlib.sha256(ccodes.tobytes()).hexdigest() == str(ref["ccodes_sha"])
    assert ix.prefix_dims == int(ref["prefix_dims"])


@pytest.mark.parametrize("name", common.golden_names())
def test_search_matches_reference(name):
    prefix, q, ref, cases = common.golden(name)
    ixs = {"batched": load(prefix), "flat": load(prefix, batched_coarse=False), "bf16": load(prefix, precision="bf16")}
    ox = O.OracleIndex.from_files(prefix)
    rows = 0
    for case in cases:
        k, nprobe, S, ranking, selection, coarse = case
        tag = common.case_tag(case)
        print("golden case", name, tag, flush=True)
        ids, dists, _ = ixs[coarse].search_ex(q, k=k, nprobe=nprobe, ranking=ranking, survivors=S,
                                              selection=selection, normalize=False)
        flags = common.oracle_mode(ox, coarse).search(q, k=k, nprobe=nprobe, ranking=ranking, survivors=S)["flags"]
        tie = (flags & common.tie_mask(ranking, selection)) != 0
        n_ok, _ = common.compare(ids, dists, ref["ids_" + tag], ref["dists_" + tag], tie)
        rows += n_ok
    assert rows > 0


@pytest.fixture(scope="module")
def mixture_index(tmp_path_factory):
    """The CPU-runnable workload's law at a reduced size: 256-centre mixture,
    sigma 0.3, d 128, normalised; nlist 256."""
    d = tmp_path_factory.mktemp("mix")
    prefix, base, q = common.small_index(d, n=60000, nq=1000, d=128, nlist=256, centres=256, seed=1)
    return prefix, base, q


def check_against_oracle(ix, ox, q, k, nprobe, S, ranking="prefix", npq=None):
    ids, dists, ex = ix.search_ex(q, k=k, nprobe=nprobe, ranking=ranking, survivors=S, normalize=False,
                                  nprobe_per_query=npq, want=("probes", "probe_dists", "survivors"))
    o = ox.search(q, k=k, nprobe=nprobe, ranking=ranking, survivors=S, nprobe_per_query=npq)
    coarse_tie = (o["flags"] & 1) != 0
    ok = ~coarse_tie
    # probes: same lists in the same (distance, list) order
    assert np.array_equal(ex["probes"][ok], o["probes"][ok])
    assert np.array_equal(ex["probe_dists"].view(np.uint32), o["probe_dists"].view(np.uint32))
    # survivors: the S smallest (rank, id) keys
    assert np.array_equal(ex["n_survivors"][ok], o["nsurv"][ok])
    assert np.array_equal(ex["survivors"][ok], o["surv"][ok])
    common.compare(ids, dists, o["ids"], o["dists"], coarse_tie)
    return ids, dists, int(coarse_tie.sum())


@pytest.mark.parametrize("k,nprobe,S", [(10, 16, 32), (10, 1, 10), (10, 64, 32), (100, 16, 100),
                                        (5, 8, 128), (32, 256, 64), (10, 64, 256), (10, 128, 500),
                                        (10, 32, 1000), (200, 16, 200)])
def test_mixture_vs_oracle(mixture_index, k, nprobe, S):
    prefix, base, q = mixture_index
    print("mixture", k, nprobe, S, flush=True)
    ix = load(prefix)
    ox = O.OracleIndex.from_files(prefix)
    check_against_oracle(ix, ox, q, k, nprobe, S)


def test_high_byte_ranking(mixture_index):
    prefix, base, q = mixture_index
    ix = load(prefix)
    ox = O.OracleIndex.from_files(prefix)
    check_against_oracle(ix, ox, q[:300], 10, 8, 32, ranking="high")


def test_per_query_nprobe(mixture_index):
    prefix, base, q = mixture_index
    ix = load(prefix)
    ox = O.OracleIndex.from_files(prefix)
    rng = np.random.default_rng(3)
    npq = rng.integers(1, 33, q.shape[0]).astype(np.uint32)
    check_against_oracle(ix, ox, q, 10, 32, 32, npq=npq)


@pytest.mark.parametrize("seed,shift", [(32, 2), (64, 1), (500, 3), (5000, 12)])
def test_seed_band_vs_oracle(mixture_index, wide_index, seed, shift):
    """With a seed band (a query's nearest lists up to `seed` rows first, then
    the rest of them, then probe-rank bands), survivors and results are the
    oracle's, with per-query nprobe too; a seed below S is raised to S."""
    rng = np.random.default_rng(seed)
    for name, (prefix, base, q) in (("mixture", mixture_index), ("wide", wide_index)):
        ix = load(prefix)
        ix.set_seed_rows(seed)
        ix.set_band_factor(shift)
        ix.set_seed_join(False)  # the remainder in a band of its own (test_seed_join_vs_oracle: joined)
        ox = _oracle(prefix)
        for (k, nprobe, S) in ((10, 16, 32), (10, 4, 128), (100, 8, 100)):
            print("seed", seed, shift, name, k, nprobe, S, flush=True)
            check_against_oracle(ix, ox, q[:300], k, min(nprobe, ix.nlist), S)
        npq = rng.integers(1, 33, 300).astype(np.uint32)
        print("seed per-query nprobe", seed, shift, name, flush=True)
        check_against_oracle(ix, ox, q[:300], 10, min(32, ix.nlist), 32, npq=npq)


@pytest.mark.parametrize("seed,shift", [(32, 2), (64, 1), (500, 3), (5000, 12)])
def test_seed_join_vs_oracle(mixture_index, seed, shift):
    """The seed remainder joined to the first probe-rank band (its rows ranked
    once with that band's queries, on the seed band's thresholds): survivors
    and results are the oracle's, with per-query nprobe and an overflowing
    pool too. nprobe <= nlist / 16 here, so the join is taken."""
    rng = np.random.default_rng(seed + 1)
    prefix, base, q = mixture_index
    ix = load(prefix)
    ix.set_seed_rows(seed)
    ix.set_band_factor(shift)
    ix.set_seed_join(True)
    ox = _oracle(prefix)
    for (k, nprobe, S) in ((10, 16, 32), (10, 4, 128), (100, 8, 100), (10, 16, 300)):
        assert ix.tile_config(300, k=k, nprobe=nprobe, survivors=S)["seed_join"] == 1
        print("seed join", seed, shift, k, nprobe, S, flush=True)
        check_against_oracle(ix, ox, q[:300], k, nprobe, S)
    npq = rng.integers(1, 17, 300).astype(np.uint32)
    print("seed join per-query nprobe", seed, shift, flush=True)
    check_against_oracle(ix, ox, q[:300], 10, 16, 32, npq=npq)
    ix.set_pool_allowance(8)
    for rep in range(2):
        print("seed join overflow rerun", seed, shift, rep, flush=True)
        check_against_oracle(ix, ox, q[:300], 10, 16, 32)


def test_pool_overflow_reruns(mixture_index):
    """A pool too small for a band's kept keys: the batch runs again with more
    room (and later searches of the shape start from it); results stay the
    oracle's, with and without a seed band."""
    prefix, base, q = mixture_index
    ox = _oracle(prefix)
    for seed in (0, 64):
        ix = load(prefix)
        ix.set_seed_rows(seed)
        ix.set_pool_allowance(8)
        for rep in range(2):
            print("overflow rerun", seed, rep, flush=True)
            check_against_oracle(ix, ox, q[:300], 10, 16, 32)


def test_ragged_and_single_batches(mixture_index):
    """Each query's result is the same alone, in a ragged batch and in the full batch."""
    prefix, base, q = mixture_index
    ix = load(prefix)
    full_i, full_d, _ = ix.search_ex(q, k=10, nprobe=16, normalize=False)
    for lo_, hi_ in ((0, 1), (5, 22), (100, 357), (999, 1000)):
        i, d, _ = ix.search_ex(q[lo_:hi_], k=10, nprobe=16, normalize=False)
        assert np.array_equal(i, full_i[lo_:hi_]) and np.array_equal(d, full_d[lo_:hi_])
    i2, d2, _ = ix.search_ex(q, k=10, nprobe=16, normalize=False)
    assert np.array_equal(i2, full_i) and np.array_equal(d2, full_d)


def test_padding_when_candidates_run_out(tmp_path):
    """Fewer candidates than k: FLT_MAX / UINT32_MAX padding (src/index.cpp:384-398)."""
    prefix, base, q = common.small_index(tmp_path, n=300, nq=20, d=64, nlist=60, centres=40, seed=4)
    ix = load(prefix)
    ox = O.OracleIndex.from_files(prefix)
    ids, dists, _ = check_against_oracle(ix, ox, q, 40, 1, 40)
    assert (ids == 0xFFFFFFFF).any() and (dists == np.finfo(np.float32).max).any()


def test_python_search_normalises(mixture_index):
    """Index.search normalises the queries of a normalised index on the GPU,
    as the reference module does before Index::search."""
    prefix, base, q = mixture_index
    ix = load(prefix)
    ox = O.OracleIndex.from_files(prefix)
    raw = (q[:200] * np.float32(3.5)).astype(np.float32)
    ids, dists = ix.search(raw, k=10, nprobe=16)
    o = ox.search(O.normalize_rows(raw), k=10, nprobe=16)
    common.compare(ids, dists, o["ids"], o["dists"], (o["flags"] & 1) != 0)


def test_argument_errors(mixture_index):
    prefix, base, q = mixture_index
    ix = load(prefix)
    with pytest.raises(ValueError, match="need k >= 1"):
        ix.search(q[:2], k=0)
    with pytest.raises(ValueError, match="need k >= 1"):
        ix.search(q[:2], nprobe=ix.nlist + 1)
    with pytest.raises(ValueError, match="queries must be"):
        ix.search(q[:2, :5])


def test_build_load_search_end_to_end(tmp_path):
    """katoann.build (host C++ builder) -> load -> search on the GPU, checked
    against brute force recall and the oracle."""
    from katoann import datasets
    base, q = datasets.decaying_clusters(8000, 100, d=128, seed=11)
    prefix = str(tmp_path / "e2e")
    katoann.build(base, prefix, nlist=32, seed=1, normalized=True)
    ix = load(prefix)
    assert ix.rotated and ix.prefix_dims == 64 and ix.size == 8000
    ids, dists = ix.search(q, k=10, nprobe=32, survivors=128)
    gt = datasets.ground_truth_np(base, q, 10)
    recall = np.mean([len(set(ids[i]) & set(gt[i])) / 10 for i in range(len(q))])
    assert recall >= 0.95
    ox = O.OracleIndex.from_files(prefix)
    o = ox.search(q, k=10, nprobe=32, survivors=128)
    common.compare(ids, dists, o["ids"], o["dists"], (o["flags"] & 1) != 0)


@pytest.fixture(scope="module")
def wide_index(tmp_path_factory):
    """d 1024 (prefix 256 dims, high-byte ranking over 1024): the ranking
    widths of the embedding-shape workloads at a size the oracle runs fast."""
    d = tmp_path_factory.mktemp("wide")
    from katoann import builder, datasets
    base, q = datasets.spectrum(12000, 200, 1024, alpha=1.0, seed=5)
    prefix = str(d / "wide")
    builder.build_numpy(base, prefix, nlist=48, kmeans_iters=4, seed=5, normalized=True)
    return prefix, base, q


# (kernel, variant, queries per item): block items; wave items with the large and
# the small row tile and with the deep query ring; wave items split into many
# query chunks; block items with the query ring
RANK_KERNELS = [(0, 0, 512), (1, 0, 512), (1, 1, 512), (1, 2, 512), (1, 3, 512), (1, 0, 16), (1, 1, 1), (1, 2, 40),
                (1, 2, 1), (1, 3, 24), (2, 0, 512), (2, 1, 512), (2, 0, 20), (2, 1, 1),
                (3, 0, 1024), (3, 1, 1024), (3, 0, 16), (3, 1, 48), (3, 0, 4096), (4, 0, 512), (4, 0, 16),
                (4, 0, 100), (5, 0, 512), (5, 0, 64), (5, 0, 128), (5, 0, 8192)]


_ORACLES = {}


def _oracle(prefix):
    if prefix not in _ORACLES:
        _ORACLES[prefix] = O.OracleIndex.from_files(prefix)
    return _ORACLES[prefix]


@pytest.mark.parametrize("kernel,variant,qc", RANK_KERNELS)
def test_rank_kernels_vs_oracle(mixture_index, wide_index, kernel, variant, qc):
    """Every ranking kernel and tile gives the oracle's survivors and results,
    for prefix ranking (P 64 and 256) and high-byte ranking (P 128 and 1024)."""
    for name, (prefix, base, q) in (("mixture", mixture_index), ("wide", wide_index)):
        ix = load(prefix)
        ix.set_rank_kernel(kernel, variant, qc)
        ox = _oracle(prefix)
        for (k, nprobe, S, ranking) in ((10, 16, 32, "prefix"), (10, 4, 128, "prefix"), (100, 8, 100, "prefix"),
                                        (10, 8, 32, "high")):
            print("rank kernel", kernel, variant, qc, name, k, nprobe, S, ranking, flush=True)
            check_against_oracle(ix, ox, q[:200], k, min(nprobe, ix.nlist), S, ranking=ranking)


def test_rank_kernel_choice_is_invisible(mixture_index):
    """The same batch gives bit-identical ids and distances under every ranking
    kernel and band factor."""
    prefix, base, q = mixture_index
    ix = load(prefix)
    ref = None
    for (kernel, variant, qc) in RANK_KERNELS:
        for shift, seed in ((1, 0), (2, 0), (4, 0), (7, 0), (2, 64), (3, 300), (12, 100)):
            ix.set_rank_kernel(kernel, variant, qc)
            ix.set_band_factor(shift)
            ix.set_seed_rows(seed)
            print("choice", kernel, variant, qc, shift, seed, flush=True)
            i, d, _ = ix.search_ex(q, k=10, nprobe=64, survivors=64, normalize=False)
            if ref is None:
                ref = (i, d)
            assert np.array_equal(i, ref[0]) and np.array_equal(d.view(np.uint32), ref[1].view(np.uint32))


def test_graph_replays_match_eager(mixture_index):
    """Searches replayed from the captured graph equal eager searches, replay
    after replay, with other shapes and eager searches run in between."""
    prefix, base, q = mixture_index
    ix = load(prefix)
    ix.set_graphs(False)
    ref = ix.search_ex(q, k=10, nprobe=32, survivors=64, normalize=False)[:2]
    other = ix.search_ex(q[:77], k=5, nprobe=8, survivors=32, normalize=False)[:2]
    ix.set_graphs(True)
    for rep in range(4):
        print("graph replay", rep, flush=True)
        for _ in range(2):  # a capture, then a replay of it
            i, d, _ = ix.search_ex(q, k=10, nprobe=32, survivors=64, normalize=False)
            assert np.array_equal(i, ref[0]) and np.array_equal(d.view(np.uint32), ref[1].view(np.uint32))
        i2, d2, _ = ix.search_ex(q[:77], k=5, nprobe=8, survivors=32, normalize=False)
        assert np.array_equal(i2, other[0]) and np.array_equal(d2.view(np.uint32), other[1].view(np.uint32))
        ix.set_graphs(False)
        i3, _, _ = ix.search_ex(q[:300], k=10, nprobe=16, normalize=False)
        ix.set_graphs(True)


def test_autotune_keeps_results(mixture_index):
    """With autotuning on, the first search of a shape times the candidate
    tiles and keeps one; every later search of that shape runs with it, and
    results equal those of the default tile."""
    prefix, base, q = mixture_index
    ix = load(prefix)
    ref = ix.search_ex(q, k=10, nprobe=64, survivors=64, normalize=False)[:2]
    default = ix.tile_config(q.shape[0], k=10, nprobe=64, survivors=64)
    ix.set_autotune(True)
    for rep in range(2):
        print("autotune", rep, flush=True)
        i, d, _ = ix.search_ex(q, k=10, nprobe=64, survivors=64, normalize=False)
        assert np.array_equal(i, ref[0]) and np.array_equal(d.view(np.uint32), ref[1].view(np.uint32))
    tuned = ix.tile_config(q.shape[0], k=10, nprobe=64, survivors=64)
    assert tuned["kernel"] in (1, 2, 3, 4, 5) and tuned["queries_per_item"] in (256, 512, 1024, 4096)
    print("default tile", default, "tuned", tuned, flush=True)


@pytest.fixture(scope="module")
def many_lists(tmp_path_factory):
    """Coarse rows of 5,000 and 16,384 lists (the selection's register rows of
    32 and 64 lists per thread), built on the GPU, written as files."""
    import torch
    from katoann import builder, datasets
    d = tmp_path_factory.mktemp("many")
    base, q = datasets.spectrum_torch(40000, 200, 128, alpha=1.0, seed=9, device="cuda")
    out = {}
    for nl in (5000, 16384):
        prefix = str(d / ("nl%d" % nl))
        builder.build_torch(base, prefix, nlist=nl, train_size=40000, kmeans_iters=3, seed=1, normalized=True)
        out[nl] = prefix
    return out, q.cpu().numpy()


@pytest.mark.parametrize("nl", [5000, 16384])
def test_selection_many_lists_vs_oracle(many_lists, nl):
    prefixes, q = many_lists
    ix = load(prefixes[nl])
    ox = _oracle(prefixes[nl])
    for nprobe in (1, 77, 4096):
        print("many lists", nl, nprobe, flush=True)
        check_against_oracle(ix, ox, q[:64], 10, nprobe, 32)


def test_concurrent_searches_match_sequential(mixture_index):
    """Searches from several host threads at once (each on its own search
    context, stream and captured graph, shapes mixed and repeated) return what
    the same searches return one at a time."""
    import threading
    prefix, base, q = mixture_index
    ix = load(prefix)
    shapes = [(10, 16, 32), (10, 64, 128), (100, 8, 100), (10, 32, 64)]
    ref = {s: ix.search_ex(q[:500], k=s[0], nprobe=s[1], survivors=s[2], normalize=False)[:2] for s in shapes}
    errors = []

    def run(t):
        try:
            for rep in range(3):
                for s in shapes[t % len(shapes):] + shapes[:t % len(shapes)]:
                    i, d, _ = ix.search_ex(q[:500], k=s[0], nprobe=s[1], survivors=s[2], normalize=False)
                    if not (np.array_equal(i, ref[s][0]) and np.array_equal(d.view(np.uint32), ref[s][1].view(np.uint32))):
                        errors.append((t, rep, s))
        except Exception as e:  # surfaced below
            errors.append((t, repr(e)))

    print("concurrent searches: 3 threads", flush=True)
    ts = [threading.Thread(target=run, args=(t,)) for t in range(3)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors


@pytest.mark.parametrize("k,nprobe,S,ranking", [(10, 64, 2000, "prefix"), (10, 256, 4000, "prefix"),
                                                (1500, 64, 1500, "prefix"), (10, 128, 4096, "high"),
                                                (1200, 32, 3000, "prefix")])
def test_survivors_above_1024_vs_oracle(mixture_index, k, nprobe, S, ranking):
    """max(survivors, k) up to 4,096, as the reference allows (src/index.cpp:914):
    the block-level pool update and refinement give the oracle's survivors and
    results."""
    prefix, base, q = mixture_index
    print("survivors", k, nprobe, S, ranking, flush=True)
    ix = load(prefix)
    check_against_oracle(ix, _oracle(prefix), q[:300], k, nprobe, S, ranking=ranking)


@pytest.mark.parametrize("k,nprobe", [(1100, 64), (4096, 256)])
def test_full_ranking_k_above_1024_vs_oracle(mixture_index, k, nprobe):
    """Full ranking keeps the k nearest of all candidates, k up to 4,096."""
    prefix, base, q = mixture_index
    print("full ranking k", k, nprobe, flush=True)
    ix = load(prefix)
    ids, dists, _ = ix.search_ex(q[:200], k=k, nprobe=nprobe, ranking="full", normalize=False)
    o = _oracle(prefix).search(q[:200], k=k, nprobe=nprobe, ranking="full")
    common.compare(ids, dists, o["ids"], o["dists"], (o["flags"] & common.tie_mask("full", "threshold")) != 0)


def test_survivors_above_4096_refused(mixture_index):
    prefix, base, q = mixture_index
    ix = load(prefix)
    with pytest.raises(ValueError, match="above 4096"):
        ix.search(q[:2], k=10, survivors=4097)


def test_device_per_query_nprobe_out_of_range(mixture_index):
    """Per-query nprobe in device memory: an entry of 0 or above nprobe is
    refused (the host path's std::invalid_argument), after kernels that only
    read probed ranks; a valid array then searches as the host path does."""
    import torch
    prefix, base, q = mixture_index
    ix = load(prefix)
    nq, k = 64, 10
    qd = torch.from_numpy(np.ascontiguousarray(q[:nq])).cuda()
    ids = torch.empty((nq, k), dtype=torch.int32, device="cuda")
    dists = torch.empty((nq, k), dtype=torch.float32, device="cuda")
    for bad in (0, 33):
        npq = torch.full((nq,), 8, dtype=torch.int32, device="cuda")
        npq[5] = bad
        print("device nprobe entry", bad, flush=True)
        with pytest.raises(ValueError, match="per-query nprobe"):
            ix.search_device(qd.data_ptr(), nq, ids.data_ptr(), dists.data_ptr(), k=k, nprobe=32,
                             npq_ptr=npq.data_ptr())
        torch.cuda.synchronize()
    npq = torch.from_numpy(np.arange(1, nq + 1, dtype=np.int32) % 32 + 1).cuda()
    ix.search_device(qd.data_ptr(), nq, ids.data_ptr(), dists.data_ptr(), k=k, nprobe=32, npq_ptr=npq.data_ptr())
    torch.cuda.synchronize()
    hi, hd, _ = ix.search_ex(q[:nq], k=k, nprobe=32, normalize=False,
                             nprobe_per_query=npq.cpu().numpy().astype(np.uint32))
    assert np.array_equal(ids.cpu().numpy().view(np.uint32), hi)
    assert np.array_equal(dists.cpu().numpy().view(np.uint32), hd.view(np.uint32))


def test_footprint_covers_search_after_overflow(mixture_index):
    """The footprint a serving loop reserves covers the buffers a search of
    that shape takes, also after an overflow rerun raised the shape's room."""
    prefix, base, q = mixture_index
    ix = load(prefix)
    ix.set_pool_allowance(8)
    srv_like = katoann.Server(ix, max_inflight=1)
    try:
        for rep in range(3):
            fp = srv_like.footprint(300, k=10, nprobe=16, survivors=32)
            used = ix.search_ex(q[:300], k=10, nprobe=16, survivors=32, normalize=False)[2]["device_bytes"]
            print("footprint", rep, fp, used, flush=True)
            if rep > 0:
                assert used <= fp
    finally:
        srv_like.close()
