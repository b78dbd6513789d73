"""Parity at the workloads' own shapes (the batch of 4,096 queries, the full
index), checked on query samples against the oracle over the lists they probe,
and for the 1M workload against the reference's own search as well:

* 100,000 x 128 Gaussian mixture (256 centres, sigma 0.3), nlist 256, nprobe
  16, survivors 32, k 10, one batch of 1,000 (the CPU-runnable search): every
  query against the oracle built from the files;
* 1,000,000 x 1,024 i^-1 spectrum, nlist 4,096, nprobe 16 / 64 / 256,
  survivors 32, k 10, batch 4,096: 256 sampled queries against the oracle
  (probes, survivors, ids, distances) and against the reference build's
  Index::search on the same files (oracle/_ref/katoann_ref_search), bit-exact
  away from tie rows, the reference's same() on them;
* 990,000 x 3,072 i^-0.8 spectrum, nlist 4,096, prefix 768 dims, survivors 32,
  k 10, batch 4,096: sampled queries against the oracle at nprobe 16 and 64.
"""
import os
import tempfile

import numpy as np
import pytest

import common
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


def test_mixture_100k_every_query(tmp_path):
    import katoann
    print("100k x 128 mixture: build", flush=True)
    prefix, base, q = common.small_index(tmp_path, n=100000, nq=1000, d=128, nlist=256, centres=256, seed=1,
                                         iters=6)
    ix = katoann.Index()
    ix.load(prefix)
    ox = O.OracleIndex.from_files(prefix)
    ids, dists, ex = ix.search_ex(q, k=10, nprobe=16, survivors=32, normalize=False,
                                  want=("probes", "survivors"))
    o = ox.search(q, k=10, nprobe=16, survivors=32)
    ok = (o["flags"] & 1) == 0
    assert np.array_equal(ex["probes"][ok], o["probes"][ok])
    assert np.array_equal(ex["survivors"][ok], o["surv"][ok])
    n_ok, _ = common.compare(ids, dists, o["ids"], o["dists"], ~ok)
    print("100k: %d of 1000 queries bit-exact, %d coarse-tie rows by same()" % (n_ok, (~ok).sum()), flush=True)
    assert n_ok >= 950


@pytest.fixture(scope="module")
def spectrum_1m():
    import torch
    from katoann import datasets
    print("1M x 1024 spectrum: data + index", flush=True)
    base, q = datasets.spectrum_torch(1_000_000, 4096, 1024, alpha=1.0, seed=2, device="cuda")
    ix, info = W.gpu_index(base, 4096)
    torch.cuda.synchronize()
    yield base, q.cpu().numpy(), ix, info
    del ix, base
    torch.cuda.empty_cache()


SAMPLE_1M = np.arange(0, 4096, 16)  # 256 queries spread over the batch


@pytest.mark.parametrize("nprobe", [16, 64, 256])
def test_spectrum_1m_vs_oracle(spectrum_1m, nprobe):
    base, q, ix, info = spectrum_1m
    _, _, sample, n_ok, n_tie = W.check_sample(ix, info, q, SAMPLE_1M, 10, nprobe, 32, tag="1M")
    assert n_ok >= 0.95 * len(sample)


@pytest.mark.skipif(not os.path.exists(O.REF_SEARCH), reason="reference build absent (make -C oracle ref)")
def test_spectrum_1m_vs_reference_build(spectrum_1m):
    """The reference's own Index::search (AVX-512 engine) on the same index
    files and the 256 sampled queries, threshold selection as its default:
    ids and distances bit-exact on rows the oracle finds free of ties, the
    reference's same() on the others."""
    import torch
    from katoann import formats
    base, q, ix, info = spectrum_1m
    nprobes = [16, 64, 256]
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        prefix = os.path.join(d, "s1m")
        order = info["order"].cpu().numpy().astype(np.uint32)
        bounds = np.concatenate([[0], np.cumsum(info["sizes"].astype(np.int64))])
        host = base.cpu().numpy()
        lists = [order[bounds[c]:bounds[c + 1]] for c in range(ix.nlist)]
        print("1M: writing the index files", flush=True)
        formats.write_index(prefix, info["centroids"].cpu().numpy(), lists, [host[i] for i in lists],
                            host.shape[0], True, info["rotation_np"], ix.prefix_dims)
        del host
        qf = os.path.join(d, "q.fbin")
        formats.write_fbin(qf, q[SAMPLE_1M])
        print("1M: reference search (load + %s)" % nprobes, flush=True)
        ref = O.ref_search(prefix, qf, 10, nprobes, 32, "prefix", "threshold")
    for nprobe, (r_ids, r_d) in zip(nprobes, ref):
        ids, dists, ex = ix.search_ex(q, k=10, nprobe=nprobe, survivors=32, normalize=False, want=("probes",))
        ox = W.sample_oracle(ix, info, ex["probes"][SAMPLE_1M])
        flags = ox.search(q[SAMPLE_1M], k=10, nprobe=nprobe, survivors=32)["flags"]
        tie = (flags & common.tie_mask("prefix", "threshold")) != 0
        n_ok, n_tie = common.compare(ids[SAMPLE_1M], dists[SAMPLE_1M], r_ids.reshape(-1, 10),
                                     r_d.reshape(-1, 10), tie)
        print("1M vs reference, nprobe %d: %d rows bit-exact, %d tie rows by same()" % (nprobe, n_ok, n_tie),
              flush=True)
        assert n_ok >= 0.9 * len(SAMPLE_1M)
    torch.cuda.synchronize()


@pytest.fixture(scope="module")
def spectrum_3072():
    import torch
    from katoann import datasets
    print("990k x 3072 spectrum: data + index", flush=True)
    base, q = datasets.spectrum_torch(990_000, 4096, 3072, alpha=0.8, seed=4, device="cuda")
    ix, info = W.gpu_index(base, 4096)
    torch.cuda.synchronize()
    del base
    torch.cuda.empty_cache()
    yield q.cpu().numpy(), ix, info
    del ix
    torch.cuda.empty_cache()


@pytest.mark.parametrize("nprobe", [16, 64])
def test_spectrum_3072_vs_oracle(spectrum_3072, nprobe):
    q, ix, info = spectrum_3072
    assert ix.prefix_dims == 768
    _, _, sample, n_ok, _ = W.check_sample(ix, info, q, np.arange(0, 4096, 32), 10, nprobe, 32, tag="3072")
    assert n_ok >= 0.95 * len(sample)
