"""Shared pieces of the GPU tests at the workloads' own shapes: seeded data
drawn on the GPU, the index built and loaded from device memory, and the
oracle restricted to the lists a query sample probes (the codes of those lists
exported from the GPU index: the load path is pinned on its own by the golden
fixtures, so the search is checked on the very codes it ran on)."""
from __future__ import annotations

import numpy as np

import common
import oracle as O


def gpu_index(base, nlist, seed=1, iters=10, train=256000):
    """(Index, info) for a (n, d) float32 CUDA tensor: built on the GPU and
    loaded from device memory; info holds sizes (numpy) and rotation (numpy)."""
    from katoann import builder
    ix, info = builder.load_torch(base, nlist, train_size=train, kmeans_iters=iters, seed=seed)
    info = dict(info, rotation_np=info["rotation"].cpu().numpy())
    return ix, info


def rows_of(info, lists):
    return int(info["sizes"][np.unique(np.asarray(lists).ravel())].astype(np.int64).sum())


def sample_oracle(ix, info, lists):
    """The oracle over the rows of `lists` (the other lists empty): exact for
    every query whose probes are among them."""
    lists = np.unique(np.asarray(lists, np.uint32).ravel())
    sizes = info["sizes"]
    sub = np.zeros_like(sizes)
    sub[lists] = sizes[lists]
    codes, ids = ix.export_lists(lists)
    lo, inv, ssq = ix.scale()
    return O.OracleIndex.from_codes(ix.dim, ix.nlist, ix.size, sub, ids, codes, ix.centroid_codes(),
                                    info["rotation_np"], lo, inv, ssq, ix.prefix_dims)


def check_sample(ix, info, q, sample, k, nprobe, S, ranking="prefix", row_budget=None, tag=""):
    """Searches the whole batch q on the GPU; on the sample's rows, compares
    probes, probe distances, survivor keys, ids and distances with the oracle,
    bit for bit (both follow the canonical tie rule); rows with a coarse tie at
    the nprobe-th list are left out (the restricted oracle may lack the list
    the other side chose). Returns (GPU ids, dists, the sample used, exact
    rows, coarse-tie rows)."""
    ids, dists, ex = ix.search_ex(q, k=k, nprobe=nprobe, ranking=ranking, survivors=S, normalize=False,
                                  want=("probes", "probe_dists", "survivors"))
    sample = np.asarray(sample)
    if row_budget:
        while len(sample) > 1 and rows_of(info, ex["probes"][sample]) > row_budget:
            sample = sample[: max(1, len(sample) * 3 // 4)]
    print("%s nprobe %d S %d: sample of %d queries, %d rows of %d lists" % (
        tag, nprobe, S, len(sample), rows_of(info, ex["probes"][sample]),
        len(np.unique(ex["probes"][sample]))), flush=True)
    ox = sample_oracle(ix, info, ex["probes"][sample])
    o = ox.search(q[sample], k=k, nprobe=nprobe, ranking=ranking, survivors=S)
    coarse_tie = (o["flags"] & 1) != 0
    ok = ~coarse_tie
    assert coarse_tie.mean() <= 0.05, "too many coarse ties in the sample: %d" % coarse_tie.sum()
    assert np.array_equal(ex["probes"][sample][ok], o["probes"][ok])
    assert np.array_equal(ex["probe_dists"][sample][ok].view(np.uint32), o["probe_dists"][ok].view(np.uint32))
    if ranking != "full":
        assert np.array_equal(ex["n_survivors"][sample][ok], o["nsurv"][ok])
        assert np.array_equal(ex["survivors"][sample][ok], o["surv"][ok])
    # both sides follow the canonical tie rule: bit-exact on every row left
    n_ok, _ = common.compare(ids[sample][ok], dists[sample][ok], o["ids"][ok], o["dists"][ok],
                             np.zeros(int(ok.sum()), bool))
    return ids, dists, sample, n_ok, int(coarse_tie.sum())


def recall(ids, gt, k=10):
    return float(np.mean([len(set(ids[i, :k].tolist()) & set(gt[i, :k].tolist())) / k for i in range(len(ids))]))
