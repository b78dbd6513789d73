"""The CPU oracle against the reference's own outputs (golden fixtures written
by tests/golden/make_golden.py from a build of the reference's sources)."""
import hashlib
import os

import numpy as np
import pytest

import common
import oracle as O
from katoann import formats


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name", common.golden_names())
def test_load_state_matches_reference(name):
    prefix, q, ref, _ = common.golden(name)
    ix = O.OracleIndex.from_files(prefix)
    lo, inv, ssq = ix.scale()
    assert np.float32(lo) == ref["lo"] and inv == ref["inv_step"] and ssq == ref["step_sq"]
    e = ix.export()
    assert sha(e["codes"]) == str(ref["codes_sha"])
    assert sha(e["ccodes"]) == str(ref["ccodes_sha"])
    assert sha(e["norm"]) == str(ref["norm_sha"])
    assert sha(e["cnorm"]) == str(ref["cnorm_sha"])
    raw = formats.read_index(prefix)
    assert sha(O.rotate_rows(raw["vecs"], raw["rotation"])) == str(ref["rotated_sha"])


@pytest.mark.parametrize("name", common.golden_names())
def test_search_matches_reference(name):
    prefix, q, ref, cases = common.golden(name)
    ox = O.OracleIndex.from_files(prefix)
    exact_rows = 0
    for case in cases:
        k, nprobe, S, ranking, selection, mode = case
        tag = common.case_tag(case)
        ix = common.oracle_mode(ox, mode)
        r_ids, r_d = ref["ids_" + tag], ref["dists_" + tag]
        # libstdc++ tie emulation: identical to the reference on every row of kLocal
        o = ix.search(q, k=k, nprobe=nprobe, ranking=ranking, survivors=S, libstdc_ties=True)
        if selection == "local":
            assert np.array_equal(o["ids"], r_ids), tag
            assert np.array_equal(o["dists"].view(np.uint32), r_d.view(np.uint32)), tag
        # canonical ties: bit-exact away from tie rows, the reference's same() on them
        c = ix.search(q, k=k, nprobe=nprobe, ranking=ranking, survivors=S)
        tie = (c["flags"] & common.tie_mask(ranking, selection)) != 0
        n_ok, _ = common.compare(c["ids"], c["dists"], r_ids, r_d, tie)
        exact_rows += n_ok
    assert exact_rows > 0


@pytest.mark.parametrize("name", common.flat_golden_names())
def test_flat_order_matches_reference(name):
    """The per-query fp32 coarse distances order every list exactly as the
    reference's coarse_flat_ does (index.cpp:832-844). The gflat fixture's
    twin centroids are equidistant in exact arithmetic, so their order is
    decided by the fp32 summation order alone."""
    d = os.path.join(common.GOLDEN, name)
    ref = np.load(os.path.join(d, "reference_outputs.npz"))["flat_order"]
    q = formats.read_fbin(os.path.join(d, "queries.fbin"))
    D = O.OracleIndex.from_files(os.path.join(d, "idx")).flat_dists(q)
    nl = D.shape[1]
    order = np.stack([np.lexsort((np.arange(nl), D[i])) for i in range(len(q))])
    assert np.array_equal(order, ref)


def test_flat_dists_fp64_bound():
    """The fp32 distances stay within a few ulps of the fp64 sum."""
    d = os.path.join(common.GOLDEN, "g100")
    q = formats.read_fbin(os.path.join(d, "queries.fbin"))
    raw = formats.read_index(os.path.join(d, "idx"))
    D = O.OracleIndex.from_files(os.path.join(d, "idx")).flat_dists(q)
    e = ((q[:, None, :].astype(np.float64) - raw["centroids"][None].astype(np.float64)) ** 2).sum(-1)
    assert np.abs(D - e).max() <= 1e-5 * e.max()
