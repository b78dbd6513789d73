"""The CPU oracle against the reference's own outputs (golden fixtures written
by tests/golden/make_golden.py from a build of the reference's sources)."""
import hashlib

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
    ix = O.OracleIndex.from_files(prefix)
    exact_rows = 0
    for (k, nprobe, S, ranking, selection) in cases:
        tag = f"{k}_{nprobe}_{S}_{ranking}_{selection}"
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
