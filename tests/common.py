"""Shared helpers of the tests: fixtures on disk, comparison rules."""
from __future__ import annotations

import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def same(a_ids, a_d, b_ids, b_d) -> bool:
    """The reference's own equality (tests/test_index.cpp:43-52): identical
    distances at every rank; ids equal wherever the distance is below the
    query's k-th distance."""
    if not np.array_equal(a_d.view(np.uint32), b_d.view(np.uint32)):
        return False
    last = a_d[:, -1:]
    mask = a_d < last
    return bool(np.array_equal(np.where(mask, a_ids, 0), np.where(mask, b_ids, 0)))


def compare(ids, dists, ref_ids, ref_dists, tie_rows):
    """Bit-exact on rows without ties; the reference's same() on tie rows.
    Returns (n_exact_rows_checked, n_tie_rows)."""
    tie = np.asarray(tie_rows, bool)
    ok = ~tie
    assert np.array_equal(ids[ok], ref_ids[ok]), "ids differ on %d tie-free rows" % int(
        (ids[ok] != ref_ids[ok]).any(1).sum())
    assert np.array_equal(dists[ok].view(np.uint32), ref_dists[ok].view(np.uint32)), "dists differ"
    if tie.any():
        assert same(ids[tie], dists[tie], ref_ids[tie], ref_dists[tie])
    return int(ok.sum()), int(tie.sum())


def tie_mask(ranking, selection):
    """Oracle flag bits that mark a query whose reference result may differ
    from the canonical one: a coarse tie (1; never set for the per-query coarse
    search, whose order is fixed), a survivor tie under threshold selection
    (2), a tie at the k-th exact distance of full ranking (4)."""
    if ranking == "full":
        return 5
    return 1 if selection == "local" else 3


def _golden_dirs():
    return sorted(n for n in os.listdir(GOLDEN) if os.path.exists(os.path.join(GOLDEN, n, "reference_outputs.npz")))


def _golden_keys(name):
    return np.load(os.path.join(GOLDEN, name, "reference_outputs.npz")).files


def golden_names():
    """Fixtures with the reference's load state and search outputs."""
    return [n for n in _golden_dirs() if "codes_sha" in _golden_keys(n)]


def flat_golden_names():
    """Fixtures with the reference's per-query coarse order of every list."""
    return [n for n in _golden_dirs() if "flat_order" in _golden_keys(n)]


_WIDE = {}


def wide_files(name):
    """A regenerated fixture's index files (tests/golden/wide_fixture.py),
    written once per process under the temporary directory: (prefix, queries)."""
    if name not in _WIDE:
        import atexit
        import shutil
        import sys
        import tempfile
        sys.path.insert(0, GOLDEN)
        import wide_fixture
        d = tempfile.mkdtemp(prefix="katoann_%s_" % name)
        atexit.register(shutil.rmtree, d, True)
        prefix = os.path.join(d, "idx")
        q = wide_fixture.write(name, prefix)
        ref = np.load(os.path.join(GOLDEN, name, "reference_outputs.npz"))
        got = wide_fixture.files_sha(prefix)
        assert got == str(ref["files_sha"]), "regenerated %s differs from the files the reference ran on" % name
        _WIDE[name] = (prefix, q)
    return _WIDE[name]


def golden(name):
    """(index prefix, queries, reference outputs, cases); a case is
    (k, nprobe, survivors, ranking, selection, mode), mode "batched", "flat"
    (the per-query coarse search, LoadOptions::batched_coarse = false) or
    "bf16" (LoadOptions::precision = kBf16)."""
    from katoann import formats
    d = os.path.join(GOLDEN, name)
    ref = np.load(os.path.join(d, "reference_outputs.npz"))
    if os.path.exists(os.path.join(d, "queries.fbin")):
        q = formats.read_fbin(os.path.join(d, "queries.fbin"))
        prefix = os.path.join(d, "idx")
    else:
        prefix, q = wide_files(name)
    cases = []
    for key in ref.files:
        if key.startswith("ids_"):
            parts = key[4:].split("_")
            k, nprobe, S, ranking, selection = parts[:5]
            coarse = parts[5] if len(parts) > 5 else "batched"
            cases.append((int(k), int(nprobe), int(S), ranking, selection, coarse))
    return prefix, q, ref, sorted(cases)


def oracle_mode(ox, mode):
    """The oracle index in a case's mode: batched coarse search, per-query
    ("flat") coarse search, or bf16 precision."""
    ox.set_bf16(mode == "bf16")
    ox.set_flat(mode in ("flat", "bf16"))
    return ox


def case_tag(case):
    k, nprobe, S, ranking, selection, coarse = case
    return f"{k}_{nprobe}_{S}_{ranking}_{selection}" + ("" if coarse == "batched" else "_" + coarse)


def small_index(tmpdir, n=20000, nq=300, d=128, nlist=64, centres=64, seed=7, normalized=True, iters=5):
    """A seeded mixture index written with the CPU builder; returns (prefix, base, queries)."""
    from katoann import builder, datasets
    base, q = datasets.mixture(n, nq, d=d, centres=centres, sigma=0.3, seed=seed)
    prefix = os.path.join(str(tmpdir), "idx_%d_%d_%d" % (n, d, nlist))
    builder.build_numpy(base, prefix, nlist=nlist, kmeans_iters=iters, seed=seed, normalized=normalized)
    return prefix, base, q
