"""The command-line front end (katoann-public_amd/cli/katoann.cpp, the
reference's tools/katoann.cpp interface): usage and argument errors on the CPU;
on the GPU, build / search / load print JSON lines whose keys, in order, are the
reference CLI's (checked against the reference build's own output where it is
present), the result files hold the library's search results, and the
percentile rule is sorted[min(n-1, ceil(q n)-1)]."""
import json
import os
import subprocess

import numpy as np
import pytest

import katoann
import oracle as O
from katoann import datasets, formats

CLI = os.path.join(os.path.dirname(katoann.LIB_PATH), "bin", "katoann")

BUILD_KEYS = ["vectors", "dim", "nlist", "build_s"]
SEARCH_HEAD_KEYS = ["index", "engine", "threads", "load_s", "memory_gb", "queries", "prefix_dims"]
SEARCH_KEYS = ["nprobe", "k", "survivors", "batch", "qps", "seconds_median", "seconds_min", "seconds_max"]
LOAD_KEYS = ["rate", "requests", "nprobe", "engine", "achieved_qps", "p50_ms", "p95_ms", "p99_ms", "max_ms",
             "batches", "mean_batch"]


def run(*args, check=True):
    r = subprocess.run([CLI, *map(str, args)], capture_output=True, text=True, timeout=300)
    if check and r.returncode != 0:
        raise AssertionError("katoann %s: rc %d\n%s" % (" ".join(map(str, args)), r.returncode, r.stderr))
    return r


def lines(r):
    return [json.loads(s) for s in r.stdout.splitlines() if s.strip()]


def test_usage_and_argument_errors():
    r = run(check=False)
    assert r.returncode == 2 and "usage: katoann build" in r.stderr
    r = run("serve", check=False)
    assert r.returncode == 2 and "usage:" in r.stderr
    r = run("search", "index", check=False)
    assert r.returncode == 1 and r.stderr.strip() == "katoann: unexpected argument index"
    r = run("search", "--index", check=False)
    assert r.returncode == 1 and r.stderr.strip() == "katoann: missing value for --index"
    r = run("search", "--queries", "q.fbin", check=False)
    assert r.returncode == 1 and r.stderr.strip() == "katoann: missing --index"
    r = run("build", "--data", "/nonexistent.fbin", "--out", "/tmp/x", check=False)
    assert r.returncode == 1 and r.stderr.strip() == "katoann: cannot open /nonexistent.fbin"


def test_truncated_fbin(tmp_path):
    p = str(tmp_path / "t.fbin")
    with open(p, "wb") as f:
        f.write(np.array([10, 4], np.uint32).tobytes() + np.zeros(7, np.float32).tobytes())
    r = run("build", "--data", p, "--out", str(tmp_path / "i"), check=False)
    assert r.returncode == 1 and r.stderr.strip() == "katoann: truncated " + p


@pytest.fixture(scope="module")
def cli_index(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    base, q = datasets.decaying_clusters(20000, 300, d=96, seed=4)
    files = dict(base=str(d / "base.fbin"), q=str(d / "q.fbin"), gt=str(d / "gt.bin"), prefix=str(d / "idx"))
    formats.write_fbin(files["base"], base)
    formats.write_fbin(files["q"], q)
    formats.write_gt(files["gt"], datasets.ground_truth_np(base, q, 10))
    files["build"] = lines(run("build", "--data", files["base"], "--out", files["prefix"], "--nlist", 64,
                               "--seed", 3, "--normalized"))
    return files, base, q


@pytest.mark.gpu
def test_build_line(cli_index):
    files, base, _ = cli_index
    (b,) = files["build"]
    assert list(b) == BUILD_KEYS and b["vectors"] == 20000 and b["dim"] == 96 and b["nlist"] == 64
    ix = katoann.Index()
    ix.load(files["prefix"])
    assert ix.size == 20000 and ix.normalized and ix.rotated


@pytest.mark.gpu
def test_search_lines_and_results(cli_index, tmp_path):
    files, _, q = cli_index
    out = str(tmp_path / "res")
    r = run("search", "--index", files["prefix"], "--queries", files["q"], "--gt", files["gt"], "--nprobe", "8,32",
            "--survivors", 64, "--batch", 100, "--repeat", 2, "--out", out, "--engine", "avx512")
    head, *rows = lines(r)
    assert list(head) == SEARCH_HEAD_KEYS and head["engine"] == "gfx950" and head["queries"] == 300
    assert [row["nprobe"] for row in rows] == [8, 32]
    for row in rows:
        assert list(row) == SEARCH_KEYS + ["recall"] and row["batch"] == 100 and row["survivors"] == 64
        assert row["seconds_min"] <= row["seconds_median"] <= row["seconds_max"]
    assert rows[1]["recall"] >= rows[0]["recall"] and rows[1]["recall"] > 0.9
    # the result file: u32 queries, u32 k, ids, dists; equal to the library's search
    ix = katoann.Index()
    ix.load(files["prefix"])
    ids, dists = ix.search(q, k=10, nprobe=32, survivors=64)
    raw = np.fromfile(out + ".np32.bin", np.uint32)
    assert raw[0] == 300 and raw[1] == 10
    assert np.array_equal(raw[2:2 + 3000].reshape(300, 10), ids)
    assert np.array_equal(raw[2 + 3000:].reshape(300, 10), dists.view(np.uint32))
    # --query-ids: the listed rows only
    qi = str(tmp_path / "qi.txt")
    open(qi, "w").write("5\n7\n299\n")
    head, row = lines(run("search", "--index", files["prefix"], "--queries", files["q"], "--query-ids", qi,
                          "--nprobe", 32, "--survivors", 64, "--repeat", 1, "--out", out + "_sub"))
    assert head["queries"] == 3 and list(row) == SEARCH_KEYS
    raw = np.fromfile(out + "_sub.np32.bin", np.uint32)
    assert np.array_equal(raw[2:2 + 30].reshape(3, 10), ids[[5, 7, 299]])


@pytest.mark.gpu
def test_search_argument_errors(cli_index, tmp_path):
    files, _, _ = cli_index
    bad = str(tmp_path / "bad.txt")
    open(bad, "w").write("300\n")
    for args, msg in [(("--ranking", "best"), "--ranking is prefix, high, or full"),
                      (("--selection", "any"), "--selection is threshold or local"),
                      (("--coarse", "tree"), "--coarse is batched or flat"),
                      (("--precision", "fp8"), "--precision is sq16 or bf16"),
                      (("--query-ids", bad), "query id out of range in " + bad)]:
        r = run("search", "--index", files["prefix"], "--queries", files["q"], *args, check=False)
        assert r.returncode == 1 and r.stderr.strip() == "katoann: " + msg, (args, r.stderr)


@pytest.mark.gpu
def test_load_lines(cli_index):
    files, _, _ = cli_index
    (a,) = lines(run("load", "--index", files["prefix"], "--queries", files["q"], "--rate", 3000, "--requests",
                     600, "--nprobe", 16, "--max-batch", 256))
    assert list(a) == LOAD_KEYS and a["requests"] == 600 and a["engine"] == "gfx950"
    assert a["p50_ms"] <= a["p95_ms"] <= a["p99_ms"] <= a["max_ms"]
    assert 1 <= a["mean_batch"] <= 256 and a["batches"] >= 600 / 256
    (b,) = lines(run("load", "--index", files["prefix"], "--queries", files["q"], "--rate", 3000, "--requests",
                     600, "--nprobe", 16, "--inflight", 3))
    assert list(b) == LOAD_KEYS + ["inflight"] and b["inflight"] == 3
    assert b["p50_ms"] <= b["p99_ms"] <= b["max_ms"]


@pytest.mark.gpu
@pytest.mark.skipif(not os.path.exists(O.REF_CLI), reason="reference CLI build absent (make -C oracle ref)")
def test_keys_match_reference_cli(cli_index):
    """The reference CLI (built from its sources) on the same files prints the
    same keys in the same order."""
    files, _, _ = cli_index
    args = ["--index", files["prefix"], "--queries", files["q"], "--nprobe", 16, "--repeat", 1, "--engine",
            "avx512"]
    ref = subprocess.run([O.REF_CLI, "search", *map(str, args), "--gt", files["gt"]], capture_output=True,
                         text=True, timeout=300, check=True)
    ours = run("search", *args, "--gt", files["gt"])
    assert [list(x) for x in lines(ref)] == [list(x) for x in lines(ours)]
    load = ["load", "--index", files["prefix"], "--queries", files["q"], "--rate", 2000, "--requests", 200,
            "--nprobe", 16, "--engine", "avx512"]
    ref = subprocess.run([O.REF_CLI, *map(str, load)], capture_output=True, text=True, timeout=300, check=True)
    assert [list(x) for x in lines(ref)] == [list(x) for x in lines(run(*load))]
