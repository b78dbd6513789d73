"""Benchmark: QPS at recall@10 >= 0.99 / 0.999 of the batched IVF search on one
MI355X, with the ranking pass's HBM roofline and the reference CPU path beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1000000] [--dim 1024]

Workload (default): 1,000,000 x 1,024 float32, Gaussian with covariance
eigenvalues i^-1 in a seeded random orthogonal basis, L2-normalised; 4,096
queries of the same law; nlist 4,096, prefix 256 dims, survivors 32, k 10,
nprobe swept 16..256, batch 4,096 (the GPU drop-in configuration). Data and
index are built on the GPU at start-up (synthetic, seeded), the index loaded
through the C ABI from device memory, queries resident in HBM.

A step = one search of one 4,096-query batch through the whole pipeline
(rotation, quantisation, coarse MFMA, probe selection, inversion, ranking,
refinement, top-k). The timed loop keeps --inflight batches (default 3) in
flight, each host thread searching on its own context and stream, as a
serving loop does; stage times and the ranking roofline come from a separate
pass of sequential eager searches. value = queries/s at the best swept point
whose recall@10 reaches 0.99.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "katoann-public_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def recall_at(ids, gt, k=10):
    hit = 0
    for i in range(ids.shape[0]):
        hit += len(set(ids[i, :k].tolist()) & set(gt[i, :k].tolist()))
    return hit / (ids.shape[0] * k)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="spectrum", choices=("spectrum", "zipf"),
                    help="spectrum: 1M x 1024 i^-1 spectrum, nlist 4096 (default); zipf: 10M x 1024 mixture of "
                         "16,384 components with Zipf(1.1) weights, nlist 16,384")
    ap.add_argument("--n", type=int, default=None)
    ap.add_argument("--dim", type=int, default=1024)
    ap.add_argument("--nq", type=int, default=4096)
    ap.add_argument("--nlist", type=int, default=None)
    ap.add_argument("--alpha", type=float, default=1.0)
    ap.add_argument("--nprobe", default=None)
    ap.add_argument("--survivors", default="32,128,256")
    ap.add_argument("--extra-points", default=None,
                    help="nprobe:survivors pairs added to the grid (default: 2048:512 on the 1M workload, where "
                         "recall 0.999 needs 512 survivors)")
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--kmeans-iters", type=int, default=10)
    ap.add_argument("--rank-rows", type=int, default=0)
    ap.add_argument("--band-shift", type=int, default=0, help="probe-rank bands grow by 2**shift (default: library's)")
    ap.add_argument("--rank-kernel", default="", help="kernel,variant,queries_per_item (default: library's)")
    ap.add_argument("--seed-rows", type=int, default=-1, help="rows of the seed band (0 = none; default: library's)")
    ap.add_argument("--inflight", type=int, default=3,
                    help="batches in flight in the timed loop (each on its own search context and stream)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-queries", type=int, default=1024)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-train", type=int, default=50000, help="training queries of the learned nprobe")
    ap.add_argument("--learned-nmax", type=int, default=1024)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=6.0)
    ap.add_argument("--serve-load", type=float, default=0.5, help="offered queries/s as a fraction of batch QPS")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the sweep, copy this many MiB device to device 3 times (a plain copy kernel of "
                         "known traffic, for scaling profiler byte counters)")
    args = ap.parse_args()
    zipf = args.workload == "zipf"
    if args.n is None:
        args.n = 10_000_000 if zipf else 1_000_000
    if args.nlist is None:
        args.nlist = 16384 if zipf else 4096
    if args.nprobe is None:
        args.nprobe = "32,64,128,256,512" if zipf else "16,32,64,128,256,512,1024,2048"
    if zipf:  # the learned leg's workload and the reference CPU run (41 GB of files) are not this one's
        args.no_learned = True
        args.no_cpu_baseline = True
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (--gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    import katoann
    from katoann import builder, datasets, formats

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    t0 = time.time()
    if zipf:
        base, queries = datasets.zipf_mixture_torch(args.n, args.nq, args.dim, comps=16384, s=1.1, seed=3, device=dev)
    else:
        base, queries = datasets.spectrum_torch(args.n, args.nq, args.dim, alpha=args.alpha, seed=2, device=dev)
    torch.cuda.synchronize()
    log("data %.1fs" % (time.time() - t0))
    t0 = time.time()
    b = builder.build_torch(base, None, nlist=args.nlist, train_size=256000, kmeans_iters=args.kmeans_iters,
                            seed=1, normalized=True)
    torch.cuda.synchronize()
    log("index build (GPU k-means + PCA) %.1fs" % (time.time() - t0))
    order = b["order"]
    sizes = b["sizes"].cpu().numpy().astype(np.uint32)
    vecs_sorted = base.index_select(0, order).contiguous()
    pd = formats.default_prefix_dims(args.dim)
    ix = katoann.Index()
    t0 = time.time()
    ix.load_arrays(args.dim, args.nlist, args.n, True, b["centroids"].cpu().numpy(), sizes,
                   order.cpu().numpy().astype(np.uint32), None, b["rotation"].cpu().numpy(), pd,
                   vectors_device_ptr=vecs_sorted.data_ptr())
    torch.cuda.synchronize()
    del vecs_sorted
    torch.cuda.empty_cache()
    log("load (GPU rotation + quantisation) %.1fs, list sizes min %d median %d max %d" % (
        time.time() - t0, sizes.min(), int(np.median(sizes)), sizes.max()))
    if args.rank_rows:
        ix.set_rank_rows(args.rank_rows)
    if args.band_shift:
        ix.set_band_factor(args.band_shift)
    if args.seed_rows >= 0:
        ix.set_seed_rows(args.seed_rows)
    if args.rank_kernel:
        ix.set_rank_kernel(*[int(x) for x in args.rank_kernel.split(",")])
    else:
        ix.set_autotune(True)  # ranking tiles timed on the first search of each shape (untimed warm-up)
    t0 = time.time()
    gt = datasets.ground_truth_torch(base, queries, 10)
    log("ground truth %.1fs" % (time.time() - t0))

    nq, k = args.nq, args.k
    ids_d = torch.empty((nq, k), dtype=torch.int32, device=dev)
    dists_d = torch.empty((nq, k), dtype=torch.float32, device=dev)
    qptr = queries.data_ptr()

    def step(np_, S):
        ix.search_device(qptr, nq, ids_d.data_ptr(), dists_d.data_ptr(), k=k, nprobe=np_, survivors=S)

    sweep = []
    grid = [(int(a), int(b)) for b in args.survivors.split(",") for a in args.nprobe.split(",")
            if int(a) <= args.nlist]
    extra = args.extra_points if args.extra_points is not None else ("" if zipf else "2048:512")
    for pt in filter(None, extra.split(",")):
        xp, xs = (int(x) for x in pt.split(":"))
        if xp <= args.nlist and (xp, xs) not in grid:
            grid.append((xp, xs))
    # with several batches in flight, each host thread searches on its own
    # context (stream, buffers, captured graph) into its own outputs
    from concurrent.futures import ThreadPoolExecutor
    pool = ThreadPoolExecutor(args.inflight) if args.inflight > 1 else None
    outs = [(torch.empty((nq, k), dtype=torch.int32, device=dev), torch.empty((nq, k), dtype=torch.float32, device=dev))
            for _ in range(args.inflight)]

    def steps_on(j, n, np_, S):
        for _ in range(n):
            ix.search_device(qptr, nq, outs[j][0].data_ptr(), outs[j][1].data_ptr(), k=k, nprobe=np_, survivors=S)

    def timed(n, np_, S):
        if pool is None:
            for _ in range(n):
                step(np_, S)  # graph replays, no stage events
            return
        shares = [n // args.inflight + (1 if j < n % args.inflight else 0) for j in range(args.inflight)]
        list(pool.map(lambda j: steps_on(j, shares[j], np_, S), range(args.inflight)))

    for np_, S in grid:
        for _ in range(args.warmup):
            step(np_, S)
        timed(args.warmup * args.inflight if pool else 0, np_, S)  # the other contexts' graphs
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        timed(args.steps, np_, S)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        # the same steps again, eagerly, with HIP events around every stage and
        # every ranking launch on the search stream (roofline durations)
        ix.set_timing(True)
        stage_sum = np.zeros(len(katoann.STAGES))
        for _ in range(args.steps):
            step(np_, S)
            stage_sum += np.array(list(ix.stage_times().values()))
        ix.set_timing(False)
        ids = ids_d.cpu().numpy().view(np.uint32)
        rec = recall_at(ids, gt, 10)
        # algorithmic bytes of the ranking pass: every touched list once, P + 4 bytes per row
        _, _, ex = ix.search_ex(queries.cpu().numpy(), k=k, nprobe=np_, survivors=S, normalize=False,
                                want=("probes",))
        touched = np.unique(ex["probes"])
        rank_bytes = float(sizes[touched].astype(np.float64).sum() * (pd + 4))
        # algorithmic int8 work of the ranking pass: 2 P ops per (query, candidate row)
        rank_ops = float(sizes[ex["probes"]].astype(np.float64).sum() * 2 * pd)
        stages = dict(zip(katoann.STAGES, (stage_sum / args.steps).tolist()))
        work = ix.last_work()
        rank_ms = stages["rank"]
        row = dict(nprobe=np_, survivors=S, ms_per_step=ms, qps=nq / (ms / 1e3), recall=rec, stages_ms=stages,
                   tile=ix.tile_config(nq, k=k, nprobe=np_, survivors=S),
                   work=work, rank_bytes=rank_bytes, rank_gbs=rank_bytes / (rank_ms * 1e-3) / 1e9 if rank_ms > 0 else 0.0,
                   rank_ops=rank_ops, rank_tops=rank_ops / (rank_ms * 1e-3) / 1e12 if rank_ms > 0 else 0.0)
        sweep.append(row)
        log(json.dumps(row))

    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    def best(target):
        ok = [r for r in sweep if r["recall"] >= target]
        return max(ok, key=lambda r: r["qps"]) if ok else None

    b99, b999 = best(0.99), best(0.999)
    head = b99 or max(sweep, key=lambda r: r["recall"])

    learned_out, learned_model = None, None
    if not args.no_learned and b99:
        try:
            learned_out, learned_model = learned_leg(ix, base, queries, gt, b, sizes, args, b99, ids_d, dists_d)
        except Exception as e:  # reported, never silently replaced
            log("learned leg failed: %r" % (e,))

    ix.set_autotune(False)  # serving runs with the tiles already measured
    serve_out = None
    if not args.no_serve:
        try:
            serve_out = serve_leg(ix, queries, args, head, b999 or max(sweep, key=lambda r: r["recall"]),
                                  learned_model, learned_out)
        except Exception as e:  # reported, never silently replaced
            log("serving leg failed: %r" % (e,))

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(base, queries, b, sizes, pd, args, head["nprobe"], head["survivors"])
        except Exception as e:  # reported, never silently replaced
            log("cpu baseline failed: %r" % (e,))
    out = {
        "metric": "QPS at recall@10>=0.99 (batch 4096, 1 GPU)",
        "value": head["qps"] if b99 else 0.0,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/int32 (16-bit codes, exact integer distances)",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": ("%dx%d mixture of 16384 components, Zipf(1.1) weights, i^-1 spectrum" % (
            args.n, args.dim) if zipf else "%dx%d i^-%g spectrum" % (args.n, args.dim, args.alpha)) +
            ", nlist %d, prefix %d, k %d, batch %d" % (args.nlist, pd, k, nq),
            "nprobe": head["nprobe"], "survivors": head["survivors"], "recall_at_10": head["recall"],
            "batches_in_flight": args.inflight},
        "qps_at_0999": b999["qps"] if b999 else None,
        "at_0999": {"nprobe": b999["nprobe"], "survivors": b999["survivors"]} if b999 else None,
        "roofline": dict(roofline(head), traffic=profiled_traffic(head, args)),
        "stages_ms": head["stages_ms"],
        "sweep": [{kk: r[kk] for kk in ("nprobe", "survivors", "qps", "recall", "ms_per_step", "rank_gbs", "rank_tops", "tile")}
                  for r in sweep],
        "learned_nprobe": learned_out,
        "serving": serve_out,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


MFMA_I8_PEAK_TOPS = 5000.0  # dense int8 MFMA, MI355X


def profiled_traffic(r, args):
    """HBM bytes per search of the ranking pass from a committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (profiles/*.md) taken on this workload,
    search shape and ranking tile."""
    t = r.get("tile") or {}
    want = "%s %dx%d nlist %d nprobe %d survivors %d batch %d tile %d,%d,%d shift %d seed %d" % (
        args.workload, args.n, args.dim, args.nlist, r["nprobe"], r["survivors"], args.nq, t.get("kernel", -1),
        t.get("variant", -1), t.get("queries_per_item", -1), t.get("log2_band_factor", -1), t.get("seed_rows", -1))
    pdir = os.path.join(ROOT, "profiles")
    for fn in sorted(os.listdir(pdir)) if os.path.isdir(pdir) else []:
        if not fn.endswith(".md"):
            continue
        lines = dict(l.split(": ", 1) for l in open(os.path.join(pdir, fn)).read().splitlines()
                     if l.startswith("traffic_"))
        if lines.get("traffic_config", "").strip() == want:
            return float(lines["traffic_bytes_per_search"])
    return None


def roofline(r):
    """Roofline of the ranking pass (k_rank_wave, all probe-rank bands of one
    search): algorithmic bytes = rows of the touched lists x (P + 4), once per
    search; algorithmic int8 ops = 2 P per (query, candidate row). The bound is
    the one its arithmetic intensity puts it under (ridge 5000 TOP/s / 8000
    GB/s = 625 ops per byte)."""
    inten = r["rank_ops"] / r["rank_bytes"] if r["rank_bytes"] else 0.0
    mfma = inten > MFMA_I8_PEAK_TOPS * 1e12 / (HBM_PEAK_GBS * 1e9)
    ach, peak, unit = (r["rank_tops"], MFMA_I8_PEAK_TOPS, "TOP/s") if mfma else (r["rank_gbs"], HBM_PEAK_GBS, "GB/s")
    return {"bound": "mfma" if mfma else "hbm", "kernel": "k_rank_wave (all probe-rank bands of one search)",
            "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak, "traffic": None,
            "intensity_ops_per_byte": inten, "hbm_gbs": r["rank_gbs"], "hbm_frac": r["rank_gbs"] / HBM_PEAK_GBS,
            "int8_tops": r["rank_tops"], "int8_frac": r["rank_tops"] / MFMA_I8_PEAK_TOPS,
            "bytes_per_search": r["rank_bytes"], "ops_per_search": r["rank_ops"],
            "ms_per_search": r["stages_ms"]["rank"], "launches_per_search": r["work"]["bands"],
            "bytes_rule": "sum over touched lists of rows x (prefix bytes + 4-byte norm), once per search",
            "ops_rule": "2 x prefix bytes per (query, row of a probed list)",
            "timing": "HIP events on the search stream around every ranking launch, averaged over the steps"}


def learned_leg(ix, base, queries, gt, b, sizes, args, plain, ids_d, dists_d):
    """Per-query nprobe from the MLP trained on the GPU (untimed): training
    queries and validation queries are drawn apart from the benchmark queries;
    the margin is the smallest reaching recall@10 0.99 on the validation set;
    then the benchmark batch is timed like the fixed-nprobe sweep."""
    import torch
    from katoann import datasets, learned
    dev = queries.device
    n, d, nlist, nq, S = base.shape[0], base.shape[1], args.nlist, args.nq, plain["survivors"]
    t0 = time.time()
    list_of_id = torch.empty(n, dtype=torch.int32, device=dev)
    list_of_id[b["order"].long()] = torch.repeat_interleave(
        torch.arange(nlist, dtype=torch.int32, device=dev), torch.from_numpy(sizes.astype(np.int64)).to(dev))
    trq = datasets.spectrum_queries_torch(args.learned_train, d, alpha=args.alpha, seed=2, draw_seed=11, device=dev)
    gtr = torch.from_numpy(datasets.ground_truth_torch(base, trq, 10).astype(np.int64)).to(dev)
    model, labels, hist = learned.train(ix, trq, list_of_id, gtr, epochs=20, seed=0)
    del trq, gtr
    val = datasets.spectrum_queries_torch(nq, d, alpha=args.alpha, seed=2, draw_seed=12, device=dev)
    gval = datasets.ground_truth_torch(base, val, 10)
    torch.cuda.synchronize()
    log("learned: training set + fit %.1fs, loss %.3f -> %.3f" % (time.time() - t0, hist[0], hist[-1]))
    nmax = min(args.learned_nmax, nlist)
    used = torch.empty(nq, dtype=torch.int32, device=dev)

    def run(qt, margin):
        ix.search_device(qt.data_ptr(), nq, ids_d.data_ptr(), dists_d.data_ptr(), k=args.k, nprobe=nmax,
                         survivors=S, model=model.struct(margin), npq_out_ptr=used.data_ptr())

    margin = None
    for mg in np.arange(-0.5, 4.01, 0.125):
        run(val, float(mg))
        if recall_at(ids_d.cpu().numpy().view(np.uint32), gval, 10) >= 0.99:
            margin = float(mg)
            break
    if margin is None:
        raise RuntimeError("no margin reaches recall 0.99 on the validation queries")
    # timed like the plain sweep: graph replays with the same batches in flight,
    # each host thread on its own context and outputs; stage times from a
    # separate eager pass
    mstruct = model.struct(margin)
    outs = [(torch.empty_like(ids_d), torch.empty_like(dists_d), torch.empty_like(used))
            for _ in range(args.inflight)]

    def run_on(j, n):
        for _ in range(n):
            ix.search_device(queries.data_ptr(), nq, outs[j][0].data_ptr(), outs[j][1].data_ptr(), k=args.k,
                             nprobe=nmax, survivors=S, model=mstruct, npq_out_ptr=outs[j][2].data_ptr())

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(args.inflight) as pool:
        list(pool.map(lambda j: run_on(j, args.warmup), range(args.inflight)))
        torch.cuda.synchronize()
        shares = [args.steps // args.inflight + (1 if j < args.steps % args.inflight else 0)
                  for j in range(args.inflight)]
        t0 = time.perf_counter()
        list(pool.map(lambda j: run_on(j, shares[j]), range(args.inflight)))
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
    ix.set_timing(True)
    stage_sum = np.zeros(len(__import__("katoann").STAGES))
    for _ in range(args.steps):
        run(queries, margin)
        stage_sum += np.array(list(ix.stage_times().values()))
    ix.set_timing(False)
    rec = recall_at(ids_d.cpu().numpy().view(np.uint32), gt, 10)
    mean_np = float(used.float().mean().item())
    import katoann
    out = {"qps": nq / (ms / 1e3), "ms_per_step": ms, "recall": rec, "margin": margin, "survivors": S,
           "nprobe_max": nmax, "mean_nprobe": mean_np, "plain_nprobe": plain["nprobe"],
           "plain_qps": plain["qps"], "speedup_vs_plain": (nq / (ms / 1e3)) / plain["qps"],
           "stages_ms": dict(zip(katoann.STAGES, (stage_sum / args.steps).tolist())),
           "training": "%d queries of the same law (own seed), coverage labels from exact top-10, 20 epochs "
                       "Adam, bf16 MFMA; margin calibrated on %d further queries" % (args.learned_train, nq)}
    log("learned: " + json.dumps(out))
    return out, (model, model.struct(margin), nmax)


def serve_leg(ix, queries, args, at99, at999, learned_model, learned_out):
    """Continuous batching under Poisson arrivals: request sizes floor(exp(U(0,
    ln 4096))) queries drawn from the benchmark queries, k in {10, 100}, target
    recall@10 0.99 (learned nprobe when trained, else the fixed nprobe reaching
    0.99) or 0.999 (the fixed nprobe reaching it, else the sweep's highest
    recall); offered load = serve_load x the batch QPS at 0.99. Latency runs
    from submission (at the scheduled arrival) to results ready."""
    import threading
    import katoann
    host_q = queries.cpu().numpy()
    rng = np.random.default_rng(42)
    mean_size = (4096 - 1) / np.log(4096)
    rate = args.serve_load * at99["qps"] / mean_size  # requests per second
    n_req = max(1, int(rate * args.serve_seconds))
    gaps = rng.exponential(1.0 / rate, n_req)
    sizes = np.minimum(4096, np.floor(np.exp(rng.uniform(0, np.log(4096), n_req)))).astype(int)
    ks = rng.choice([10, 100], n_req)
    hi = rng.random(n_req) < 0.5
    starts = rng.integers(0, host_q.shape[0] - sizes + 1)
    tickets = [None] * n_req
    lag = [0.0]
    srv = katoann.Server(ix, max_inflight=4, max_batch=4096)

    def params(i):
        a = at999 if hi[i] else at99
        return dict(k=int(ks[i]), nprobe=a["nprobe"], survivors=max(a["survivors"], int(ks[i])))

    def producer():
        t0 = time.perf_counter()
        t = 0.0
        for i in range(n_req):
            t += gaps[i]
            now = time.perf_counter() - t0
            if now < t:
                time.sleep(t - now)
            else:
                lag[0] = max(lag[0], now - t)
            p = params(i)
            model = None
            if not hi[i] and learned_model is not None:
                model = learned_model[1]
                p["nprobe"] = learned_model[2]
                p["survivors"] = max(learned_out["survivors"], p["k"])
            tickets[i] = srv.submit(host_q[starts[i]:starts[i] + sizes[i]], model=model, normalize=False, **p)

    th = threading.Thread(target=producer)
    t0 = time.perf_counter()
    th.start()
    th.join()
    srv.drain()
    wall = time.perf_counter() - t0
    for t in tickets:
        srv.wait(t)
    st = srv.stats()
    srv.close()
    out = {"requests": n_req, "queries": int(sizes.sum()), "offered_qps": float(sizes.sum() / gaps.sum()),
           "achieved_qps": st["achieved_qps"], "p50_ms": st["p50_ms"], "p95_ms": st["p95_ms"],
           "p99_ms": st["p99_ms"], "max_ms": st["max_ms"], "mean_batch": st["mean_batch"],
           "batches": st["batches"], "peak_inflight": st["peak_inflight"], "peak_reserved_bytes": st["peak_reserved"],
           "budget_bytes": st["budget"], "wall_s": wall, "max_submit_lag_ms": lag[0] * 1e3,
           "mix": "sizes floor(exp(U(0, ln 4096))), k {10, 100}, half at target 0.99 (%s), half at 0.999 "
                  "(nprobe %d, survivors %d)" % ("learned nprobe" if learned_model is not None else
                                                 "nprobe %d" % at99["nprobe"], at999["nprobe"], at999["survivors"])}
    log("serving: " + json.dumps(out))
    return out


def cpu_baseline(base, queries, b, sizes, pd, args, nprobe, survivors):
    """The reference's own CPU search (oracle/_ref/katoann, AVX-512 VNNI engine
    where AMX is absent) on this index and a sample of the same queries, timed by
    its CLI (one warm-up pass, median of 3) on this host's cores."""
    import oracle as O
    from katoann import formats
    if not os.path.exists(O.REF_CLI):
        raise RuntimeError("reference build missing (oracle/_ref/katoann)")
    nq = min(args.cpu_queries, queries.shape[0])
    threads = len(os.sched_getaffinity(0))
    with tempfile.TemporaryDirectory(dir="/tmp") as d:
        prefix = os.path.join(d, "cpu")
        order = b["order"].cpu().numpy().astype(np.uint32)
        bounds = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))]).astype(np.int64)
        host = base.cpu().numpy()
        lists = [order[bounds[c]:bounds[c + 1]] for c in range(len(sizes))]
        t0 = time.time()
        formats.write_index(prefix, b["centroids"].cpu().numpy(), lists, [host[i] for i in lists], host.shape[0],
                            True, b["rotation"].cpu().numpy(), pd)
        formats.write_fbin(os.path.join(d, "q.fbin"), queries[:nq].cpu().numpy())
        log("cpu baseline: files written %.1fs" % (time.time() - t0))
        cmd = [O.REF_CLI, "search", "--index", prefix, "--queries", os.path.join(d, "q.fbin"), "--k", str(args.k),
               "--nprobe", str(nprobe), "--survivors", str(survivors), "--batch", "4096", "--engine", "avx512",
               "--threads", str(threads)]
        t0 = time.time()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=1200)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-500:])
        lines = [json.loads(x) for x in r.stdout.splitlines() if x.startswith("{")]
        res = [x for x in lines if "qps" in x][0]
        hdr = lines[0]
        log("cpu baseline: %s (wall %.1fs incl. load %.1fs)" % (json.dumps(res), time.time() - t0, hdr["load_s"]))
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        model = "unknown"
    return {"value": res["qps"], "unit": "queries/s", "cores": threads, "kind": "reference",
            "sample": "%d of the %d queries, nprobe %d, survivors %d, batch %d, median of 3 after a warm-up "
                      "(reference CLI, AVX-512 VNNI engine); %s" % (nq, queries.shape[0], nprobe, survivors,
                                                                     min(nq, 4096), model),
            "load_s": hdr["load_s"]}


if __name__ == "__main__":
    main()
