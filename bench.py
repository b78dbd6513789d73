"""Benchmark: QPS at recall@10 >= 0.99 / 0.999 of the batched IVF search on one
MI355X, plain and learned nprobe, p50/p99 of the serving loop, the ranking
pass's roofline, and the reference's CPU search beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload zipf|spectrum|s3072] [--extra LIST]

Headline workload (default, --workload zipf): 10,000,000 x 1,024 float32 from a
16,384-component Gaussian mixture with Zipf(1.1) component weights (means and
noise with covariance eigenvalues i^-1 in one seeded random orthogonal basis),
L2-normalised; 4,096-query batches of the same law; nlist 16,384, prefix 256
dims, k 10, nprobe swept 32..512 with survivors 32 / 128 / 256; persistent
ranking kernels replayed from hipGraphs, autotuned tiles. Extra keys (--extra,
default spectrum,s3072): the 1,000,000 x 1,024 i^-1-spectrum drop-in workload
(nlist 4,096) and the 990,000 x 3,072 i^-0.8-spectrum workload (nlist 4,096,
prefix 768) with its learned nprobe trained on 200,000 queries.

Data and index are built on the GPU at start-up (synthetic, seeded), the index
loaded through the C ABI from device memory, queries resident in HBM. A step =
one search of one 4,096-query batch through the whole pipeline (rotation,
quantisation, coarse MFMA, probe selection, inversion, ranking, refinement,
top-k). The timed loop keeps --inflight batches (default 3) in flight, each host
thread searching on its own context and stream, as the serving loop does; stage
times and the ranking roofline come from a separate pass of sequential eager
searches. value = queries/s at the best swept point whose recall@10 reaches
0.99 on the headline workload.
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
MFMA_I8_PEAK_TOPS = 5000.0  # dense int8 MFMA, MI355X


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def recall_at(ids, gt, k=10):
    hit = 0
    for i in range(ids.shape[0]):
        hit += len(set(ids[i, :k].tolist()) & set(gt[i, :k].tolist()))
    return hit / (ids.shape[0] * k)


# ---------------------------------------------------------------- workloads

def workload(name, dev):
    """(base, queries) on the GPU, the index's nlist, a generator of more
    queries of the same law (nq, draw_seed), its description and sweep."""
    from katoann import datasets
    if name == "zipf":
        base, q = datasets.zipf_mixture_torch(10_000_000, 4096, 1024, comps=16384, s=1.1, seed=3, device=dev)
        more = lambda nq, ds: datasets.zipf_queries_torch(nq, 1024, comps=16384, s=1.1, seed=3, draw_seed=ds,
                                                          device=dev)
        return dict(base=base, q=q, nlist=16384, more=more, nprobe="32,64,128,256,512", survivors="32,128,256",
                    extra_points="", learned_train=50000,
                    label="10000000x1024 mixture of 16384 components, Zipf(1.1) weights, i^-1 spectrum")
    if name == "spectrum":
        base, q = datasets.spectrum_torch(1_000_000, 4096, 1024, alpha=1.0, seed=2, device=dev)
        more = lambda nq, ds: datasets.spectrum_queries_torch(nq, 1024, alpha=1.0, seed=2, draw_seed=ds, device=dev)
        return dict(base=base, q=q, nlist=4096, more=more, nprobe="64,256,1024", survivors="32,128",
                    extra_points="2048:512", learned_train=50000, label="1000000x1024 i^-1 spectrum")
    if name == "s3072":
        base, q = datasets.spectrum_torch(990_000, 4096, 3072, alpha=0.8, seed=4, device=dev)
        more = lambda nq, ds: datasets.spectrum_queries_torch(nq, 3072, alpha=0.8, seed=4, draw_seed=ds, device=dev)
        return dict(base=base, q=q, nlist=4096, more=more, nprobe="64,256,1024,2048", survivors="128,512,2048",
                    extra_points="", learned_train=200000, learned_nmax=4096, label="990000x3072 i^-0.8 spectrum")
    raise SystemExit("unknown workload " + name)


# ---------------------------------------------------------------- one workload

def run_workload(name, args, dev, headline):
    import torch
    import katoann
    from katoann import builder, datasets, formats
    W = workload(name, dev)
    base, queries = W["base"], W["q"]
    n, d = base.shape
    torch.cuda.synchronize()
    t0 = time.time()
    b = builder.build_torch(base, None, nlist=W["nlist"], train_size=256000, kmeans_iters=args.kmeans_iters,
                            seed=1, normalized=True)
    torch.cuda.synchronize()
    log("%s: index build %.1fs" % (name, time.time() - t0))
    order = b["order"]
    sizes = b["sizes"].cpu().numpy().astype(np.uint32)
    vecs_sorted = base.index_select(0, order).contiguous()
    pd = formats.default_prefix_dims(d)
    ix = katoann.Index()
    t0 = time.time()
    ix.load_arrays(d, W["nlist"], n, True, b["centroids"].cpu().numpy(), sizes, order.cpu().numpy().astype(np.uint32),
                   None, b["rotation"].cpu().numpy(), pd, vectors_device_ptr=vecs_sorted.data_ptr())
    torch.cuda.synchronize()
    del vecs_sorted
    torch.cuda.empty_cache()
    log("%s: load %.1fs, list sizes min %d median %d max %d" % (name, time.time() - t0, sizes.min(),
                                                                int(np.median(sizes)), sizes.max()))
    ix.set_autotune(True)  # ranking tiles timed on the first search of each shape (untimed warm-up)
    t0 = time.time()
    gt = datasets.ground_truth_torch(base, queries, 10)
    log("%s: ground truth %.1fs" % (name, time.time() - t0))
    nq, k = queries.shape[0], args.k
    grid = [(int(a), int(s)) for s in W["survivors"].split(",") for a in W["nprobe"].split(",")
            if int(a) <= W["nlist"]]
    for pt in filter(None, W["extra_points"].split(",")):
        xp, xs = (int(x) for x in pt.split(":"))
        if xp <= W["nlist"] and (xp, xs) not in grid:
            grid.append((xp, xs))
    if args.profile_point:  # one point, with the tile pinned (profiler runs)
        grid = [tuple(int(x) for x in args.profile_point.split(":"))]
        if args.tile:
            kern, var, qc, shift, seed, join = (tuple(int(x) for x in args.tile.split(",")) + (0,))[:6]
            ix.set_autotune(False)
            ix.set_rank_kernel(kern, var, qc)
            ix.set_band_factor(shift)
            ix.set_seed_rows(seed)
            ix.set_seed_join(join)
    sweep, outs = sweep_points(ix, queries, gt, grid, sizes, pd, args)
    b99, b999 = best(sweep, 0.99), best(sweep, 0.999)
    head = b99 or max(sweep, key=lambda r: r["recall"])
    res = dict(workload=W["label"] + ", nlist %d, prefix %d, k %d, batch %d" % (W["nlist"], pd, k, nq),
               qps_at_099=b99["qps"] if b99 else None, qps_at_0999=b999["qps"] if b999 else None,
               at_099=point(b99), at_0999=point(b999), head=head, b999=b999,
               sweep=[{kk: r[kk] for kk in ("nprobe", "survivors", "qps", "recall", "ms_per_step", "rank_gbs",
                                            "rank_tops", "tile")} for r in sweep])
    res["roofline"] = roofline_pass(ix, queries, head, sizes, pd, d, W["nlist"], name, args)
    if args.profile_point:
        ix.close()
        return res
    learned_model = None
    if not args.no_learned and b99 and (headline or name == "s3072"):
        try:
            ntrain = args.learned_train or W["learned_train"]
            res["learned_nprobe"], learned_model = learned_leg(ix, base, queries, gt, b, sizes, args, b99,
                                                               W["more"], ntrain, W["nlist"],
                                                               W.get("learned_nmax", 1024))
        except Exception as e:  # reported, never silently replaced
            log("%s: learned leg failed: %r" % (name, e))
            res["learned_nprobe"] = {"error": repr(e)}
    ix.set_autotune(False)  # serving runs with the tiles already measured
    if headline and not args.no_serve:
        try:
            res["serving"] = serve_leg(ix, queries, args, head, b999 or max(sweep, key=lambda r: r["recall"]),
                                       learned_model, res.get("learned_nprobe"))
        except Exception as e:
            log("serving leg failed: %r" % (e,))
            res["serving"] = {"error": repr(e)}
    if headline and not args.no_cpu_baseline:
        try:
            res["cpu_baseline"] = cpu_baseline(ix, base, queries, gt, b, sizes, pd, args, head)
        except Exception as e:
            log("cpu baseline failed: %r" % (e,))
            res["cpu_baseline"] = {"error": repr(e)}
    del learned_model, outs
    ix.close()
    del ix, base, queries, b
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return res


def point(r):
    return None if r is None else {"nprobe": r["nprobe"], "survivors": r["survivors"], "recall": r["recall"],
                                   "qps": r["qps"]}


def best(sweep, target):
    ok = [r for r in sweep if r["recall"] >= target]
    return max(ok, key=lambda r: r["qps"]) if ok else None


def sweep_points(ix, queries, gt, grid, sizes, pd, args):
    import torch
    import katoann
    nq, k = queries.shape[0], args.k
    dev = queries.device
    qptr = queries.data_ptr()
    from concurrent.futures import ThreadPoolExecutor
    pool = ThreadPoolExecutor(args.inflight) if args.inflight > 1 else None
    outs = [(torch.empty((nq, k), dtype=torch.int32, device=dev), torch.empty((nq, k), dtype=torch.float32, device=dev))
            for _ in range(max(1, args.inflight))]

    def steps_on(j, n_, np_, S):
        for _ in range(n_):
            ix.search_device(qptr, nq, outs[j][0].data_ptr(), outs[j][1].data_ptr(), k=k, nprobe=np_, survivors=S)

    def timed(n_, np_, S):
        if pool is None:
            steps_on(0, n_, np_, S)
            return
        shares = [n_ // args.inflight + (1 if j < n_ % args.inflight else 0) for j in range(args.inflight)]
        list(pool.map(lambda j: steps_on(j, shares[j], np_, S), range(args.inflight)))

    sweep = []
    for np_, S in grid:
        steps_on(0, args.warmup, np_, S)  # the autotuner runs on the first search of the shape
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
            steps_on(0, 1, np_, S)
            stage_sum += np.array(list(ix.stage_times().values()))
        ix.set_timing(False)
        ids = outs[0][0].cpu().numpy().view(np.uint32)
        rec = recall_at(ids, gt, 10)
        rank_bytes, rank_ops = rank_work(ix, queries, k, np_, S, sizes, pd)
        stages = dict(zip(katoann.STAGES, (stage_sum / args.steps).tolist()))
        rank_ms = stages["rank"]
        row = dict(nprobe=np_, survivors=S, ms_per_step=ms, qps=nq / (ms / 1e3), recall=rec, stages_ms=stages,
                   tile=ix.tile_config(nq, k=k, nprobe=np_, survivors=S), work=ix.last_work(),
                   rank_bytes=rank_bytes, rank_ops=rank_ops,
                   rank_gbs=rank_bytes / (rank_ms * 1e-3) / 1e9 if rank_ms > 0 else 0.0,
                   rank_tops=rank_ops / (rank_ms * 1e-3) / 1e12 if rank_ms > 0 else 0.0)
        sweep.append(row)
        log(json.dumps(row))
    if pool:
        pool.shutdown()
    return sweep, outs


def rank_work(ix, queries, k, np_, S, sizes, pd):
    """Algorithmic bytes of the ranking pass (every touched list once, P + 4
    bytes per row) and int8 ops (2 P per (query, candidate row))."""
    _, _, ex = ix.search_ex(queries.cpu().numpy(), k=k, nprobe=np_, survivors=S, normalize=False, want=("probes",))
    touched = np.unique(ex["probes"])
    return (float(sizes[touched].astype(np.float64).sum() * (pd + 4)),
            float(sizes[ex["probes"]].astype(np.float64).sum() * 2 * pd))


# ---------------------------------------------------------------- roofline

def profiled(name, r):
    """The committed rocprofv3 summary (profiles/*.md) of this workload and
    search shape: its tile and HBM traffic, or None."""
    want = "%s nprobe %d survivors %d" % (name, r["nprobe"], r["survivors"])
    pdir = os.path.join(ROOT, "profiles")
    for fn in sorted(os.listdir(pdir)) if os.path.isdir(pdir) else []:
        if not fn.endswith(".md"):
            continue
        lines = dict(l.split(": ", 1) for l in open(os.path.join(pdir, fn)).read().splitlines()
                     if l.startswith("traffic_"))
        cfg = lines.get("traffic_config", "").strip()
        if cfg.startswith(want + " tile "):
            tile = [int(x) for x in cfg[len(want + " tile "):].split(",")]
            return dict(tile=tile, bytes=float(lines["traffic_bytes_per_search"]), profile=fn,
                        kernel_avg_us=float(lines.get("traffic_kernel_avg_us", "nan")))
    return None


def roofline_pass(ix, queries, head, sizes, pd, d, nlist, name, args):
    """Roofline of the ranking pass at the headline point, with the tile of the
    committed profile of this shape pinned (so the profile's HBM traffic is the
    traffic of the kernels timed here), else with the tuned tile: algorithmic
    bytes / ops of one search over the summed duration of its ranking launches
    (HIP events on the search stream); the coarse (int8 MFMA) and refinement
    (HBM gather) stages' fractions beside it."""
    import katoann
    nq, k = queries.shape[0], args.k
    prof = None if args.profile_point else profiled(name, head)
    tile = head["tile"]
    if prof:
        kern, var, qc, shift, seed, join = (tuple(prof["tile"]) + (0,))[:6]
        ix.set_autotune(False)
        ix.set_rank_kernel(kern, var, qc)
        ix.set_band_factor(shift)
        ix.set_seed_rows(seed)
        ix.set_seed_join(join)
        tile = dict(kernel=kern, variant=var, queries_per_item=qc, log2_band_factor=shift, seed_rows=seed,
                    seed_join=join)
        # a pinned tile leaves the tuned ones for other shapes untouched: the
        # tuner's table is per shape, and this one is now measured with the pin
    import torch
    ids = torch.empty((nq, k), dtype=torch.int32, device=queries.device)
    dd = torch.empty((nq, k), dtype=torch.float32, device=queries.device)
    for _ in range(args.warmup):
        ix.search_device(queries.data_ptr(), nq, ids.data_ptr(), dd.data_ptr(), k=k, nprobe=head["nprobe"],
                         survivors=head["survivors"])
    ix.set_timing(True)
    st = np.zeros(len(katoann.STAGES))
    for _ in range(args.steps):
        ix.search_device(queries.data_ptr(), nq, ids.data_ptr(), dd.data_ptr(), k=k, nprobe=head["nprobe"],
                         survivors=head["survivors"])
        st += np.array(list(ix.stage_times().values()))
    ix.set_timing(False)
    launches = ix.last_work()["bands"]  # ranking launches of the timed tile
    st = dict(zip(katoann.STAGES, (st / args.steps).tolist()))
    if prof:  # back to the index's defaults with tuning for what follows
        ix.set_rank_kernel(1, 0, 1024)
        ix.set_band_factor(5)
        ix.set_seed_rows(1024)
        ix.set_seed_join(1)
        ix.set_autotune(True)
    rank_ms = st["rank"]
    byts, ops = head["rank_bytes"], head["rank_ops"]
    gbs, tops = byts / (rank_ms * 1e-3) / 1e9, ops / (rank_ms * 1e-3) / 1e12
    inten = ops / byts if byts else 0.0
    mfma = inten > MFMA_I8_PEAK_TOPS * 1e12 / (HBM_PEAK_GBS * 1e9)
    ach, peak, unit = (tops, MFMA_I8_PEAK_TOPS, "TOP/s") if mfma else (gbs, HBM_PEAK_GBS, "GB/s")
    dc = (d + 63) // 64 * 64
    S = max(head["survivors"], k)
    coarse_ops = 8.0 * nq * nlist * dc
    refine_bytes = float(nq) * S * (2 * dc + 8)
    out = {"bound": "mfma" if mfma else "hbm", "kernel": "ranking pass (all probe-rank bands of one search)",
           "achieved": ach, "peak": peak, "unit": unit, "frac": ach / peak,
           "traffic": prof["bytes"] if prof else None,
           "traffic_profile": prof["profile"] if prof else None, "tile": tile,
           "intensity_ops_per_byte": inten, "hbm_gbs": gbs, "hbm_frac": gbs / HBM_PEAK_GBS,
           "int8_tops": tops, "int8_frac": tops / MFMA_I8_PEAK_TOPS, "bytes_per_search": byts,
           "ops_per_search": ops, "ms_per_search": rank_ms, "launches_per_search": launches,
           "bytes_rule": "sum over touched lists of rows x (prefix bytes + 4-byte norm), once per search",
           "ops_rule": "2 x prefix bytes per (query, row of a probed list)",
           "timing": "HIP events on the search stream around every ranking launch, averaged over the steps",
           "stages_ms": st,
           "coarse": {"bound": "mfma", "ops_per_search": coarse_ops, "ms": st["coarse"],
                      "int8_tops": coarse_ops / (st["coarse"] * 1e-3) / 1e12,
                      "frac": coarse_ops / (st["coarse"] * 1e-3) / 1e12 / MFMA_I8_PEAK_TOPS,
                      "ops_rule": "8 x batch x nlist x Dc (four byte products)"},
           "refine": {"bound": "hbm", "bytes_per_search": refine_bytes, "ms": st["refine"],
                      "gbs": refine_bytes / (st["refine"] * 1e-3) / 1e9,
                      "frac": refine_bytes / (st["refine"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                      "bytes_rule": "batch x S x (2 Dc + 8): both code planes and the norm of every survivor"}}
    log("roofline: " + json.dumps(out))
    return out


# ---------------------------------------------------------------- learned nprobe

def learned_leg(ix, base, queries, gt, b, sizes, args, plain, more, ntrain, nlist, W_nmax=1024):
    """Per-query nprobe from the MLP trained on the GPU (untimed): training
    queries and validation queries of the same law are drawn apart from the
    benchmark queries; the margin is the smallest reaching recall@10 0.99 on the
    validation set; then the benchmark batch is timed like the fixed-nprobe
    sweep."""
    import torch
    import katoann
    from katoann import datasets, learned
    dev = queries.device
    n, d, nq, S = base.shape[0], base.shape[1], queries.shape[0], plain["survivors"]
    t0 = time.time()
    list_of_id = torch.empty(n, dtype=torch.int32, device=dev)
    list_of_id[b["order"].long()] = torch.repeat_interleave(
        torch.arange(nlist, dtype=torch.int32, device=dev), torch.from_numpy(sizes.astype(np.int64)).to(dev))
    trq = more(ntrain, 11)
    gtr = torch.from_numpy(datasets.ground_truth_torch(base, trq, 10).astype(np.int64)).to(dev)
    model, labels, hist = learned.train(ix, trq, list_of_id, gtr, epochs=20, seed=0)
    del trq, gtr
    val = more(nq, 12)
    gval = datasets.ground_truth_torch(base, val, 10)
    torch.cuda.synchronize()
    log("learned: %d training queries + fit %.1fs, loss %.3f -> %.3f" % (ntrain, time.time() - t0, hist[0], hist[-1]))
    # the model's nprobe ceiling sets the probe selection's width (a top-nmax
    # per query before the model runs): each candidate ceiling is calibrated on
    # the validation queries, the fastest that reaches recall 0.99 is kept
    ceilings = ([args.learned_nmax] if args.learned_nmax
                else [c for c in (plain["nprobe"], 256, 512, W_nmax) if c <= W_nmax])
    used = torch.empty(nq, dtype=torch.int32, device=dev)
    ids_d = torch.empty((nq, args.k), dtype=torch.int32, device=dev)
    dists_d = torch.empty((nq, args.k), dtype=torch.float32, device=dev)
    from concurrent.futures import ThreadPoolExecutor
    best_fit = None
    for nmax in sorted(set(min(c, nlist) for c in ceilings)):
        def run(qt, margin, nmax=nmax):
            ix.search_device(qt.data_ptr(), nq, ids_d.data_ptr(), dists_d.data_ptr(), k=args.k, nprobe=nmax,
                             survivors=S, model=model.struct(margin), npq_out_ptr=used.data_ptr())

        margin = None
        for mg in np.arange(-0.5, 4.01, 0.125):
            run(val, float(mg))
            if recall_at(ids_d.cpu().numpy().view(np.uint32), gval, 10) >= 0.99:
                margin = float(mg)
                break
        if margin is None:
            log("learned: ceiling %d reaches no recall 0.99 on the validation queries" % nmax)
            continue
        mstruct = model.struct(margin)
        outs = [(torch.empty_like(ids_d), torch.empty_like(dists_d), torch.empty_like(used))
                for _ in range(args.inflight)]

        def run_on(j, n_, nmax=nmax, mstruct=mstruct, outs=outs):
            for _ in range(n_):
                ix.search_device(queries.data_ptr(), nq, outs[j][0].data_ptr(), outs[j][1].data_ptr(), k=args.k,
                                 nprobe=nmax, survivors=S, model=mstruct, npq_out_ptr=outs[j][2].data_ptr())

        with ThreadPoolExecutor(args.inflight) as pool:
            list(pool.map(lambda j: run_on(j, args.warmup), range(args.inflight)))
            torch.cuda.synchronize()
            shares = [args.steps // args.inflight + (1 if j < args.steps % args.inflight else 0)
                      for j in range(args.inflight)]
            t0 = time.perf_counter()
            list(pool.map(lambda j: run_on(j, shares[j]), range(args.inflight)))
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t0) * 1e3 / args.steps
        log("learned: ceiling %d margin %.3f: %.3f ms per batch" % (nmax, margin, ms))
        if best_fit is None or ms < best_fit[0]:
            best_fit = (ms, nmax, margin, mstruct, run)
    if best_fit is None:
        raise RuntimeError("no margin reaches recall 0.99 on the validation queries")
    ms, nmax, margin, mstruct, run = best_fit
    ix.set_timing(True)
    stage_sum = np.zeros(len(katoann.STAGES))
    for _ in range(args.steps):
        run(queries, margin)
        stage_sum += np.array(list(ix.stage_times().values()))
    ix.set_timing(False)
    rec = recall_at(ids_d.cpu().numpy().view(np.uint32), gt, 10)
    out = {"qps": nq / (ms / 1e3), "ms_per_step": ms, "recall": rec, "margin": margin, "survivors": S,
           "nprobe_max": nmax, "ceilings_tried": sorted(set(min(c, nlist) for c in ceilings)),
           "mean_nprobe": float(used.float().mean().item()), "plain_nprobe": plain["nprobe"],
           "plain_qps": plain["qps"], "speedup_vs_plain": (nq / (ms / 1e3)) / plain["qps"],
           "stages_ms": dict(zip(katoann.STAGES, (stage_sum / args.steps).tolist())),
           "training": "%d queries of the same law (own seed), coverage labels from the exact top-10, 20 epochs "
                       "Adam, bf16 MFMA; margin calibrated on %d further queries" % (ntrain, nq)}
    log("learned: " + json.dumps(out))
    return out, (model, mstruct, nmax)


# ---------------------------------------------------------------- serving

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
    use_model = learned_model is not None and isinstance(learned_out, dict) and "survivors" in learned_out
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
            if not hi[i] and use_model:
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
           "peak_used_bytes": st["peak_used"], "budget_bytes": st["budget"], "wall_s": wall,
           "max_submit_lag_ms": lag[0] * 1e3,
           "mix": "sizes floor(exp(U(0, ln 4096))), k {10, 100}, half at target 0.99 (%s), half at 0.999 "
                  "(nprobe %d, survivors %d)" % ("learned nprobe" if use_model else "nprobe %d" % at99["nprobe"],
                                                at999["nprobe"], at999["survivors"])}
    log("serving: " + json.dumps(out))
    return out


# ---------------------------------------------------------------- CPU baseline

def cpu_info():
    model, cores = "unknown", None
    try:
        model = [l.split(":", 1)[1].strip() for l in open("/proc/cpuinfo") if l.startswith("model name")][0]
    except Exception:
        pass
    try:
        r = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=20)
        cores = len({l for l in r.stdout.splitlines() if l and not l.startswith("#")})
    except Exception:
        pass
    return model, cores


def cpu_baseline(ix, base, queries, gt, b, sizes, pd, args, head):
    """The reference's own CPU search (oracle/_ref/katoann, AVX-512 VNNI engine
    where AMX is absent) on a bounded sample of the headline workload: the
    first queries of the batch (up to --cpu-queries) with the index's lists
    they probe at the headline point (every centroid, the other lists empty, so
    these queries search exactly as on the full index); timed by its CLI (one
    warm-up pass, median of 3, batches of up to 4,096) on this host's cores,
    with its ids checked against the GPU's."""
    import torch
    import oracle as O
    from katoann import formats
    if not os.path.exists(O.REF_CLI):
        raise RuntimeError("reference build missing (oracle/_ref/katoann)")
    nq_all, k = queries.shape[0], args.k
    np_, S = head["nprobe"], head["survivors"]
    host_q = queries.cpu().numpy()
    ids_g, dists_g, ex = ix.search_ex(host_q, k=k, nprobe=np_, survivors=S, normalize=False, want=("probes",))
    nq = min(args.cpu_queries, nq_all)
    while True:  # the sample: the first nq queries, bounded by the rows of their lists
        lists = np.unique(ex["probes"][:nq])
        rows = int(sizes[lists].astype(np.int64).sum())
        if rows <= args.cpu_rows or nq <= 16:
            break
        nq //= 2
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    model, phys = cpu_info()
    order = b["order"]
    bounds = np.concatenate([[0], np.cumsum(sizes.astype(np.int64))])
    keep = np.zeros(len(sizes), bool)
    keep[lists] = True
    ids_all = order.cpu().numpy().astype(np.uint32)
    # the load's scale comes from the smallest and largest rotated coordinate of
    # the whole index: the rows that may hold them (the 64 smallest row minima
    # and 64 largest row maxima of a float32 rotation) join the sub-index in
    # their own lists, which the sample's queries do not probe when they are not
    # probed lists already, so the sub-index quantises exactly as the full one
    rot = b["rotation"]
    lo_v, lo_i, hi_v, hi_i = [], [], [], []
    for s0 in range(0, base.shape[0], 1 << 20):
        y = base[s0:s0 + (1 << 20)] @ rot.T
        mn, mx = y.min(1).values, y.max(1).values
        v, i = torch.topk(mn, min(64, mn.numel()), largest=False)
        lo_v.append(v), lo_i.append(i + s0)
        v, i = torch.topk(mx, min(64, mx.numel()))
        hi_v.append(v), hi_i.append(i + s0)
        del y
    lo_v, lo_i, hi_v, hi_i = (torch.cat(t) for t in (lo_v, lo_i, hi_v, hi_i))
    ext = torch.cat([lo_i[torch.topk(lo_v, min(64, lo_v.numel()), largest=False).indices],
                     hi_i[torch.topk(hi_v, min(64, hi_v.numel())).indices]]).cpu().numpy()
    list_of = np.repeat(np.arange(len(sizes)), sizes.astype(np.int64))[np.argsort(ids_all, kind="stable")]
    extra = np.unique(ext[~keep[list_of[ext]]]).astype(np.uint32)
    # the sub-index holds `rows` points: their ids renumbered 0..rows-1 in the
    # order of the original ids (ties between equal distances fall the same
    # way), the ground truth and the GPU's ids mapped alike (a neighbour
    # outside the sub-index maps to an id no search returns)
    kept = np.sort(np.concatenate([ids_all[bounds[c]:bounds[c + 1]] for c in lists] + [extra]))
    renum = np.full(base.shape[0], 0xFFFFFFFE, np.uint32)
    renum[kept] = np.arange(kept.size, dtype=np.uint32)
    sel = torch.cat([order[bounds[c]:bounds[c + 1]] for c in lists]).long()
    vec = base.index_select(0, sel).cpu().numpy()
    xvec = base.index_select(0, torch.from_numpy(extra.astype(np.int64)).to(base.device)).cpu().numpy()
    xlist = list_of[extra]
    rows += int(extra.size)
    with tempfile.TemporaryDirectory(dir="/tmp") as dirn:
        prefix = os.path.join(dirn, "cpu")
        at = 0
        lid, lvec = [], []
        for c in range(len(sizes)):
            if keep[c]:
                m = int(sizes[c])
                lid.append(renum[ids_all[bounds[c]:bounds[c + 1]]])
                lvec.append(vec[at:at + m])
                at += m
            else:  # empty, or the extreme rows of an unprobed list (ascending ids)
                xs = np.nonzero(xlist == c)[0]
                lid.append(renum[extra[xs]])
                lvec.append(xvec[xs].reshape(len(xs), base.shape[1]))
        t0 = time.time()
        # written as not normalised: the CLI then ranks the queries as given,
        # exactly as the GPU search above (normalize=False) did
        formats.write_index(prefix, b["centroids"].cpu().numpy(), lid, lvec, rows, False,
                            b["rotation"].cpu().numpy(), pd)
        del vec, lvec, xvec
        formats.write_fbin(os.path.join(dirn, "q.fbin"), host_q[:nq])
        formats.write_gt(os.path.join(dirn, "gt.bin"), renum[gt[:nq]])
        log("cpu baseline: %d queries, %d rows of %d lists written in %.1fs" % (nq, rows, len(lists), time.time() - t0))
        outp = os.path.join(dirn, "res")
        cmd = [O.REF_CLI, "search", "--index", prefix, "--queries", os.path.join(dirn, "q.fbin"), "--gt",
               os.path.join(dirn, "gt.bin"), "--k", str(k), "--nprobe", str(np_), "--survivors", str(S), "--batch",
               "4096", "--engine", "avx512", "--threads", str(threads), "--out", outp]
        t0 = time.time()
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=args.cpu_timeout)
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-500:])
        lines = [json.loads(x) for x in r.stdout.splitlines() if x.startswith("{")]
        res = [x for x in lines if "qps" in x][0]
        hdr = lines[0]
        raw = np.fromfile(outp + ".np%d.bin" % np_, dtype=np.uint32)
        c_ids = raw[2:2 + nq * k].reshape(nq, k)
        c_d = raw[2 + nq * k:].view(np.float32).reshape(nq, k)
        log("cpu baseline: %s (wall %.1fs incl. load %.1fs)" % (json.dumps(res), time.time() - t0, hdr["load_s"]))
    same_rows = int((np.all(c_ids == renum[ids_g[:nq]], 1) &
                     np.all(c_d.view(np.uint32) == dists_g[:nq].view(np.uint32), 1)).sum())
    return {"value": res["qps"], "unit": "queries/s", "cores": threads, "kind": "reference",
            "sample": "the first %d of the %d queries in one batch (the reference CLI's --batch 4096), nprobe %d, "
                      "survivors %d, on the %d lists they probe with every centroid kept (%d rows with the rows "
                      "that fix the index's scale); median of 3 "
                      "after a warm-up; the full 41 GB of index files is not written within the benchmark's "
                      "run" % (nq, nq_all, np_, S, len(lists), rows),
            "recall_at_10": res.get("recall"), "rows_identical_to_gpu": same_rows, "rows_compared": nq,
            "engine": hdr["engine"], "cpu_model": model, "physical_cores": phys,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"), "load_s": hdr["load_s"]}


# ---------------------------------------------------------------- main

EXTRA_KEYS = {"spectrum": "spectrum_1Mx1024", "s3072": "spectrum_990kx3072"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="zipf", choices=("zipf", "spectrum", "s3072"),
                    help="headline workload: zipf = 10M x 1024 Zipf(1.1) mixture, nlist 16,384 (default); "
                         "spectrum = 1M x 1024 i^-1 spectrum, nlist 4,096; s3072 = 990k x 3072 i^-0.8 spectrum")
    ap.add_argument("--extra", default="spectrum,s3072", help="further workloads reported as extra keys ('' none)")
    ap.add_argument("--k", type=int, default=10)
    ap.add_argument("--kmeans-iters", type=int, default=10)
    ap.add_argument("--inflight", type=int, default=3,
                    help="batches in flight in the timed loop (each on its own search context and stream)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-queries", type=int, default=512)
    ap.add_argument("--cpu-rows", type=int, default=2_000_000, help="rows of the CPU sample's lists at most")
    ap.add_argument("--cpu-timeout", type=int, default=600)
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--learned-train", type=int, default=0, help="training queries (default per workload)")
    ap.add_argument("--learned-nmax", type=int, default=0, help="nprobe ceiling of the model (default per workload)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=6.0)
    ap.add_argument("--serve-load", type=float, default=0.5, help="offered queries/s as a fraction of batch QPS")
    ap.add_argument("--profile-point", default="",
                    help="NPROBE:S: measure only this point of the headline workload (profiler runs)")
    ap.add_argument("--tile", default="", help="kernel,variant,queries_per_item,log2_band_factor,seed_rows[,seed_join] to pin "
                                               "with --profile-point")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the run, three device copies of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X (--gpus 1)", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    res = run_workload(args.workload, args, dev, headline=True)
    if args.profile_point:
        if args.calib_copy_mib:
            src = torch.empty(args.calib_copy_mib << 20, dtype=torch.uint8, device=dev).fill_(1)
            for _ in range(3):
                dst = src.clone()
                torch.cuda.synchronize()
                del dst
        r = res["head"]
        print(json.dumps({"profile_point": args.profile_point, "tile": r["tile"], "ms_per_step": r["ms_per_step"],
                          "recall": r["recall"], "roofline": res["roofline"]}), flush=True)
        return
    extras = {}
    for name in filter(None, args.extra.split(",")):
        if name == args.workload:
            continue
        try:
            extras[name] = run_workload(name, args, dev, headline=False)
            for kk in ("head", "b999"):
                extras[name].pop(kk, None)
        except Exception as e:  # reported, never silently replaced
            log("extra workload %s failed: %r" % (name, e))
            extras[name] = {"error": repr(e)}
    head = res["head"]
    ok = res["qps_at_099"] is not None
    out = {
        "metric": "QPS at recall@10>=0.99 (batch 4096, 1 GPU)",
        "value": res["qps_at_099"] if ok else 0.0,
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
        "config": {"workload": res["workload"], "nprobe": head["nprobe"], "survivors": head["survivors"],
                   "recall_at_10": head["recall"], "batches_in_flight": args.inflight},
        "qps_at_0999": res["qps_at_0999"],
        "at_0999": res["at_0999"],
        "roofline": res["roofline"],
        "stages_ms": head["stages_ms"],
        "sweep": res["sweep"],
        "learned_nprobe": res.get("learned_nprobe"),
        "serving": res.get("serving"),
        "cpu_baseline": res.get("cpu_baseline"),
        "extra": {EXTRA_KEYS.get(kk, kk): v for kk, v in extras.items()},
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
