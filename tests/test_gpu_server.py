"""The continuous-batching server on the GPU: every request's result equals the
same request searched alone (bit for bit: a query's result does not depend on
the batch it runs in), admission keeps the in-flight footprint within the HBM
budget, and requests that can never fit or are malformed are refused at
submission."""
import threading
import time

import numpy as np
import pytest

import common
import katoann

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def served(tmp_path_factory):
    d = tmp_path_factory.mktemp("srv")
    prefix, base, q = common.small_index(d, n=60000, nq=3000, d=128, nlist=256, centres=256, seed=3)
    ix = katoann.Index()
    ix.load(prefix)
    return ix, q


def _requests(q, n, seed, with_model=None):
    rng = np.random.default_rng(seed)
    out, pos = [], 0
    for i in range(n):
        m = int(np.exp(rng.uniform(0, np.log(600))))
        if pos + m > q.shape[0]:
            pos = 0
        k = int(rng.choice([10, 100]))
        out.append(dict(queries=q[pos:pos + m], k=k, nprobe=int(rng.choice([8, 32])), survivors=max(32, k),
                        model=with_model if (with_model is not None and i % 3 == 0) else None))
        pos += m
    return out


def _alone(ix, r):
    nq = r["queries"].shape[0]
    if r["model"] is None:
        ids, dists, _ = ix.search_ex(r["queries"], k=r["k"], nprobe=r["nprobe"], survivors=r["survivors"])
        return ids, dists
    import torch
    q = torch.from_numpy(np.ascontiguousarray(r["queries"])).cuda()
    ids = torch.empty((nq, r["k"]), dtype=torch.int32, device="cuda")
    dists = torch.empty((nq, r["k"]), dtype=torch.float32, device="cuda")
    ix.search_device(q.data_ptr(), nq, ids.data_ptr(), dists.data_ptr(), k=r["k"], nprobe=r["nprobe"],
                     survivors=r["survivors"], model=r["model"], normalize=bool(ix.normalized))
    torch.cuda.synchronize()
    return ids.cpu().numpy().view(np.uint32), dists.cpu().numpy()


def _serve(srv, reqs, threads=3, jitter=0.002, seed=0):
    """Submit from several threads with random gaps; returns results by index."""
    tickets = [None] * len(reqs)
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(jitter, len(reqs))

    def run(lo):
        for i in range(lo, len(reqs), threads):
            time.sleep(gaps[i])
            r = reqs[i]
            tickets[i] = srv.submit(r["queries"], k=r["k"], nprobe=r["nprobe"], survivors=r["survivors"],
                                    model=r["model"])

    ts = [threading.Thread(target=run, args=(t,)) for t in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return [srv.wait(t) for t in tickets]


def _model():
    """An untrained network (any fixed one gives a valid per-query nprobe); the
    network object must outlive every search that uses its struct."""
    import torch
    from katoann import learned
    m = learned.NprobeMLP(8, 64, seed=4)
    m.y_mean, m.y_std = 3.0, 1.0
    torch.cuda.synchronize()
    return m, m.struct(0.5, nmin=1)


def test_results_same_alone_and_in_mix(served):
    ix, q = served
    net, model = _model()
    reqs = _requests(q, 60, seed=1, with_model=model)
    print("server mix: %d requests, %d queries" % (len(reqs), sum(r["queries"].shape[0] for r in reqs)), flush=True)
    alone = [_alone(ix, r) for r in reqs]
    with katoann.Server(ix, max_inflight=4, max_batch=1024) as srv:
        got = _serve(srv, reqs)
        srv.drain()
        st = srv.stats()
    print(st, flush=True)
    for i, ((ia, da), (ig, dg, lat)) in enumerate(zip(alone, got)):
        assert np.array_equal(ia, ig), i
        assert np.array_equal(da.view(np.uint32), dg.view(np.uint32)), i
        assert lat >= 0.0
    assert st["requests"] == len(reqs)
    assert st["queries"] == sum(r["queries"].shape[0] for r in reqs)
    assert 1 <= st["batches"] <= len(reqs)
    assert st["p50_ms"] <= st["p95_ms"] <= st["p99_ms"] <= st["max_ms"]
    assert st["peak_reserved"] <= st["budget"]


def test_merged_batches_same_as_alone(served):
    """A burst of same-parameter requests queued behind one slot is merged into
    larger batches; results are unchanged."""
    ix, q = served
    reqs = [dict(queries=q[i * 37:(i + 1) * 37], k=10, nprobe=16, survivors=32, model=None) for i in range(40)]
    alone = [_alone(ix, r) for r in reqs]
    with katoann.Server(ix, max_inflight=1, max_batch=4096) as srv:
        tickets = [srv.submit(r["queries"], k=10, nprobe=16, survivors=32) for r in reqs]
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
    print(st, flush=True)
    for (ia, da), (ig, dg, _) in zip(alone, got):
        assert np.array_equal(ia, ig) and np.array_equal(da.view(np.uint32), dg.view(np.uint32))
    assert st["batches"] < len(reqs)  # the burst queued behind the first batch ran merged
    assert st["peak_inflight"] == 1


def test_admission_within_budget(served):
    ix, q = served
    reqs = _requests(q, 30, seed=2)
    alone = [_alone(ix, r) for r in reqs]
    with katoann.Server(ix, max_inflight=4, max_batch=1024, hbm_budget=1 << 40) as probe:
        biggest = max(probe.footprint(r["queries"].shape[0], k=r["k"], nprobe=r["nprobe"], survivors=r["survivors"])
                      for r in reqs)
        huge = probe.footprint(1000, k=100, nprobe=256, survivors=1000)
    budget = int(biggest * 1.3)
    assert huge > budget
    with katoann.Server(ix, max_inflight=4, max_batch=1024, hbm_budget=budget) as srv:
        got = _serve(srv, reqs, jitter=0.0005)
        st = srv.stats()
        with pytest.raises(ValueError, match="above the budget"):
            srv.submit(q[:1000], k=100, nprobe=256, survivors=1000)
    print(st, "budget", budget, flush=True)
    assert st["budget"] == budget and st["peak_reserved"] <= budget
    for (ia, da), (ig, dg, _) in zip(alone, got):
        assert np.array_equal(ia, ig) and np.array_equal(da.view(np.uint32), dg.view(np.uint32))


def test_refusals(served):
    ix, q = served
    with katoann.Server(ix, max_inflight=2, max_batch=512) as srv:
        with pytest.raises(ValueError, match="nprobe"):
            srv.submit(q[:4], k=10, nprobe=ix.nlist + 1)
        with pytest.raises(ValueError, match="max_batch"):
            srv.submit(q[:513], k=10, nprobe=8)
        with pytest.raises(ValueError, match="unknown ticket"):
            srv.poll(12345)
        t = srv.submit(q[:4], k=10, nprobe=8)
        ids, dists, _ = srv.wait(t)
        assert ids.shape == (4, 10)
        srv.drain()


def test_overflow_rerun_reserves_its_room(served):
    """A survivor pool too small for a band (8 slots of room): the search may
    not grow past its batch's reservation, so the batch reserves the room its
    rerun needs and runs again; results equal the searches alone, and the
    reservation always covered the buffers a batch ran with."""
    ix, q = served
    ix.set_pool_allowance(8)
    try:
        reqs = [dict(queries=q[i * 200:(i + 1) * 200], k=10, nprobe=16, survivors=32, model=None) for i in range(6)]
        alone = [_alone(ix, r) for r in reqs]
        ix.set_pool_allowance(8)  # forget the room the searches alone found
        srv = katoann.Server(ix, max_inflight=2, max_batch=200)
        try:
            print("server overflow reruns", flush=True)
            got = _serve(srv, reqs, threads=2)
            st = srv.stats()
        finally:
            srv.close()
        for (ids, dists, _), (ai, ad) in zip(got, alone):
            assert np.array_equal(ids, ai) and np.array_equal(dists.view(np.uint32), ad.view(np.uint32))
        assert st["peak_used"] > 0 and st["peak_reserved"] >= st["peak_used"]
        assert st["peak_reserved"] <= st["budget"]
    finally:
        ix.set_pool_allowance(0)
