"""The allocation-free entry point (katoann_search_noalloc): through the C ABI,
its results equal katoann_search_ex's with device pointers, bit for bit, on
two calls in a row (the second with fewer queries, on the scratch the first
left behind) and with per-query nprobe; a search whose buffers exceed the
static scratch is refused with a clear error."""
import numpy as np
import pytest

import common
import katoann

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def index(tmp_path_factory):
    d = tmp_path_factory.mktemp("noalloc")
    prefix, base, q = common.small_index(d, n=60000, nq=1000, d=128, nlist=256, centres=256, seed=8)
    ix = katoann.Index()
    ix.load(prefix)
    return ix, q


def _both(ix, q, **kw):
    import torch
    nq, k = q.shape[0], kw["k"]
    qd = torch.from_numpy(np.ascontiguousarray(q)).cuda()
    out = []
    for fn in (ix.search_device, ix.search_noalloc):
        # sentinels no search writes: id 0xFFFFFFFE (ids are below 60,000; padding is
        # 0xFFFFFFFF) and distance -1 (distances are >= 0)
        ids = torch.full((nq, k), -2, dtype=torch.int32, device="cuda")
        dists = torch.full((nq, k), -1.0, dtype=torch.float32, device="cuda")
        fn(qd.data_ptr(), nq, ids.data_ptr(), dists.data_ptr(), **kw)
        torch.cuda.synchronize()
        out.append((ids.cpu().numpy().view(np.uint32), dists.cpu().numpy()))
    return out


@pytest.mark.parametrize("k,nprobe,S", [(10, 16, 32), (100, 32, 128), (10, 64, 2000)])
def test_noalloc_equals_search_ex(index, k, nprobe, S):
    ix, q = index
    for nq in (1000, 137):  # the second call on the scratch the first one left
        print("noalloc", k, nprobe, S, nq, flush=True)
        (ai, ad), (bi, bd) = _both(ix, q[:nq], k=k, nprobe=nprobe, survivors=S)
        assert np.array_equal(ai, bi)
        assert np.array_equal(ad.view(np.uint32), bd.view(np.uint32))
        assert (ai != 0xFFFFFFFE).all() and (ad >= 0).all()


def test_noalloc_per_query_nprobe(index):
    import torch
    ix, q = index
    nq = 500
    npq = torch.from_numpy((np.arange(nq, dtype=np.int32) % 24) + 1).cuda()
    print("noalloc per-query nprobe", flush=True)
    (ai, ad), (bi, bd) = _both(ix, q[:nq], k=10, nprobe=24, survivors=32, npq_ptr=npq.data_ptr())
    assert np.array_equal(ai, bi) and np.array_equal(ad.view(np.uint32), bd.view(np.uint32))


def test_noalloc_refuses_above_scratch(index):
    import torch
    ix, q = index
    nq = 8192  # pools of 8192 x 4096 keys: 256 MiB and more
    big = np.tile(q, (nq // q.shape[0] + 1, 1))[:nq]
    qd = torch.from_numpy(np.ascontiguousarray(big)).cuda()
    ids = torch.empty((nq, 10), dtype=torch.int32, device="cuda")
    dists = torch.empty((nq, 10), dtype=torch.float32, device="cuda")
    print("noalloc above scratch", flush=True)
    with pytest.raises(ValueError, match="bytes of scratch"):
        ix.search_noalloc(qd.data_ptr(), nq, ids.data_ptr(), dists.data_ptr(), k=10, nprobe=64, survivors=4096)
    # and the entry point still serves a search that fits
    (ai, ad), (bi, bd) = _both(ix, q[:50], k=10, nprobe=16, survivors=32)
    assert np.array_equal(ai, bi) and np.array_equal(ad.view(np.uint32), bd.view(np.uint32))
