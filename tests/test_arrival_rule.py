"""The refinement's arrival rule (src/index.cpp:1117-1125): survivors arrive in
order and enter a K-entry max-heap of (dist, id), admitted while it has room,
else iff dist < the heap's largest dist (strict), evicting the largest
(dist, id). The GPU evaluates it in closed form (k_refine): with M the K-th
smallest dist, the heap ends with every survivor with dist < M, plus, of those
with dist == M that arrived when fewer than K earlier arrivals had dist <= M,
the ones with the smallest ids. Checked here against a literal simulation on
inputs dense in ties."""
import numpy as np


def heap_rule(dists, ids, K):
    heap = []  # (dist, id) pairs; max by (dist, id)
    for d, i in zip(dists, ids):
        if len(heap) < K:
            heap.append((d, i))
        elif d < max(heap)[0]:
            heap.remove(max(heap))
            heap.append((d, i))
    return sorted(heap)


def closed_form(dists, ids, K):
    n = len(dists)
    if n <= K:
        return sorted(zip(dists, ids))
    M = np.sort(dists)[K - 1]
    out = []
    le_before = 0
    for d, i in zip(dists, ids):
        if d < M or (d == M and le_before < K):
            out.append((d, i))
        if d <= M:
            le_before += 1
    return sorted(out)[:K]


def test_closed_form_matches_heap():
    rng = np.random.default_rng(5)
    for trial in range(3000):
        n = int(rng.integers(1, 40))
        K = int(rng.integers(1, 12))
        levels = int(rng.integers(1, 6))
        dists = rng.integers(0, levels, n).astype(np.float32)
        ids = rng.permutation(1000)[:n].astype(np.uint32)
        assert heap_rule(dists.tolist(), ids.tolist(), K) == closed_form(dists, ids, K), (trial, n, K)
