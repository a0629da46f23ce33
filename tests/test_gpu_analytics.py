"""GPU BFS and PageRank over the gapped layout against the host oracles:
BFS levels bit-exact, PageRank (FP32) within 1e-6 L1 of the FP64 oracle run
for the same number of iterations."""
import numpy as np
import pytest

import lpma_amd as L
import oracle
from gpu_util import say

pytestmark = pytest.mark.gpu


def build(scale, E, seed, deletes=0):
    V = 1 << scale
    keys = L.shuffle(L.gen_rmat(scale, E, seed=seed), seed + 1)
    pma = L.PMA(V)
    pma.insert(keys[: E // 2])
    for i in range(E // 2, E, 5000):
        pma.insert(keys[i:i + 5000])
    ref = oracle.sort_unique(keys)
    if deletes:
        rng = np.random.default_rng(seed)
        d = ref[rng.choice(ref.size, deletes, replace=False)]
        pma.delete(d)
        ref = oracle.apply(ref, (), d)
    return V, pma, ref


@pytest.mark.parametrize("scale,E,deletes", [(10, 6000, 0), (14, 150000, 20000), (16, 400000, 100000)])
def test_bfs_levels_bit_exact(scale, E, deletes):
    say(f"bfs scale={scale}")
    V, pma, ref = build(scale, E, 5, deletes)
    for src in (0, 1, V // 2):
        lv, depth, scanned = pma.bfs(src)
        got = lv.cpu().numpy().view(np.uint32)
        want = oracle.bfs(ref, V, src)
        assert np.array_equal(got, want)
        reached = want != 0xFFFFFFFF
        assert depth == int(want[reached].max()) + 1
        ro, _ = oracle.csr_of(ref, V)
        deg = np.diff(ro.astype(np.int64))
        assert scanned == int(deg[reached].sum())
    pma.close()


@pytest.mark.parametrize("scale,E,deletes", [(10, 6000, 0), (14, 150000, 20000), (16, 400000, 100000),
                                             (18, 2000000, 300000)])
def test_pagerank_fp32_within_1e6_of_fp64(scale, E, deletes):
    say(f"pagerank scale={scale}")
    V, pma, ref = build(scale, E, 7, deletes)
    pr, iters, delta = pma.pagerank(0.85, 1e-6, 50)
    got = pr.cpu().numpy().astype(np.float64)
    want, _ = oracle.pagerank(ref, V, 0.85, iters)
    assert abs(got.sum() - 1.0) < 1e-5
    l1 = np.abs(got - want).sum()
    assert l1 < 1e-6, (l1, iters, delta)
    pma.close()


def test_bfs_frontier_of_empty_rows():
    """Level 1 is 6000 rows of which all but every 97th are empty, so one run
    of the expansion (2048 slots) spans more frontier entries than it stages
    in LDS and the rest are found by the global search."""
    say("bfs empty-row frontier")
    from gpu_util import keys_of
    V = 1 << 14
    hub = np.arange(1, 6001)
    srcs = [np.zeros(hub.size, dtype=np.int64)]
    dsts = [hub]
    for i in range(1, 6001, 97):
        srcs.append(np.full(12, i))
        dsts.append(6001 + (i * 13 + np.arange(12) * 131) % 9000)
    keys = oracle.sort_unique(keys_of(np.concatenate(srcs), np.concatenate(dsts)))
    pma = L.PMA(V)
    pma.insert(keys)
    for src in (0, 97):
        lv, depth, scanned = pma.bfs(src)
        want = oracle.bfs(keys, V, src)
        assert np.array_equal(lv.cpu().numpy().view(np.uint32), want)
        reached = want != 0xFFFFFFFF
        ro, _ = oracle.csr_of(keys, V)
        deg = np.diff(ro.astype(np.int64))
        assert scanned == int(deg[reached].sum())
        assert depth == int(want[reached].max()) + 1
    pma.close()


def test_pagerank_partial_last_bin():
    """V = 20000: three destination bins of 8192 vertices, the last one
    partial; uniform edges with a tenth of the vertices dangling."""
    say("pagerank V=20000")
    from gpu_util import keys_of
    V = 20000
    rng = np.random.default_rng(11)
    src = rng.integers(0, V, 300000)
    src = src[src % 10 != 3]  # vertices 3, 13, ... have no out-edges
    dst = rng.integers(0, V, src.size)
    keep = src != dst
    keys = oracle.sort_unique(keys_of(src[keep], dst[keep]))
    pma = L.PMA(V)
    pma.insert(keys)
    pr, iters, delta = pma.pagerank(0.85, 1e-6, 50)
    got = pr.cpu().numpy().astype(np.float64)
    want, _ = oracle.pagerank(keys, V, 0.85, iters)
    assert abs(got.sum() - 1.0) < 1e-5
    l1 = np.abs(got - want).sum()
    assert l1 < 1e-6, (l1, iters, delta)
    pma.close()


def test_pagerank_more_bins_than_staged_pass():
    """V = 2^24 + 5: 2049 destination bins, past the staged bin pass's LDS,
    so the unstaged bin pass (no padding) runs."""
    say("pagerank V=2^24+5")
    from gpu_util import keys_of
    V = (1 << 24) + 5
    rng = np.random.default_rng(12)
    src = rng.integers(0, V, 200000)
    dst = np.concatenate([rng.integers(0, 1 << 12, 100000), rng.integers(V - 4096, V, 100000)])
    keep = src != dst
    keys = oracle.sort_unique(keys_of(src[keep], dst[keep]))
    pma = L.PMA(V)
    pma.insert(keys)
    pr, iters, delta = pma.pagerank(0.85, 1e-6, 50)
    got = pr.cpu().numpy().astype(np.float64)
    want, _ = oracle.pagerank(keys, V, 0.85, iters)
    assert abs(got.sum() - 1.0) < 1e-5
    l1 = np.abs(got - want).sum()
    assert l1 < 1e-6, (l1, iters, delta)
    pma.close()


def test_pagerank_atomic_push_above_binned_limit():
    """V = 2^27 + 1: more destination bins than the binned push's LDS cursors
    hold, so PageRank takes the per-edge compensated atomic push."""
    say("pagerank V=2^27+1")
    from gpu_util import keys_of
    V = (1 << 27) + 1
    rng = np.random.default_rng(13)
    src = rng.integers(0, V, 100000)
    dst = np.concatenate([rng.integers(0, 4096, 50000), rng.integers(V - 4096, V, 50000)])
    keep = src != dst
    keys = oracle.sort_unique(keys_of(src[keep], dst[keep]))
    pma = L.PMA(V)
    pma.insert(keys)
    pr, iters, delta = pma.pagerank(0.85, 1e-6, 50)
    got = pr.cpu().numpy().astype(np.float64)
    del pr
    want, _ = oracle.pagerank(keys, V, 0.85, iters)
    assert abs(got.sum() - 1.0) < 1e-5
    l1 = np.abs(got - want).sum()
    assert l1 < 1e-6, (l1, iters, delta)
    pma.close()


@pytest.mark.parametrize("slots", [4, 64, 256])
def test_bfs_and_pagerank_other_segment_sizes(slots):
    """The expansion and the bin passes read live lines through the segment
    counts, whose geometry follows the segment size."""
    say(f"analytics segment_slots={slots}")
    V = 1 << 14
    keys = L.shuffle(L.gen_rmat(14, 150000, seed=21), 22)
    pma = L.PMA(V, segment_slots=slots)
    pma.insert(keys[:100000])
    pma.insert(keys[100000:])
    ref = oracle.sort_unique(keys)
    rng = np.random.default_rng(23)
    d = ref[rng.choice(ref.size, 20000, replace=False)]
    pma.delete(d)
    ref = oracle.apply(ref, (), d)
    for src in (0, 3):
        lv, depth, scanned = pma.bfs(src)
        assert np.array_equal(lv.cpu().numpy().view(np.uint32), oracle.bfs(ref, V, src))
    pr, iters, delta = pma.pagerank(0.85, 1e-6, 50)
    want, _ = oracle.pagerank(ref, V, 0.85, iters)
    l1 = np.abs(pr.cpu().numpy().astype(np.float64) - want).sum()
    assert l1 < 1e-6, (l1, iters, delta)
    pma.close()


@pytest.mark.parametrize("V", [1, 2, 3])
def test_tiny_and_empty_graphs(V):
    """No edges: every vertex dangles, so PageRank is uniform and BFS reaches
    only its root; then one edge 0 -> 1 where there is a vertex 1."""
    say(f"analytics tiny V={V}")
    from gpu_util import keys_of
    pma = L.PMA(V)
    for keys in ([], [(0, 1)] if V > 1 else []):
        if keys:
            pma.insert(keys_of([k[0] for k in keys], [k[1] for k in keys]))
        ref = oracle.sort_unique(keys_of([k[0] for k in keys], [k[1] for k in keys])) if keys else \
            np.empty(0, dtype=np.uint64)
        lv, depth, scanned = pma.bfs(0)
        want = oracle.bfs(ref, V, 0)
        assert np.array_equal(lv.cpu().numpy().view(np.uint32), want)
        assert scanned == len(keys)
        pr, iters, delta = pma.pagerank(0.85, 1e-6, 50)
        got = pr.cpu().numpy().astype(np.float64)
        exp, _ = oracle.pagerank(ref, V, 0.85, iters)
        assert np.abs(got - exp).sum() < 1e-6, (got, exp)
    pma.close()
