"""CPU pinning of the analytics oracles: BFS levels against scipy's
unweighted shortest paths, PageRank against the dominant eigenvector of the
Google matrix (numpy) and against the closed form on a directed cycle."""
import numpy as np
import pytest

import oracle

U32 = np.uint64(32)


def keys_of(s, d):
    return (np.asarray(s, dtype=np.uint64) << U32) | np.asarray(d, dtype=np.uint64)


def random_graph(rng, V, E):
    return oracle.sort_unique(keys_of(rng.integers(0, V, E), rng.integers(0, V, E)))


def test_bfs_matches_scipy_shortest_paths():
    sp = pytest.importorskip("scipy.sparse")
    csgraph = pytest.importorskip("scipy.sparse.csgraph")
    rng = np.random.default_rng(0)
    for V, E in ((50, 120), (300, 900), (2000, 5000)):
        keys = random_graph(rng, V, E)
        ro, col = oracle.csr_of(keys, V)
        A = sp.csr_matrix((np.ones(col.size), col, ro.astype(np.int64)), shape=(V, V))
        for src in (0, V // 3, V - 1):
            lv = oracle.bfs(keys, V, src)
            dist = csgraph.shortest_path(A, directed=True, unweighted=True, indices=src)
            want = np.where(np.isinf(dist), 0xFFFFFFFF, dist).astype(np.uint64)
            assert np.array_equal(lv.astype(np.uint64), want)


def google_fixed_point(keys, V, alpha):
    ro, col = oracle.csr_of(keys, V)
    G = np.zeros((V, V))
    for u in range(V):
        od = ro[u + 1] - ro[u]
        if od == 0:
            G[:, u] = 1.0 / V
        else:
            for e in range(ro[u], ro[u + 1]):
                G[col[e], u] += 1.0 / od
    M = alpha * G + (1 - alpha) / V
    w, vecs = np.linalg.eig(M)
    x = np.real(vecs[:, np.argmax(np.real(w))])
    return x / x.sum()


def test_pagerank_converges_to_google_eigenvector():
    rng = np.random.default_rng(1)
    for V, E in ((6, 10), (40, 100), (120, 300)):
        keys = random_graph(rng, V, E)
        pr, delta = oracle.pagerank(keys, V, 0.85, 300)
        assert delta < 1e-12
        assert np.abs(pr - google_fixed_point(keys, V, 0.85)).sum() < 1e-9


def test_pagerank_cycle_is_uniform_and_one_step_closed_form():
    V = 7
    keys = keys_of(np.arange(V), (np.arange(V) + 1) % V)
    pr, delta = oracle.pagerank(keys, V, 0.85, 5)
    assert np.allclose(pr, 1.0 / V) and delta < 1e-15
    # star: one step from uniform, by hand
    keys = keys_of([1, 2, 3], [0, 0, 0])
    pr, _ = oracle.pagerank(keys, 4, 0.85, 1)
    # vertex 0 is dangling: gamma = (0.85*0.25 + 0.15)/4
    g = (0.85 * 0.25 + 0.15) / 4
    want = np.array([0.85 * 0.75 + g, g, g, g])
    assert np.allclose(pr, want / want.sum(), atol=1e-15)
