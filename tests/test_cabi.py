"""CPU tests of the C ABI library: it loads, exports every entry point that
include/lpma.h declares, rejects bad arguments before touching a device, and
its host-side input generators are correct and reproducible.  No test here
launches device work."""
import os
import re
import tempfile

import numpy as np
import pytest

import lpma_amd as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
U32 = np.uint64(32)


def declared_functions():
    text = open(os.path.join(ROOT, "include", "lpma.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lpma_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = L.lib()
    names = declared_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    bound = {s[0] for s in L.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_strerror_and_defaults():
    lib = L.lib()
    for code in range(7):
        assert lib.lpma_strerror(code)
    assert b"unknown" in lib.lpma_strerror(99)
    cfg = L.Config()
    lib.lpma_config_default(cfg)
    assert cfg.segment_slots == 32  # rpma.cuh:1769
    assert cfg.density_lower_leaf == cfg.density_lower_root == 0.42  # rpma.cuh:76-79
    assert cfg.density_upper_leaf == cfg.density_upper_root == 1.0


def test_null_arguments_are_rejected_without_a_device():
    lib = L.lib()
    assert lib.lpma_create(None, 10, None, None) == 1
    assert lib.lpma_destroy(None) == 1
    assert lib.lpma_update(None, None, 0, 1, None) == 1
    assert lib.lpma_update_mixed(None, None, 0, None, 0, None) == 1
    assert lib.lpma_get_view(None, None) == 1
    assert lib.lpma_export_compacted(None, None, None, None) == 1
    assert lib.lpma_edge_query(None, None, 0, None, None) == 1
    assert lib.lpma_get_stats(None, None) == 1
    assert lib.lpma_set_profiling(None, 1) == 1
    assert lib.lpma_group_create(None, None, 0) == 1
    assert lib.lpma_group_destroy(None) == 1
    assert lib.lpma_group_update(None, None, None, None, None, None) == 1
    assert lib.lpma_group_get_stats(None, None) == 1
    assert lib.lpma_group_partitions(None) == 0
    assert lib.lpma_server_create(None, None, None) == 1
    assert lib.lpma_server_destroy(None) == 1
    assert lib.lpma_server_submit_bfs(None, 0, 0, None) == 1
    assert lib.lpma_server_wait(None, 1, None) == 1
    assert lib.lpma_snapshot_create(None, None, None) == 1
    assert lib.lpma_snapshot_destroy(None) == 1
    assert lib.lpma_snapshot_bytes(None) == 0
    assert lib.lpma_update_sequence(None, None, None, None, None, 0, None) == 1


def test_mt19937_64_known_answers():
    """std::mt19937_64 with the default seed 5489: the first output and the
    10000th output, which C++11 [rand.predef] fixes to 9981545732273789042."""
    x = L.mt64(5489, 10000)
    assert int(x[0]) == 14514284786278117030
    assert int(x[-1]) == 9981545732273789042


def test_uniform_generator():
    V = 1 << 10
    a = L.gen_uniform(V, 50000, 1)
    assert np.array_equal(a, L.gen_uniform(V, 50000, 1))
    assert not np.array_equal(a, L.gen_uniform(V, 50000, 2))
    assert np.unique(a).size == a.size
    src, dst = a >> U32, a & np.uint64(0xFFFFFFFF)
    assert (src < V).all() and (dst < V).all() and (src != dst).all()
    # roughly uniform sources
    h = np.bincount(src.astype(np.int64), minlength=V)
    assert h.min() > 0.5 * h.mean() and h.max() < 1.6 * h.mean()


def test_rmat_generator_quadrants():
    scale = 12
    a = L.gen_rmat(scale, 100000, 0.57, 0.19, 0.19, 2)
    assert np.unique(a).size == a.size
    src, dst = a >> U32, a & np.uint64(0xFFFFFFFF)
    assert (src < (1 << scale)).all() and (dst < (1 << scale)).all() and (src != dst).all()
    top_s = (src >> np.uint64(scale - 1)) & np.uint64(1)
    top_d = (dst >> np.uint64(scale - 1)) & np.uint64(1)
    frac = [np.mean((top_s == i) & (top_d == j)) for i in (0, 1) for j in (0, 1)]
    # top-level quadrant frequencies stay close to (a, b, c, d) despite dedup
    for f, p in zip(frac, (0.57, 0.19, 0.19, 0.05)):
        assert abs(f - p) < 0.05, frac
    assert np.array_equal(a, L.gen_rmat(scale, 100000, 0.57, 0.19, 0.19, 2))


def test_rmat_samples_are_the_draws_before_dedup():
    """gen_rmat keeps the first occurrence of each sampled edge, so it must
    equal the first-occurrence order of gen_rmat_samples from the same seed;
    samples keep repeats and follow the quadrant probabilities."""
    scale = 8
    s = L.gen_rmat_samples(scale, 60000, 0.57, 0.19, 0.19, 5)
    assert np.array_equal(s, L.gen_rmat_samples(scale, 60000, 0.57, 0.19, 0.19, 5))
    src, dst = s >> U32, s & np.uint64(0xFFFFFFFF)
    assert (src < (1 << scale)).all() and (dst < (1 << scale)).all() and (src != dst).all()
    uniq, first = np.unique(s, return_index=True)
    assert uniq.size < s.size  # repeats kept
    order = s[np.sort(first)]
    d = L.gen_rmat(scale, order.size, 0.57, 0.19, 0.19, 5)
    assert np.array_equal(d, order)
    top_s = (src >> np.uint64(scale - 1)) & np.uint64(1)
    top_d = (dst >> np.uint64(scale - 1)) & np.uint64(1)
    frac = [np.mean((top_s == i) & (top_d == j)) for i in (0, 1) for j in (0, 1)]
    for f, p in zip(frac, (0.57, 0.19, 0.19, 0.05)):
        assert abs(f - p) < 0.03, frac


def test_shuffle_and_absent():
    a = L.gen_uniform(1 << 8, 5000, 3)
    b = L.shuffle(a.copy(), 9)
    assert np.array_equal(np.sort(a), np.sort(b)) and not np.array_equal(a, b)
    assert np.array_equal(b, L.shuffle(a.copy(), 9))
    present = np.sort(a)
    n = L.gen_absent(1 << 8, present, 3000, 4)
    assert np.unique(n).size == n.size
    assert not np.isin(n, present).any()


def test_edgelist_reader_round_trip():
    a = L.gen_uniform(100, 500, 5)
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "g.edgelist")
        with open(p, "w") as f:
            f.write(f"100 {a.size}\n")
            for k in a:
                f.write(f"{int(k) >> 32} {int(k) & 0xFFFFFFFF}\n")
        V, keys = L.read_edgelist(p)
        assert V == 100 and np.array_equal(keys, a)
        V, keys = L.read_edgelist(p, max_edges=10)
        assert keys.size == 10 and np.array_equal(keys, a[:10])
        with open(p, "w") as f:
            f.write("10 1\n3 12\n")
        with pytest.raises(L.LpmaError):
            L.read_edgelist(p)
