"""CPU tests that pin the oracle before it is trusted as the GPU checker.

* the set contract (oracle.apply) against Python set algebra on small random
  cases (brute force), against numpy's union1d/setdiff1d on medium ones, and
  across thread counts;
* row offsets against numpy bincount;
* the PMA model: its contents equal the set contract after every batch, its
  layout satisfies the structural invariants, its spread equals the
  reference's formula (rpma.cuh:676-679) written out independently, and its
  layout matches hand-derived golden cases (tests/golden/layout_cases.json);
* planted mistakes in oracle outputs are caught by the comparisons the GPU
  tests use.
"""
import json
import os

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
U32 = np.uint64(32)


def keys_of(s, d):
    return (np.asarray(s, dtype=np.uint64) << U32) | np.asarray(d, dtype=np.uint64)


def set_apply(S, I, D):
    return np.array(sorted((set(int(x) for x in S) - set(int(x) for x in D)) | set(int(x) for x in I)),
                    dtype=np.uint64)


def test_apply_brute_force_small():
    rng = np.random.default_rng(0)
    for trial in range(400):
        V = int(rng.integers(1, 6))
        S = oracle.sort_unique(keys_of(rng.integers(0, V, 12), rng.integers(0, V, 12)))
        n_i = int(rng.integers(0, 8))
        n_d = int(rng.integers(0, 8))
        I = keys_of(rng.integers(0, V, n_i), rng.integers(0, V, n_i))
        D = keys_of(rng.integers(0, V, n_d), rng.integers(0, V, n_d))
        got = oracle.apply(S, I, D)
        assert np.array_equal(got, set_apply(S, I, D)), (S, I, D, got)


def test_apply_against_numpy_medium():
    rng = np.random.default_rng(1)
    V = 1 << 10
    S = oracle.sort_unique(keys_of(rng.integers(0, V, 200000), rng.integers(0, V, 200000)))
    I = keys_of(rng.integers(0, V, 50000), rng.integers(0, V, 50000))
    D = np.concatenate([S[::7], keys_of(rng.integers(0, V, 1000), rng.integers(0, V, 1000))])
    want = np.union1d(np.setdiff1d(S, D), I)
    for threads in (1, 2, 4, 7):
        assert np.array_equal(oracle.apply(S, I, D, threads=threads), want)
        # into a caller's buffer (reused, stale contents beyond the result)
        buf = np.full(S.size + I.size + 5, 7, dtype=np.uint64)
        assert np.array_equal(oracle.apply(S, I, D, threads=threads, out=buf), want)
    with pytest.raises(ValueError):
        oracle.apply(S, I, D, out=np.empty(S.size, dtype=np.uint64))  # too small
    with pytest.raises(ValueError):
        big = np.concatenate([S, np.zeros(I.size + 1, dtype=np.uint64)])
        oracle.apply(big[:S.size], I, D, out=big)  # overlaps S


def test_apply_edges_of_domain():
    S = oracle.sort_unique(keys_of([0, 1, 2], [5, 5, 5]))
    assert oracle.apply(S).tolist() == S.tolist()                   # empty batch is identity
    assert oracle.apply(S, (), S).size == 0                         # delete everything
    assert oracle.apply(np.empty(0, np.uint64), S).tolist() == S.tolist()  # insert into empty
    both = keys_of([9], [9])
    assert oracle.apply(S, both, both).tolist() == sorted(S.tolist() + both.tolist())  # delete then insert


def test_planted_mistakes_are_caught():
    """The GPU tests compare with array_equal; a dropped, duplicated or
    swapped key in the oracle's output must make that comparison fail."""
    rng = np.random.default_rng(2)
    S = oracle.sort_unique(keys_of(rng.integers(0, 64, 500), rng.integers(0, 64, 500)))
    good = oracle.apply(S, keys_of([1], [2]), S[:3])
    for mutate in (lambda a: a[1:], lambda a: np.insert(a, 5, a[5]),
                   lambda a: np.concatenate([a[:3], a[4:5], a[3:4], a[5:]]), lambda a: a + np.uint64(1)):
        bad = mutate(good.copy())
        assert not np.array_equal(bad, good)
        assert not (bad.size == good.size and oracle.is_sorted_unique(bad) and np.array_equal(bad, good))


def test_row_offsets():
    rng = np.random.default_rng(3)
    V = 100
    S = oracle.sort_unique(keys_of(rng.integers(0, V, 3000), rng.integers(0, V, 3000)))
    ro = oracle.row_offsets(S, V)
    cnt = np.bincount((S >> U32).astype(np.int64), minlength=V)
    assert ro[0] == 0 and np.array_equal(np.diff(ro.astype(np.int64)), cnt)
    # markers are skipped
    withm = np.union1d(S, np.array([oracle.marker_key(r) for r in range(V)], dtype=np.uint64))
    assert np.array_equal(oracle.row_offsets(withm, V), ro)
    with pytest.raises(ValueError):
        oracle.row_offsets(keys_of([V], [0]), V)
    # the threaded form (rows split over threads) gives the same offsets and error
    for threads in (2, 3, 7, 16):
        assert np.array_equal(oracle.row_offsets(S, V, threads=threads), ro)
        assert np.array_equal(oracle.row_offsets(withm, V, threads=threads), ro)
        with pytest.raises(ValueError):
            oracle.row_offsets(keys_of([3, V], [0, 0]), V, threads=threads)
    # empty rows at the ends and a single hub row
    hub = keys_of(np.full(50, 40), np.arange(50))
    for threads in (1, 4):
        got = oracle.row_offsets(hub, V, threads=threads)
        assert got[40] == 0 and got[41] == 50 and got[V] == 50


def model_check(m, V, S):
    """Structural invariants of the model's layout."""
    k = m.keys()
    c = m.cnt()
    sep = m.sep()
    nseg = 1 << m.height
    assert k.size == nseg * S
    assert sep[0] == 0
    live = []
    for s in range(nseg):
        seg = k[s * S:(s + 1) * S]
        assert c[s] <= S
        assert (seg[: c[s]] != oracle.KEY_NONE).all() and (seg[c[s]:] == oracle.KEY_NONE).all()
        hi = sep[s + 1] if s + 1 < nseg else None
        for x in seg[: c[s]]:
            assert x >= sep[s] and (hi is None or x < hi)
        live.extend(seg[: c[s]].tolist())
    assert oracle.is_sorted_unique(np.array(live, dtype=np.uint64))
    ro = m.row_offsets()
    for r in range(V):
        assert k[ro[r + 1]] == oracle.marker_key(r)
    assert m.live == len(live)


@pytest.mark.parametrize("S", [2, 4, 32, 64])
def test_model_contents_follow_contract(S):
    rng = np.random.default_rng(10 + S)
    V = 200
    m = oracle.PmaModel(V, S)
    ref = np.empty(0, dtype=np.uint64)
    markers = np.array([oracle.marker_key(r) for r in range(V)], dtype=np.uint64)
    for b in range(12):
        I = keys_of(rng.integers(0, V, 300), rng.integers(0, V, 300))
        if b % 3 == 2:
            D = ref[rng.integers(0, max(ref.size, 1), 200)] if ref.size else np.empty(0, np.uint64)
            m.update_mixed(I, D)
            ref = oracle.apply(ref, I, D)
        else:
            m.insert(I)
            ref = oracle.apply(ref, I)
        assert np.array_equal(m.export(), np.union1d(ref, markers))
        model_check(m, V, S)
    while ref.size:
        D = ref[: 500]
        m.delete(D)
        ref = oracle.apply(ref, (), D)
        assert np.array_equal(m.export(), np.union1d(ref, markers))
        model_check(m, V, S)


def test_model_brute_force_tiny():
    """Tiny universes, many random batch sequences: contents match the set
    contract and the layout stays valid."""
    rng = np.random.default_rng(4)
    for trial in range(150):
        V = int(rng.integers(1, 5))
        S = int(rng.choice([2, 4]))
        m = oracle.PmaModel(V, S)
        ref = np.empty(0, dtype=np.uint64)
        markers = np.array([oracle.marker_key(r) for r in range(V)], dtype=np.uint64)
        for b in range(6):
            ni, nd = int(rng.integers(0, 5)), int(rng.integers(0, 5))
            I = keys_of(rng.integers(0, V, ni), rng.integers(0, V, ni))
            D = keys_of(rng.integers(0, V, nd), rng.integers(0, V, nd))
            m.update_mixed(I, D)
            ref = oracle.apply(ref, I, D)
            assert np.array_equal(m.export(), np.union1d(ref, markers))
            model_check(m, V, S)


def reference_spread(sorted_keys, nsegw, S):
    """rpma.cuh:676-679 written out: element i goes to segment
    frac*i/deno and lane (frac*i % deno)/frac with frac = W/S, deno = m."""
    m = len(sorted_keys)
    out = np.full(nsegw * S, oracle.KEY_NONE, dtype=np.uint64)
    frac = nsegw
    for i, k in enumerate(sorted_keys):
        seg = (frac * i) // m
        lane = ((frac * i) % m) // frac
        out[seg * S + lane] = k
    return out


@pytest.mark.parametrize("V,S", [(5, 2), (40, 4), (100, 32), (1000, 32), (77, 8)])
def test_model_initial_spread_is_reference_formula(V, S):
    m = oracle.PmaModel(V, S)
    markers = [oracle.marker_key(r) for r in range(V)]
    assert np.array_equal(m.keys(), reference_spread(markers, 1 << m.height, S))
    # a resize re-spreads everything with the same formula
    rng = np.random.default_rng(V)
    I = oracle.sort_unique(keys_of(rng.integers(0, V, 20 * V), rng.integers(0, V, 20 * V)))
    m.insert(I)
    if m.rebuilds >= 2:
        allk = np.union1d(I, np.array(markers, dtype=np.uint64))
        assert np.array_equal(m.keys(), reference_spread(allk.tolist(), 1 << m.height, S))


def test_model_golden_layouts():
    with open(os.path.join(HERE, "golden", "layout_cases.json")) as f:
        cases = json.load(f)["cases"]
    for case in cases:
        m = oracle.PmaModel(case["V"], case["S"])
        for step in case["steps"]:
            keys = keys_of([e[0] for e in step["keys"]], [e[1] for e in step["keys"]])
            if step["op"] == "insert":
                m.insert(keys)
            else:
                m.delete(keys)
            want = [oracle.KEY_NONE if x is None else
                    (oracle.marker_key(x[1]) if x[0] == "M" else (x[0] << 32) | x[1]) for x in step["layout"]]
            assert m.height == step["height"], case["name"]
            assert m.keys().tolist() == want, (case["name"], step)
