"""Parity at the full sizes of the benchmark workloads.

The reference checks its update stream at the benchmark's own size
(test_update_rpma.cu:131-165: the whole stream, then the sorted keys against
the host's).  These tests do the same for every workload the benchmark runs:

* R-MAT, 2^20 vertices, 2^24 distinct edges, half bulk-loaded, the rest
  streamed as inserts and the same volume of live edges deleted: batches of
  10^5 (all 168) and 10^6 (all 18) with the export, the CSR offsets and
  get_data compared with the oracle after every batch; batches of 10^3, the
  first 1000 inserts and 200 deletes, with the CSR offsets and every row the
  batch touched compared after every batch and the whole export every 100.
* 64 PMAs, edges log-uniform in [10^4, 10^7], 100 rounds of lognormal
  batches (70 % inserts): after every round every member's edge count and
  CSR offsets, and a rotating eighth of the members in full; all members in
  full at the end.
* The served workload, R-MAT 2^24 vertices from 2^28 samples: a Poisson
  trace through the server; every read and every BFS equals the oracle at
  the number of update requests the server reports it saw.
* Analytics, R-MAT 2^24 vertices from 2^29 samples, batches of 10^6: the edge
  set after every batch; BFS levels bit-exact and PageRank (run to its
  converged iteration count) within 1e-6 L1 of the FP64 oracle.
"""
import os

import numpy as np
import pytest
import torch

import lpma_amd as L
import oracle
from gpu_util import check_get_data, say
from lpma_amd.workloads import BatchedWorkload, rmat_stream

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, len(os.sched_getaffinity(0))))
MASK32 = np.uint64(0xFFFFFFFF)
U32 = np.uint64(32)


def full_compare(pma, ref, V):
    torch.cuda.synchronize()
    assert pma.edges == ref.size, (pma.edges, ref.size)
    ek, ro = pma.export()
    got = L.to_host_u64(ek)
    del ek
    assert got.size == ref.size
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, f"edges differ at {bad[:5]}: got {got[bad[:5]]}, want {ref[bad[:5]]}"
    want_ro = oracle.row_offsets(ref, V, threads=THREADS)
    assert np.array_equal(ro.cpu().numpy().view(np.uint32), want_ro)
    check_get_data(pma, ref, want_ro, V)
    assert pma.check() == 0


def rows_of(ref, ro, srcs):
    """The oracle's neighbor lists of `srcs` (offsets, dst ids)."""
    lens = (ro[srcs + 1] - ro[srcs]).astype(np.int64)
    offs = np.zeros(srcs.size + 1, dtype=np.int64)
    np.cumsum(lens, out=offs[1:])
    idx = np.repeat(ro[srcs].astype(np.int64) - offs[:-1], lens) + np.arange(offs[-1], dtype=np.int64)
    return offs, (ref[idx] & MASK32).astype(np.uint32)


def light_compare(pma, ref, V, batch):
    """CSR offsets of every row, and the neighbor lists of the rows the
    batch touched (the only rows whose contents can change)."""
    want_ro = oracle.row_offsets(ref, V, threads=THREADS)
    assert np.array_equal(pma.row_offsets().cpu().numpy().view(np.uint32), want_ro)
    srcs = np.unique((np.asarray(batch, dtype=np.uint64) >> U32).astype(np.int64))
    offs, vals = pma.neighbors(torch.from_numpy(srcs.astype(np.int32)))
    w_offs, w_vals = rows_of(ref, want_ro, srcs)
    assert np.array_equal(offs.cpu().numpy().astype(np.int64), w_offs)
    assert np.array_equal(vals.cpu().numpy().view(np.uint32), w_vals)


def run_stream(B, steps=None, full_every=1):
    V = 1 << 20
    bulk, ins, dels = rmat_stream(batch=B)
    seq = [(1, i) for i in range(len(ins))] + [(0, i) for i in range(len(dels))]
    if steps is not None:
        seq = steps(seq, len(ins))
    pma = L.PMA(V)
    pma.bulk_load(L.to_device(bulk))
    ref = oracle.sort_unique(bulk)
    full_compare(pma, ref, V)
    say(f"stream B={B}: bulk load of {bulk.size} matches; {len(seq)} batches")
    for j, (op, i) in enumerate(seq):
        b = ins[i] if op else dels[i]
        pma.update(L.to_device(b), op)
        ref = oracle.apply(ref, b, (), threads=THREADS) if op else oracle.apply(ref, (), b, threads=THREADS)
        if j % full_every == 0 or j == len(seq) - 1:
            full_compare(pma, ref, V)
        else:
            light_compare(pma, ref, V, b)
        if j % 100 == 0:
            say(f"  batch {j}: {ref.size} edges, capacity {pma.view().capacity}")
    s = pma.stats()
    pma.close()
    return ref, s


@pytest.mark.timeout(900)
def test_rmat20_stream_batch_1e5_every_batch():
    ref, s = run_stream(100000)
    # all inserts then as many deletes: the 2^24 - 2^23 edges the deletes missed
    assert ref.size > 0 and s.rebuilds >= 1


@pytest.mark.timeout(900)
def test_rmat20_stream_batch_1e6_every_batch():
    run_stream(1000000)


@pytest.mark.timeout(900)
def test_rmat20_stream_batch_1e3_window():
    run_stream(1000, steps=lambda seq, n_ins: seq[:1000] + seq[n_ins:n_ins + 200], full_every=100)


# ------------------------------------------------------------- batched form

@pytest.mark.timeout(900)
def test_batched_64_pmas_100_rounds():
    w = BatchedWorkload()
    say(f"batched: {w.n} graphs, {sum(w.edges)} edges, {w.total_updates} updates in {w.rounds} rounds")
    # the configuration the autotuner picks for this workload: 64-slot
    # segments, lower 0.08 at the leaves to 0.30 at the root, upper 1.0 to 0.75
    fleet = [L.PMA(V, segment_slots=64, density_lower_leaf=0.08, density_lower=0.30, density_upper_leaf=1.0,
                   density_upper=0.75) for V in w.V]
    refs = []
    for p, init in zip(fleet, w.initial):
        p.bulk_load(L.to_device(init))
        refs.append(oracle.sort_unique(init))
    group = L.Group(fleet)
    for i in range(0, w.n, 8):
        full_compare(fleet[i], refs[i], w.V[i])
    for r in range(w.rounds):
        ins = [w.ins[i][r] for i in range(w.n)]
        dels = [w.dels[i][r] for i in range(w.n)]
        group.update(ins, dels)
        for i in range(w.n):
            refs[i] = oracle.apply(refs[i], ins[i], dels[i], threads=THREADS)
        torch.cuda.synchronize()
        for i in range(w.n):
            assert fleet[i].edges == refs[i].size, (r, i)
            if (i + r) % 8 == 0:
                full_compare(fleet[i], refs[i], w.V[i])
            else:
                want_ro = oracle.row_offsets(refs[i], w.V[i], threads=THREADS)
                assert np.array_equal(fleet[i].row_offsets().cpu().numpy().view(np.uint32), want_ro), (r, i)
        if r % 20 == 0:
            say(f"  round {r}: {sum(x.size for x in refs)} edges")
    for i in range(w.n):
        full_compare(fleet[i], refs[i], w.V[i])
    group.close()
    for p in fleet:
        p.close()
    L.pool_trim()


# ------------------------------------------------------------- served form

def restricted_rows(S, verts):
    """The keys of S whose source is in `verts` (sorted)."""
    return S[np.isin((S >> U32).astype(np.int64), verts)]


@pytest.mark.timeout(900)
def test_served_rmat24_trace_matches_oracle():
    """Reads are checked against the oracle restricted to the rows they
    read (applying a batch commutes with restricting to a set of rows);
    every BFS against the oracle's full edge set at its version."""
    scale, V = 24, 1 << 24
    samples = L.gen_rmat_device(scale, 1 << 28, seed=4)
    pma = L.PMA(V)
    pma.bulk_load(samples)
    host = L.to_host_u64(samples)
    del samples
    S0 = oracle.sort_unique(host)
    del host
    full_compare(pma, S0, V)
    say(f"served: {S0.size} edges after the bulk load")
    rng = np.random.default_rng(44)
    ins_pool = L.to_host_u64(L.gen_rmat_device(scale, 1 << 22, seed=4004))
    del_pool = S0[rng.integers(0, S0.size, 1 << 22)]
    reqs = []
    t = 0.0
    for k in range(90):
        t += rng.exponential(1.0 / 1500)
        kind = rng.choice(3, p=(0.60, 0.35, 0.05)) if k >= 3 else k  # one of each kind first
        if kind == 0:
            n = max(2, int(rng.lognormal(np.log(1e4) - 0.5, 1.0)))
            reqs.append((t, "update", ins_pool[rng.integers(0, ins_pool.size, n // 2)],
                         del_pool[rng.integers(0, del_pool.size, n - n // 2)]))
        elif kind == 1:
            reqs.append((t, "read", rng.integers(0, V, int(rng.integers(1, 1025))).astype(np.uint32)))
        else:
            reqs.append((t, "bfs", int(rng.integers(0, V))))
    srv = L.Server(pma)
    import time
    t0 = time.perf_counter()
    tickets = []
    for (at, kind, *payload) in reqs:
        while time.perf_counter() - t0 < at:
            pass
        if kind == "update":
            tickets.append(srv.submit_update(payload[0], payload[1]))
        elif kind == "read":
            tickets.append(srv.submit_read(payload[0]))
        else:
            tickets.append(srv.submit_bfs(payload[0], levels=True))
    results = [srv.wait(tk) for tk in tickets]
    srv.close()
    updates = [(r[2], r[3]) for r in reqs if r[1] == "update"]
    assert all(res["status"] == 0 for res in results)

    # reads: the rows of every read vertex, replayed through the updates
    reads = [(res["applied"], req[2], res) for req, res in zip(reqs, results) if req[1] == "read"]
    verts = np.unique(np.concatenate([v for _, v, _ in reads]).astype(np.int64))
    R = restricted_rows(S0, verts)
    Rro = None
    applied = 0
    checked = 0
    for a, vs, res in sorted(reads, key=lambda x: x[0]):
        while applied < a:
            i_, d_ = updates[applied]
            R = oracle.apply(R, restricted_rows(i_, verts), restricted_rows(d_, verts))
            applied += 1
            Rro = None
        if Rro is None:
            Rro = oracle.row_offsets(R, V, threads=THREADS)
        offs, vals = rows_of(R, Rro, vs.astype(np.int64))
        assert np.array_equal(res["offsets"].astype(np.int64), offs), a
        assert np.array_equal(res["values"], vals), a
        checked += 1
    say(f"served: {checked} reads match the oracle at their versions")

    # searches: the oracle's full edge set at each search's version
    bfs = sorted([(res["applied"], req[2], res) for req, res in zip(reqs, results) if req[1] == "bfs"],
                 key=lambda x: x[0])
    S = S0
    applied = 0
    for a, root, res in bfs:
        while applied < a:
            S = oracle.apply(S, updates[applied][0], updates[applied][1], threads=THREADS)
            applied += 1
        assert np.array_equal(res["values"], oracle.bfs(S, V, root)), (a, root)
    say(f"served: {len(bfs)} searches match the oracle at their versions")
    assert checked >= 5 and len(bfs) >= 1
    while applied < len(updates):
        S = oracle.apply(S, updates[applied][0], updates[applied][1], threads=THREADS)
        applied += 1
    full_compare(pma, S, V)
    pma.close()
    L.pool_trim()


# ---------------------------------------------------------------- analytics

@pytest.mark.timeout(900)
def test_analytics_rmat24_full_size():
    scale, V, B = 24, 1 << 24, 1000000
    samples = L.gen_rmat_device(scale, 1 << 29, seed=5)
    pma = L.PMA(V)
    pma.bulk_load(samples)
    host = L.to_host_u64(samples)
    del samples
    S = oracle.sort_unique(host)
    del host
    say(f"analytics: {S.size} edges, capacity {pma.view().capacity}")
    fresh = L.to_host_u64(L.gen_rmat_device(scale, 2 * B, seed=5005))
    rng = np.random.default_rng(55)
    for b in range(4):
        if b % 2 == 0:
            batch = fresh[(b // 2) * B:(b // 2 + 1) * B]
            pma.insert(L.to_device(batch))
            S = oracle.apply(S, batch, (), threads=THREADS)
        else:
            batch = S[rng.integers(0, S.size, B)]
            pma.delete(L.to_device(batch))
            S = oracle.apply(S, (), batch, threads=THREADS)
        torch.cuda.synchronize()
        assert pma.edges == S.size
        if b in (0, 3):
            full_compare(pma, S, V)
            lv, depth, _ = pma.bfs(0)
            assert np.array_equal(lv.cpu().numpy().view(np.uint32), oracle.bfs(S, V, 0)), b
            pr, iters, delta = pma.pagerank(0.85, 1e-6, 50)
            assert delta < 1e-6 and iters < 50
            ref_pr, _ = oracle.pagerank(S, V, 0.85, iters)
            l1 = float(np.abs(pr.cpu().numpy().astype(np.float64) - ref_pr).sum())
            say(f"analytics batch {b}: BFS depth {depth} exact, PageRank {iters} iterations, L1 {l1:.2e}")
            assert l1 < 1e-6, l1
    pma.close()
    L.pool_trim()
