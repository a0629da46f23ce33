"""GPU tests of the served form: requests submitted from several threads are
served with continuous batching, and every result equals the oracle at the
point of the update sequence the server reports for it -- which is what the
same request returns when run alone on a structure at that point."""
import threading

import numpy as np
import pytest

import lpma_amd as L
import oracle
from gpu_util import keys_of, say

pytestmark = pytest.mark.gpu


def rand_keys(rng, V, n):
    return keys_of(rng.integers(0, V, n), rng.integers(0, V, n))


def adjacency(keys, V, verts):
    ro, col = oracle.csr_of(keys, V)
    offs = [0]
    vals = []
    for v in verts:
        v = int(v)
        if v < V:
            vals.append(col[ro[v]:ro[v + 1]])
            offs.append(offs[-1] + int(ro[v + 1] - ro[v]))
        else:
            offs.append(offs[-1])
    return np.array(offs, dtype=np.uint32), (np.concatenate(vals) if vals else np.empty(0, np.uint32))


def make_requests(rng, V, n, live):
    reqs = []
    for i in range(n):
        x = rng.random()
        if x < 0.6:
            k = int(rng.integers(1, 400))
            ins = rand_keys(rng, V, k // 2 + 1)
            dels = np.concatenate([live[rng.integers(0, live.size, k // 2)], rand_keys(rng, V, 3)])
            reqs.append(("update", ins, dels))
        elif x < 0.95:
            reqs.append(("read", rng.integers(0, V, int(rng.integers(1, 200))).astype(np.uint32)))
        else:
            reqs.append(("bfs", int(rng.integers(0, V))))
    return reqs


def submit_all(srv, reqs, threads=3):
    """Submits the requests from several threads in a fixed interleaving;
    returns (ticket, request) pairs in submission order."""
    lock = threading.Lock()
    order = []
    nxt = [0]

    def run():
        while True:
            with lock:
                i = nxt[0]
                if i >= len(reqs):
                    return
                nxt[0] += 1
                r = reqs[i]
                if r[0] == "update":
                    t = srv.submit_update(r[1], r[2])
                elif r[0] == "read":
                    t = srv.submit_read(r[1])
                else:
                    t = srv.submit_bfs(r[1], levels=True)
                order.append((t, r))

    ts = [threading.Thread(target=run) for _ in range(threads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    return order


def check_results(srv, order, V, initial):
    # oracle states after each update request, in submission (= ticket) order
    states = [initial]
    for t, r in order:
        if r[0] == "update":
            states.append(oracle.apply(states[-1], r[1], r[2]))
    results = {t: srv.wait(t) for t, _ in order}
    n_upd = 0
    for t, r in order:
        res = results[t]
        assert res["status"] == 0, (r[0], res["status"])
        if r[0] == "update":
            n_upd += 1
            assert res["applied"] == n_upd  # applied in submission order
        elif r[0] == "read":
            offs, vals = adjacency(states[res["applied"]], V, r[1])
            assert np.array_equal(res["offsets"], offs)
            assert np.array_equal(res["values"], vals)
        else:
            want = oracle.bfs(states[res["applied"]], V, r[1])
            assert np.array_equal(res["values"], want)
    return states, results


def test_served_mix_equals_oracle_at_each_version():
    say("server mix")
    rng = np.random.default_rng(50)
    V = 2000
    initial = oracle.sort_unique(rand_keys(rng, V, 20000))
    pma = L.PMA(V)
    pma.insert(initial)
    srv = L.Server(pma, bfs_workers=3)
    reqs = make_requests(rng, V, 400, initial)
    order = submit_all(srv, reqs)
    states, results = check_results(srv, order, V, initial)
    st = srv.stats()
    n_upd = sum(1 for _, r in order if r[0] == "update")
    assert st["completed"] == st["submitted"]
    assert st["applied_updates"] == n_upd
    assert st["update_runs"] <= n_upd
    srv.close()
    assert np.array_equal(L.to_host_u64(pma.export()[0]), states[-1])
    assert pma.check() == 0

    # the same reads alone, on a structure holding the state they saw
    for t, r in order[:200]:
        if r[0] == "update":
            continue
        ref = states[results[t]["applied"]]
        alone = L.PMA(V)
        alone.insert(ref)
        if r[0] == "read":
            import torch
            offs, vals = alone.neighbors(torch.from_numpy(r[1].astype(np.int32)))
            assert np.array_equal(offs.cpu().numpy().view(np.uint32), results[t]["offsets"])
            assert np.array_equal(vals.cpu().numpy().view(np.uint32), results[t]["values"])
        else:
            lv, _, _ = alone.bfs(r[1])
            assert np.array_equal(lv.cpu().numpy().view(np.uint32), results[t]["values"])
        alone.close()
    pma.close()


def test_merged_updates_keep_submission_order():
    """A burst of conflicting updates on few keys: the final set is the
    sequential application, whatever runs the server merged them into."""
    say("server conflicting burst")
    rng = np.random.default_rng(51)
    V = 64
    pool = rand_keys(rng, V, 100)
    pma = L.PMA(V)
    srv = L.Server(pma, bfs_workers=1)
    reqs = []
    for i in range(300):
        reqs.append(("update", pool[rng.integers(0, 100, 20)], pool[rng.integers(0, 100, 20)]))
    order = submit_all(srv, reqs, threads=1)
    states, _ = check_results(srv, order, V, np.empty(0, np.uint64))
    st = srv.stats()
    srv.close()
    assert np.array_equal(L.to_host_u64(pma.export()[0]), states[-1])
    assert st["update_runs"] < 300  # requests were merged
    pma.close()


def test_bfs_admission_by_budget():
    """With room for one snapshot and one search, searches run one at a time
    (they wait for memory, they do not fail); one that can never fit fails."""
    say("server budget")
    rng = np.random.default_rng(52)
    V = 5000
    initial = oracle.sort_unique(rand_keys(rng, V, 40000))
    pma = L.PMA(V)
    pma.insert(initial)
    snap = int(L.lib().lpma_snapshot_bytes(pma._h))
    bfs = 24 * (V + 2) + (1 << 20)
    srv = L.Server(pma, bfs_workers=4, hbm_budget_bytes=snap + bfs + bfs // 2)
    order = submit_all(srv, [("bfs", int(r)) for r in rng.integers(0, V, 12)], threads=2)
    check_results(srv, order, V, initial)
    st = srv.stats()
    assert st["max_concurrent_bfs"] == 1 and st["admission_waits"] > 0
    srv.close()
    tiny = L.Server(pma, bfs_workers=2, hbm_budget_bytes=bfs)
    t = tiny.submit_bfs(0)
    assert tiny.wait(t)["status"] == 3  # LPMA_ERR_OUT_OF_MEMORY
    tiny.close()
    pma.close()


def test_invalid_update_fails_alone():
    say("server invalid update")
    rng = np.random.default_rng(53)
    V = 100
    pma = L.PMA(V)
    srv = L.Server(pma)
    good1 = rand_keys(rng, V, 50)
    bad = keys_of([V + 5], [1])
    good2 = rand_keys(rng, V, 50)
    t1, t2, t3 = srv.submit_update(good1), srv.submit_update(bad), srv.submit_update(good2)
    r1, r2, r3 = srv.wait(t1), srv.wait(t2), srv.wait(t3)
    assert r1["status"] == 0 and r2["status"] == 2 and r3["status"] == 0
    srv.close()
    assert np.array_equal(L.to_host_u64(pma.export()[0]), oracle.apply(oracle.apply(np.empty(0, np.uint64), good1),
                                                                      good2))
    pma.close()
