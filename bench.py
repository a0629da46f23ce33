#!/usr/bin/env python3
"""Benchmark of the batched edge-update path on one MI355X.

Headline workload: R-MAT graph, 2^20 vertices, 2^24 distinct edges
(a,b,c,d = 0.57,0.19,0.19,0.05, seed 2, self loops and repeats rejected),
shuffled with seed 12; the first half is bulk-loaded, the second half is
streamed as insert batches of B edges, then the same volume of live edges
(a seeded uniform sample without replacement of all 2^24) is deleted in
batches of B.  A step is one update batch.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
    python bench.py --workload batched|served|analytics

Prints one JSON line.  `value` is edge updates per second over the WHOLE
stream (84 insert + 84 delete batches at B = 1e5; `steps` reports the batches
timed, whatever --steps asks: the stream's state depends on every batch
before it), inputs resident in HBM, wall clock between device
synchronisations, from a fresh bulk load; W warm-up batches run first on a
separate instance.  p50/p99 are per-batch HIP-event spans of that pass.
`roofline` describes the dominant rebalance kernel, the one that moves the
most algorithmic bytes (bytes per launch / average launch time, from HIP
events the library records around each launch on the update stream in a
second, profiled pass over the same stream).  `cpu_baseline` times the CPU
oracle (sorted merge + row offsets, pthreads on every host core the process
may use, and on one) on the stream's first insert batches from the
bulk-loaded state and its first delete batches from the state the GPU
deletes from.  Unless --legs is empty, the line also carries two compact
legs measured in the same run: `batched` (64 PMAs, one group update per
round, against one update call per PMA per round) and `served` (the request
server against a one-request-at-a-time loop at the top offered rate).

`--workload batched` measures the batched form alone at full length: 64
independent R-MAT graphs (edge counts log-uniform in [1e4, 1e7], vertices =
edges/16 rounded to a power of two, seeds 3+i), 100 rounds in which every
graph gets one batch (lognormal size, mu = ln 5000, sigma = 1.5, cap 1e6, 70%
inserts / 30% deletes).  `--workload served` sweeps the served form over
offered rates; `--workload analytics` runs PageRank and BFS on the live
structure.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
ROOT = os.path.dirname(os.path.abspath(__file__))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_workload(scale, edges, B):
    from lpma_amd.workloads import rmat_stream
    t = time.time()
    bulk, ins_batches, del_batches = rmat_stream(scale=scale, edges=edges, seed=2, shuffle_seed=12, batch=B,
                                                 del_seed=13)
    log(f"workload ready in {time.time() - t:.1f}s: bulk {bulk.size}, {len(ins_batches)} insert + "
        f"{len(del_batches)} delete batches of {B}")
    return bulk, ins_batches, del_batches


def stream_steps(ins_batches, del_batches):
    """(op, batch index) of the whole stream: every insert batch, then every
    delete batch."""
    return [(1, i) for i in range(len(ins_batches))] + [(0, i) for i in range(len(del_batches))]


def new_pma(V, seg, slots_per_lane, density):
    import lpma_amd as L
    if density is None:
        pma = L.PMA(V, segment_slots=seg)
    else:  # (lower leaf, lower root, upper leaf, upper root)
        pma = L.PMA(V, segment_slots=seg, density_lower=density[1], density_upper=density[3],
                    density_lower_leaf=density[0], density_upper_leaf=density[2])
    pma.set_launch(slots_per_lane)
    return pma


def warm_up(bulk, ins_d, del_d, steps, V, seg, slots_per_lane=0, density=None):
    """W untimed batches on an instance of their own (the timed pass starts
    from a fresh bulk load, so the whole stream is timed)."""
    import torch
    import lpma_amd as L
    if not steps:
        return
    pma = new_pma(V, seg, slots_per_lane, density)
    pma.bulk_load(L.to_device(bulk))
    for op, i in steps:
        pma.update(ins_d[i] if op else del_d[i], op)
    torch.cuda.synchronize()
    pma.close()


def run_gpu(bulk, ins_d, del_d, steps, V, seg, profile, slots_per_lane=0, graph=False, density=None):
    """Bulk load, then time `steps` (op, batch) in order."""
    import torch
    import lpma_amd as L
    stream = torch.cuda.current_stream()
    pma = new_pma(V, seg, slots_per_lane, density)
    t = time.time()
    pma.bulk_load(L.to_device(bulk))
    torch.cuda.synchronize()
    log(f"bulk load of {bulk.size} edges: {time.time() - t:.3f}s, capacity {pma.view().capacity}")
    # the stream's resizes find their arrays in the device pool (allocated
    # here, before the timed region, for the peak of bulk + every insert)
    pma.reserve_capacity(int(bulk.size) + sum(int(b.numel()) for b in ins_d))
    if graph:
        pma.set_graph(True)
        pma.reserve(max(int(b.numel()) for b in list(ins_d) + list(del_d)))
    pma.reset_stats()
    if profile:
        pma.set_profiling(True)
        pma.profile()  # clear
    n_updates = 0
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in steps]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j, (op, i) in enumerate(steps):
        b = ins_d[i] if op else del_d[i]
        ev[j][0].record(stream)
        pma.update(b, op)
        ev[j][1].record(stream)
        n_updates += b.numel()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([a.elapsed_time(b) for a, b in ev])
    stats = pma.stats()
    prof = pma.profile() if profile else None
    # consistency of the final state: sorted, unique, counts and markers
    assert pma.check() == 0
    res = dict(wall=wall, n_updates=n_updates, lat_ms=lat, stats=stats, prof=prof, edges=pma.edges,
               capacity=int(pma.view().capacity), height=int(pma.view().tree_height))
    pma.close()
    return res


def host_cpu():
    """(cores this process may use, CPU model name) of the host."""
    cores = len(os.sched_getaffinity(0))
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model


def cpu_baseline(bulk, ins_batches, del_batches, V, n_ins, n_del, threads):
    """The CPU oracle on the GPU run's own batches: the first n_ins insert
    batches from the bulk-loaded state, and the first n_del delete batches
    from the state the GPU deletes from (every edge of the stream live)."""
    import oracle
    times = []
    n = 0
    empty = np.empty(0, dtype=np.uint64)
    S = oracle.sort_unique(bulk)
    cap = S.size + sum(int(b.size) for b in ins_batches) + 1
    # two output buffers, touched once before timing, used in turn
    bufs = [np.zeros(cap, dtype=np.uint64), np.zeros(cap, dtype=np.uint64)]
    for i in range(n_ins):
        t = time.perf_counter()
        S = oracle.apply(S, ins_batches[i], empty, threads=threads, out=bufs[i & 1])
        oracle.row_offsets(S, V, threads=threads)
        times.append(time.perf_counter() - t)
        n += ins_batches[i].size
    S = oracle.sort_unique(np.concatenate([bulk] + list(ins_batches)))
    for i in range(n_del):
        t = time.perf_counter()
        S = oracle.apply(S, empty, del_batches[i], threads=threads, out=bufs[(n_ins + i) & 1])
        oracle.row_offsets(S, V, threads=threads)
        times.append(time.perf_counter() - t)
        n += del_batches[i].size
    return n / sum(times), times


# HBM traffic per launch, measured by rocprofv3 counter passes over a named
# configuration (profiles/traffic_*.json); reported only for that configuration
def measured_traffic(kernel, cfg):
    """(bytes per launch, source) of `kernel` from the counter summary whose
    recorded configuration equals `cfg`, else (None, reason)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "traffic_*.json"))):
        try:
            with open(path) as f:
                t = json.load(f)
        except (OSError, ValueError):
            continue
        if t.get("config") != cfg:
            continue
        k = t.get("kernels", {}).get(kernel)
        if k is None:
            return None, f"{kernel} not in {os.path.basename(path)}"
        return k["hbm_bytes_per_launch"], {"file": "profiles/" + os.path.basename(path),
                                           "window": t.get("window"), "method": t.get("method")}
    return None, "no counter summary for this configuration"


def _ptr_arrays(tensors_per_graph, r):
    """ctypes (pointer, count) arrays of round r over all graphs."""
    import ctypes
    n = len(tensors_per_graph)
    p = (ctypes.c_void_p * n)()
    c = (ctypes.c_uint64 * n)()
    for i in range(n):
        t = tensors_per_graph[i][r]
        c[i] = t.numel()
        p[i] = t.data_ptr() if t.numel() else None
    return p, c


def build_fleet(w, seg, density, slots_per_lane):
    import torch
    import lpma_amd as L
    ll, lr, ul, ur = density
    fleet = []
    for i in range(w.n):
        p = L.PMA(w.V[i], segment_slots=seg, density_lower=lr, density_upper=ur, density_lower_leaf=ll,
                  density_upper_leaf=ul)
        p.bulk_load(L.to_device(w.initial[i]))
        p.set_launch(slots_per_lane)
        fleet.append(p)
    torch.cuda.synchronize()
    return fleet


def run_rounds(w, fleet, group, ins_d, del_d, rounds, stream):
    """Runs `rounds` (indices) through the group (batched form) or one
    update call per graph (plain form); returns per-round event times (ms)
    and the wall time."""
    import ctypes
    import torch
    import lpma_amd as L
    lib = L.lib()
    sp = ctypes.c_void_p(stream.cuda_stream)
    args = [(_ptr_arrays(ins_d, r), _ptr_arrays(del_d, r)) for r in rounds]
    hs = [p._h for p in fleet]
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in rounds]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j, r in enumerate(rounds):
        (pi, ni), (pd, nd) = args[j]
        ev[j][0].record(stream)
        if group is not None:
            rc = lib.lpma_group_update(group._h, pi, ni, pd, nd, sp)
            if rc:
                raise L.LpmaError(rc, "group_update")
        else:
            for i, h in enumerate(hs):
                rc = lib.lpma_update_mixed(h, pi[i], ni[i], pd[i], nd[i], sp)
                if rc:
                    raise L.LpmaError(rc, "update_mixed")
        ev[j][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return np.array([a.elapsed_time(b) for a, b in ev]), wall


def batched_measure(args):
    """The batched form against the plain form on the same rounds; returns
    the JSON object of the batched workload."""
    import torch
    import lpma_amd as L
    from lpma_amd.workloads import BatchedWorkload
    L.lib()
    total = args.warmup + args.steps
    t = time.time()
    w = BatchedWorkload(rounds=total)
    log(f"batched workload: {w.n} graphs, {sum(w.edges)} initial edges, {w.total_updates} updates in {total} "
        f"rounds, generated in {time.time() - t:.1f}s")
    dev = torch.device("cuda", 0)
    ins_d = [[L.to_device(b, dev) for b in w.ins[i]] for i in range(w.n)]
    del_d = [[L.to_device(b, dev) for b in w.dels[i]] for i in range(w.n)]
    stream = torch.cuda.current_stream()
    seg, density, spl = args.segment_slots, (0.42, 0.42, 1.0, 1.0), 0
    tuned = None
    if args.tune:
        from lpma_amd.tune import autotune
        t = time.time()
        best, trials = autotune(w.initial, w.V, w.ins, w.dels, rounds=min(4, total), log=log)
        seg, density, spl = best.segment_slots, best.density, best.slots_per_lane
        tuned = {"best": best.as_dict(), "trials": [x.as_dict() for x in trials],
                 "sample": f"first {min(4, total)} rounds, fresh groups per trial", "seconds": time.time() - t}
        log(f"autotuner picked S={seg} density={density} slots/lane={spl or 'default'} in {time.time() - t:.1f}s")
    warm = list(range(args.warmup))
    timed = list(range(args.warmup, total))
    n_timed = int(w.sizes[args.warmup:].sum())

    res = {}
    finals = {}
    forms = ("plain", "batched", "batched_graph") if args.forms == "both" else (args.forms,)
    for form in forms:
        t = time.time()
        fleet = build_fleet(w, seg, density, spl)
        log(f"{form}: fleet bulk-loaded in {time.time() - t:.1f}s")
        group = L.Group(fleet) if form != "plain" else None
        if group is not None:
            group.set_launch(spl)
            group.set_graph(form == "batched_graph" or (args.forms == "batched" and bool(args.graph)))
            # size the buffers (and capture the round graphs) before timing
            group.reserve(int(w.sizes.sum(axis=1).max()))
        run_rounds(w, fleet, group, ins_d, del_d, warm, stream)
        lat, wall = run_rounds(w, fleet, group, ins_d, del_d, timed, stream)
        res[form] = dict(value=n_timed / wall, wall=wall, p50=float(np.percentile(lat, 50)),
                         p99=float(np.percentile(lat, 99)), lat=lat)
        finals[form] = [(int(p.view().live_keys), int(p.view().tree_height)) for p in fleet]
        assert all(p.check() == 0 for p in fleet)
        log(f"{form}: {res[form]['value']:.3e} updates/s, round p50 {res[form]['p50']:.3f} ms "
            f"p99 {res[form]['p99']:.3f} ms")
        if group is not None:
            group.close()
        for p in fleet:
            p.close()
        torch.cuda.synchronize()
    if len(forms) == 1:
        r = res[forms[0]]
        return {"form": forms[0], "graph": args.graph, "value": r["value"], "p50": r["p50"], "p99": r["p99"]}
    assert finals["plain"] == finals["batched"] == finals["batched_graph"], "the forms ended in different states"
    b, p, bg = res["batched"], res["plain"], res["batched_graph"]
    out = {
        "metric": "edge updates/sec, batched form (64 PMAs, one group update per round) vs plain form",
        "value": b["value"],
        "unit": "edge updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": b["wall"] * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded R-MAT generators)",
        "config": {"workload": "batched_64_rmat_graphs_lognormal_batches", "graphs": w.n,
                   "initial_edges": int(sum(w.edges)), "updates_timed": n_timed, "segment_slots": seg,
                   "density_leaf_root": {"lower": [density[0], density[1]], "upper": [density[2], density[3]]},
                   "seg_slots_per_lane": spl or "default", "rounds": "stream-ordered, one sync per round",
                   "plain_form": "same configuration, one update call per graph per round",
                   "parallelism": "single GPU"},
        "plain_value": p["value"],
        "speedup_vs_plain": b["value"] / p["value"],
        "round_p50_ms": b["p50"], "round_p99_ms": b["p99"],
        "plain_round_p50_ms": p["p50"], "plain_round_p99_ms": p["p99"],
        "graph_rounds_value": bg["value"], "graph_round_p50_ms": bg["p50"], "graph_round_p99_ms": bg["p99"],
        "autotune": tuned,
    }
    return out


def run_batched(args):
    print(json.dumps(batched_measure(args)))


def served_requests(rng, rate, duration, V, ins_pool, del_pool, mean_batch):
    """Poisson arrivals at `rate`/s for `duration` s: 60% update batches
    (lognormal size with the given mean, sigma 1, half inserts from the
    fresh-edge pool, half deletes drawn from the initial edges), 35%
    neighbor reads of 1-1024 uniform vertices, 5% BFS from a uniform root."""
    n = max(1, int(rate * duration))
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    kinds = rng.choice(3, n, p=(0.60, 0.35, 0.05))
    sigma = 1.0
    mu = np.log(mean_batch) - sigma * sigma / 2
    reqs = []
    for i in range(n):
        if kinds[i] == 0:
            k = max(2, int(rng.lognormal(mu, sigma)))
            a = ins_pool[rng.integers(0, ins_pool.size, k // 2)]
            d = del_pool[rng.integers(0, del_pool.size, k - k // 2)]
            reqs.append((t[i], "update", a, d))
        elif kinds[i] == 1:
            reqs.append((t[i], "read", rng.integers(0, V, int(rng.integers(1, 1025))).astype(np.uint32)))
        else:
            reqs.append((t[i], "bfs", int(rng.integers(0, V))))
    return reqs


def serve_trace(pma, reqs, workers, search_workgroups):
    """The served form: requests submitted at their arrival times to the
    server (continuous batching, BFS on snapshots beside the writer).
    Returns [(kind, result, payload, arrival_s)] with result latency in us."""
    import lpma_amd as L
    srv = L.Server(pma, bfs_workers=workers, search_workgroups=search_workgroups)
    tickets = []
    t0 = time.perf_counter()
    for (at, kind, *payload) in reqs:
        while time.perf_counter() - t0 < at:
            pass
        if kind == "update":
            tickets.append((kind, srv.submit_update(payload[0], payload[1]), payload, at))
        elif kind == "read":
            tickets.append((kind, srv.submit_read(payload[0]), payload, at))
        else:
            tickets.append((kind, srv.submit_bfs(payload[0]), payload, at))
    results = [(kind, srv.wait(t, release=True), payload, at) for kind, t, payload, at in tickets]
    st = srv.stats()
    srv.close()
    return results, st


def plain_trace(pma, reqs):
    """The plain form: one loop serving each request alone, in arrival order,
    on the live structure (one update call per batch, reads and BFS through
    the handle); a request waits while earlier ones run."""
    import torch
    import lpma_amd as L
    results = []
    t0 = time.perf_counter()
    applied = 0
    for (at, kind, *payload) in reqs:
        while time.perf_counter() - t0 < at:
            pass
        start = time.perf_counter() - t0
        if kind == "update":
            pma.update_mixed(L.to_device(payload[0]), L.to_device(payload[1]))
            applied += 1
        elif kind == "read":
            offs, vals = pma.neighbors(torch.from_numpy(payload[0].astype(np.int32)))
        else:
            pma.bfs(payload[0])
        torch.cuda.synchronize()
        done = time.perf_counter() - t0
        results.append((kind, {"status": 0, "latency_us": (done - at) * 1e6, "queue_us": (start - at) * 1e6,
                               "applied": applied}, payload, at))
    return results


def summarize_trace(results, st=None):
    lat = {k: np.array([r["latency_us"] for kind, r, _, _ in results if kind == k]) for k in ("update", "read", "bfs")}
    upd = [(at + r["latency_us"] / 1e6, p[0].size + p[1].size) for kind, r, p, at in results if kind == "update"]
    first = min(at for _, _, _, at in results)
    last_upd = max(d for d, _ in upd) if upd else first + 1e-9
    row = {"requests": len(results), "failed": sum(1 for _, r, _, _ in results if r["status"] != 0),
           "update_requests": len(upd), "sustained_updates_per_s": sum(k for _, k in upd) / (last_upd - first),
           "last_update_done_s": last_upd - first,
           "last_done_s": max(at + r["latency_us"] / 1e6 for _, r, _, at in results) - first}
    for k, a in lat.items():
        if a.size:
            row[f"{k}_p50_ms"] = float(np.percentile(a, 50)) / 1e3
            row[f"{k}_p99_ms"] = float(np.percentile(a, 99)) / 1e3
    if st is not None:
        row.update({"update_runs": st["update_runs"], "read_runs": st["read_runs"], "snapshots": st["snapshots"],
                    "max_concurrent_bfs": st["max_concurrent_bfs"], "peak_admitted_bytes": st["peak_bytes"],
                    "engine_ms": {"updates": st["update_us"] / 1e3, "reads": st["read_us"] / 1e3,
                                  "admission": st["dispatch_us"] / 1e3, "update_staging": st["stage_us"] / 1e3,
                                  "update_call": st["pipeline_us"] / 1e3, "update_sync": st["finish_us"] / 1e3}})
    return row


def served_measure(args):
    """The request server against a one-request-at-a-time loop on the same
    traces; returns the JSON object of the served workload."""
    import torch
    import lpma_amd as L
    L.lib()
    scale, samples = args.served_scale, args.served_samples
    V = 1 << scale
    t = time.time()
    keys = L.gen_rmat_device(scale, samples, 0.57, 0.19, 0.19, 4)
    pma = L.PMA(V)
    pma.bulk_load(keys)
    torch.cuda.synchronize()
    E = pma.edges
    log(f"served: {samples} R-MAT samples on 2^{scale} vertices -> {E} edges, bulk load {time.time() - t:.1f}s")
    rng = np.random.default_rng(44)
    ins_pool = L.to_host_u64(L.gen_rmat_device(scale, 1 << 24, 0.57, 0.19, 0.19, 4004))
    del_pool = L.to_host_u64(keys[torch.randint(0, keys.numel(), (1 << 24,), device=keys.device)])
    rates = [float(x) for x in args.rates.split(",")]
    sweep = {}
    for mode in args.served_modes.split(","):
        rows = []
        for rate in rates:
            reqs = served_requests(rng, rate, args.duration, V, ins_pool, del_pool, 1e4)
            if mode == "served":
                results, st = serve_trace(pma, reqs, args.bfs_workers, args.search_workgroups)
                row = summarize_trace(results, st)
            else:
                results = plain_trace(pma, reqs)
                row = summarize_trace(results)
            row["offered_req_per_s"] = rate
            rows.append(row)
            log(f"{mode} @ {rate:g} req/s: {row['sustained_updates_per_s']:.3e} updates/s, "
                + ", ".join(f"{k} p50/p99 {row.get(k + '_p50_ms', 0):.2f}/{row.get(k + '_p99_ms', 0):.2f} ms"
                            for k in ("update", "read", "bfs")) + f", all done at {row['last_done_s']:.2f}s"
                + (f", {row['update_runs']} update runs, {row['snapshots']} snapshots, engine ms {row['engine_ms']}"
                   if "update_runs" in row else ""))
        sweep[mode] = rows
    top = sweep["served"][-1]
    base = sweep["plain"][-1] if "plain" in sweep else top
    out = {
        "metric": "served form: sustained edge updates/s with p50/p99 latency per request type (R-MAT 2^24 V, 2^28 samples)",
        "value": top["sustained_updates_per_s"],
        "unit": "edge updates/s",
        "n_gpus": 1,
        "steps": len(rates),
        "warmup": 0,
        "ms_per_step": top["last_done_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (device R-MAT generator, Poisson arrivals)",
        "config": {"workload": "served_rmat24_poisson_mix_60upd_35read_5bfs", "vertices": V, "samples": samples,
                   "edges": int(E), "duration_s_per_rate": args.duration, "rates_req_per_s": rates,
                   "bfs_workers": args.bfs_workers, "search_workgroups": args.search_workgroups,
                   "update_mean_batch": 1e4, "parallelism": "single GPU"},
        "sweep": sweep,
        "speedup_vs_plain": top["sustained_updates_per_s"] / base["sustained_updates_per_s"],
        "update_p99_ms": top.get("update_p99_ms"),
        "plain_update_p99_ms": base.get("update_p99_ms"),
        "plain_form": "one loop serving each request alone in arrival order on the live structure",
        "parity": "reads and searches equal the oracle at their reported versions at this size: "
                  "tests/test_gpu_fullsize.py::test_served_rmat24_trace_matches_oracle",
    }
    pma.close()
    del keys
    L.pool_trim()
    return out


def run_served(args):
    print(json.dumps(served_measure(args)))


def run_analytics(args):
    """Analytics on the live structure: R-MAT 2^24 vertices from 2^29
    samples (seed 5), alternating insert and delete batches of 1e6, each
    followed by PageRank (0.85, uniform start, FP32, 50 iterations or L1
    change < 1e-6) and BFS from vertex 0 over the gapped layout."""
    import torch
    import lpma_amd as L
    L.lib()
    scale, samples = args.analytics_scale, args.analytics_samples
    V = 1 << scale
    t = time.time()
    keys = L.gen_rmat_device(scale, samples, 0.57, 0.19, 0.19, 5)
    pma = L.PMA(V)
    pma.bulk_load(keys)
    torch.cuda.synchronize()
    log(f"analytics: {samples} samples on 2^{scale} vertices -> {pma.edges} edges, capacity "
        f"{pma.view().capacity}, built in {time.time() - t:.1f}s")
    fresh = L.gen_rmat_device(scale, args.analytics_batches * args.analytics_batch, 0.57, 0.19, 0.19, 5005)
    rows = []
    for b in range(args.analytics_batches):
        B = args.analytics_batch
        if b % 2 == 0:
            batch = fresh[b * B:(b + 1) * B]
            op = "insert"
        else:
            # live edges: a uniform sample of the current edge set
            ek, _ = pma.export()
            batch = ek[torch.randint(0, ek.numel(), (B,), device=ek.device)]
            del ek
            op = "delete"
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        pma.update(batch, L.OP_INSERT if op == "insert" else L.OP_DELETE)
        e1.record()
        torch.cuda.synchronize()
        upd_ms = e0.elapsed_time(e1)
        E = pma.edges
        cap = int(pma.view().capacity)
        t0 = time.perf_counter()
        pr, iters, delta = pma.pagerank(0.85, 1e-6, 50)
        torch.cuda.synchronize()
        pr_s = time.perf_counter() - t0
        t0 = time.perf_counter()
        lv, depth, scanned = pma.bfs(0)
        torch.cuda.synchronize()
        bfs_s = time.perf_counter() - t0
        # algorithmic bytes of the binned push: per iteration the live keys
        # and row markers (8 B each), per edge its (dst, contribution) pair
        # written and read back (16 B), per vertex the degree, contribution,
        # rank and partial-sum vectors (about 28 B); once per call the count
        # pass over the live keys
        pr_bytes = iters * (8 * (E + V) + 16 * E + 28 * V) + 8 * (E + V)
        reached = int((lv >= 0).sum().item())
        row = {"batch": b, "op": op, "update_ms": upd_ms, "edges": E, "capacity_slots": cap,
               "pagerank_ms": pr_s * 1e3, "pagerank_iterations": iters, "pagerank_last_l1_change": delta,
               "pagerank_edges_per_s": iters * E / pr_s, "pagerank_alg_GBps": pr_bytes / pr_s / 1e9,
               "bfs_ms": bfs_s * 1e3, "bfs_depth": depth, "bfs_reached": reached, "bfs_edges_scanned": scanned,
               "bfs_edges_per_s": scanned / bfs_s,
               # per scanned edge its key (8 B) and a visited-word probe (4 B)
               "bfs_alg_GBps": 12 * scanned / bfs_s / 1e9}
        rows.append(row)
        log(f"analytics batch {b} ({op}): update {upd_ms:.2f} ms, PageRank {iters} it in {pr_s * 1e3:.1f} ms "
            f"({row['pagerank_edges_per_s']:.3e} edges/s, {row['pagerank_alg_GBps']:.0f} GB/s), BFS depth {depth} "
            f"{bfs_s * 1e3:.1f} ms ({row['bfs_edges_per_s']:.3e} edges/s, {row['bfs_alg_GBps']:.0f} GB/s)")
    cpu = None
    if not args.no_cpu:
        # CPU baseline: the FP64 oracle's PageRank and BFS on the final graph
        # (their parity with the GPU at this size is a GPU test:
        # tests/test_gpu_fullsize.py::test_analytics_rmat24_full_size)
        import oracle
        ek, _ = pma.export()
        host = L.to_host_u64(ek)
        del ek
        t0 = time.perf_counter()
        oracle.bfs(host, V, 0)
        cpu_bfs_s = time.perf_counter() - t0
        it = args.analytics_check_iters
        t0 = time.perf_counter()
        oracle.pagerank(host, V, 0.85, it)
        cpu_pr_s = time.perf_counter() - t0
        E = host.size
        cpu = {"value": E * it / cpu_pr_s, "unit": "PageRank edges/s", "cores": 1, "kind": "port",
               "sample": f"FP64 oracle PageRank, {it} iterations, and BFS from 0 ({cpu_bfs_s:.1f} s, "
                         f"{E / cpu_bfs_s:.3e} edges/s) on the final {E}-edge graph, single thread"}
        log(f"analytics cpu baseline: PageRank {cpu_pr_s:.1f}s, BFS {cpu_bfs_s:.1f}s")
    last = rows[-1]
    out = {
        "metric": "analytics on the live structure: PageRank and BFS edges traversed/s over the gapped layout "
                  "(R-MAT 2^24 V, 2^29 samples, batches of 1e6)",
        "value": last["pagerank_edges_per_s"],
        "unit": "PageRank edges/s",
        "n_gpus": 1,
        "steps": len(rows),
        "warmup": 0,
        "ms_per_step": float(np.mean([r["update_ms"] + r["pagerank_ms"] + r["bfs_ms"] for r in rows])),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device R-MAT generator)",
        "config": {"workload": "analytics_rmat24_2p29_samples_batches_1e6", "vertices": V, "samples": samples,
                   "batch": args.analytics_batch, "pagerank": "0.85, uniform start, 50 it or L1 < 1e-6",
                   "parallelism": "single GPU"},
        "batches": rows,
        "bfs_edges_per_s": last["bfs_edges_per_s"],
        "bfs_alg_GBps": last["bfs_alg_GBps"],
        "pagerank_alg_GBps": last["pagerank_alg_GBps"],
        "parity": "BFS levels bit-exact and PageRank within 1e-6 L1 of the FP64 oracle at this size: "
                  "tests/test_gpu_fullsize.py::test_analytics_rmat24_full_size",
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=("headline", "batched", "served", "analytics"), default="headline")
    ap.add_argument("--analytics-scale", type=int, default=24)
    ap.add_argument("--analytics-samples", type=int, default=1 << 29)
    ap.add_argument("--analytics-batches", type=int, default=4)
    ap.add_argument("--analytics-batch", type=int, default=1000000)
    ap.add_argument("--analytics-check-iters", type=int, default=5)
    ap.add_argument("--served-scale", type=int, default=24)
    ap.add_argument("--served-samples", type=int, default=1 << 28)
    ap.add_argument("--rates", default="500,1000,2000,4000,8000,16000", help="served: offered requests/s to sweep")
    ap.add_argument("--duration", type=float, default=1.5, help="served: seconds of arrivals per rate")
    ap.add_argument("--bfs-workers", type=int, default=2)
    ap.add_argument("--search-workgroups", type=int, default=256,
                    help="served: workgroups each concurrent search expands its frontier with")
    ap.add_argument("--served-modes", default="served,plain")
    ap.add_argument("--graph", type=int, default=0,
                    help="replay rounds as hipGraphs (headline: the timed pass; --forms batched: the rounds)")
    ap.add_argument("--forms", choices=("both", "plain", "batched"), default="both")
    ap.add_argument("--density", default="",
                    help="headline: density bounds 'lower_leaf,lower_root,upper_leaf,upper_root' "
                         "(default: the reference's 0.42 / 1.0 at every level)")
    ap.add_argument("--tune", type=int, default=1, help="batched workload: autotune segment size / density / shape")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--batch", type=int, default=100000)
    ap.add_argument("--scale", type=int, default=20)
    ap.add_argument("--edges", type=int, default=1 << 24)
    ap.add_argument("--segment-slots", type=int, default=32)
    ap.add_argument("--seg-slots-per-lane", type=int, default=0, help="single-segment rebalance shape (0: default)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-sweep", type=int, default=1, help="also run batch sizes 1e3..1e6 of the R-MAT 2^20 stream")
    ap.add_argument("--cpu-batches", type=int, default=15)
    ap.add_argument("--legs", default="batched,served",
                    help="headline: compact legs of the other forms to add to the JSON line ('' for none)")
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one MI355X: sharding the store by key range across GPUs is designed but not "
              "built, so --gpus must be 1.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "served":
        return run_served(args)
    if args.workload == "analytics":
        return run_analytics(args)
    if args.workload == "batched":
        args.steps = 95 if args.steps is None else args.steps
        args.warmup = 5 if args.warmup is None else args.warmup
        return run_batched(args)
    args.warmup = 4 if args.warmup is None else args.warmup
    print(json.dumps(headline(args)))


def headline(args):
    import torch
    import lpma_amd as L
    L.lib()
    V = 1 << args.scale
    B = args.batch
    bulk, ins_b, del_b = make_workload(args.scale, args.edges, B)
    dev = torch.device("cuda", 0)
    ins_d = [L.to_device(b, dev) for b in ins_b]
    del_d = [L.to_device(b, dev) for b in del_b]
    seq = stream_steps(ins_b, del_b)
    dens = tuple(float(x) for x in args.density.split(",")) if args.density else None
    warm_up(bulk, ins_d, del_d, seq[:args.warmup], V, args.segment_slots, args.seg_slots_per_lane, dens)

    # pass 1: the whole stream, throughput and per-batch latency, no profiling events
    r = run_gpu(bulk, ins_d, del_d, seq, V, args.segment_slots, profile=False,
                slots_per_lane=args.seg_slots_per_lane, graph=bool(args.graph), density=dens)
    # pass 2: the same stream with per-launch HIP events for the kernel roofline
    rp = run_gpu(bulk, ins_d, del_d, seq, V, args.segment_slots, profile=True,
                 slots_per_lane=args.seg_slots_per_lane, density=dens)

    if args.copy_calib:
        src = torch.empty(args.copy_calib << 20, dtype=torch.uint8, device=dev)
        dstc = torch.empty_like(src)
        torch.cuda.synchronize()
        dstc.copy_(src)
        torch.cuda.synchronize()
        del src, dstc

    value = r["n_updates"] / r["wall"]
    lat = r["lat_ms"]
    st, pf = rp["stats"], rp["prof"]
    # dominant rebalance kernel: the one that moves the most algorithmic bytes
    # (a lone PMA's one-wave windows run inside the single-segment launch,
    # which then has no wave-class launch of its own)
    wave_in_seg = pf.launches_wave == 0
    kern = {
        "k_rebalance_seg": (pf.rebalance_seg_ms, st.alg_bytes_seg + (st.alg_bytes_wave if wave_in_seg else 0),
                            pf.launches_seg),
        "k_rebalance_wave": (pf.rebalance_wave_ms, 0 if wave_in_seg else st.alg_bytes_wave, pf.launches_wave),
        "k_rebalance_block": (pf.rebalance_block_ms, st.alg_bytes_block, pf.launches_block),
    }
    name = max(kern, key=lambda k: kern[k][1])
    ms, byts, launches = kern[name]
    avg_s = ms / 1e3 / max(launches, 1)
    bytes_per_launch = byts / max(launches, 1)
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    reb_ms = pf.rebalance_seg_ms + pf.rebalance_wave_ms + pf.rebalance_block_ms + pf.large_ms
    reb_bytes = st.alg_bytes_seg + st.alg_bytes_wave + st.alg_bytes_block + st.alg_bytes_large
    phases = {k: round(getattr(pf, k), 3) for k in ("sort_ms", "locate_ms", "elect_ms", "rebalance_seg_ms",
                                                    "rebalance_wave_ms", "rebalance_block_ms", "large_ms",
                                                    "rebuild_ms", "pyramid_ms")}
    windows = {str(l): int(st.windows_per_level[l]) for l in range(32) if st.windows_per_level[l]}
    prof_cfg = {"scale": args.scale, "edges": args.edges, "batch": B, "segment_slots": args.segment_slots,
                "density": args.density or "reference", "seg_slots_per_lane": args.seg_slots_per_lane}
    traffic, traffic_src = measured_traffic(name, prof_cfg)

    # the same R-MAT stream at every batch size: batches of 1e3 (a fixed
    # window of the first 1000), 1e4, 1e5 and 1e6 (the whole stream), each
    # from a fresh bulk load, with the rebalance kernels' algorithmic GB/s
    sweep = []
    if args.batch_sweep:
        for Bs in (1000, 10000, 100000, 1000000):
            if Bs == B and not args.graph:
                rs, rsp = r, rp
                n_steps = len(seq)
            else:
                bulk_s, ins_s, del_s = make_workload(args.scale, args.edges, Bs)
                ins_ds = [L.to_device(b, dev) for b in ins_s]
                del_ds = [L.to_device(b, dev) for b in del_s]
                seq_s = stream_steps(ins_s, del_s)
                if Bs == 1000:
                    seq_s = seq_s[:1000]
                n_steps = len(seq_s)
                # timing pass (no per-phase events: at small batches their
                # records cost as much as the kernels), then a profiled pass of
                # the same batches for the rebalance kernels' time and bytes
                rs = run_gpu(bulk_s, ins_ds, del_ds, seq_s, V, args.segment_slots, profile=False,
                             slots_per_lane=args.seg_slots_per_lane)
                rsp = run_gpu(bulk_s, ins_ds, del_ds, seq_s, V, args.segment_slots, profile=True,
                              slots_per_lane=args.seg_slots_per_lane)
                del ins_ds, del_ds
            stt, pff = rsp["stats"], rsp["prof"]
            rms = pff.rebalance_seg_ms + pff.rebalance_wave_ms + pff.rebalance_block_ms + pff.large_ms
            rbytes = stt.alg_bytes_seg + stt.alg_bytes_wave + stt.alg_bytes_block + stt.alg_bytes_large
            sweep.append({"batch": Bs, "batches": n_steps, "updates_per_s": rs["n_updates"] / rs["wall"],
                          "p50_batch_ms": float(np.percentile(rs["lat_ms"], 50)),
                          "p99_batch_ms": float(np.percentile(rs["lat_ms"], 99)),
                          "rebalance_alg_GBps": rbytes / (rms / 1e3) / 1e9 if rms > 0 else 0.0,
                          "rebuilds": int(stt.rebuilds)})
            log(f"batch {Bs}: {sweep[-1]['updates_per_s']:.3e} updates/s, p50 {sweep[-1]['p50_batch_ms']:.3f} ms, "
                f"p99 {sweep[-1]['p99_batch_ms']:.3f} ms, rebalance {sweep[-1]['rebalance_alg_GBps']:.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        cores, model = host_cpu()
        nb = min(args.cpu_batches, len(ins_b), len(del_b))
        runs = {}
        for threads in sorted({cores, 1}, reverse=True):
            t = time.time()
            runs[threads] = cpu_baseline(bulk, ins_b, del_b, V, nb, nb, threads)
            log(f"cpu baseline ({threads} threads): {runs[threads][0]:.3e} updates/s in {time.time() - t:.1f}s")
        best = max(runs, key=lambda k: runs[k][0])
        cpu = {"value": runs[best][0], "unit": "edge updates/s", "cores": best, "kind": "port",
               "host_cores": cores, "cpu_model": model,
               "sample": f"CPU oracle (sorted merge + CSR offsets) on the GPU run's own batches: its first {nb} "
                         f"insert batches of {B} from the bulk-loaded state and its first {nb} delete batches from "
                         f"the state it deletes from, on all {cores} host cores and on one; the faster is reported",
               "p50_batch_ms": float(np.percentile(runs[best][1], 50)) * 1e3,
               "p99_batch_ms": float(np.percentile(runs[best][1], 99)) * 1e3,
               "by_threads": {str(k): {"value": v, "p50_batch_ms": float(np.percentile(ts, 50)) * 1e3,
                                       "p99_batch_ms": float(np.percentile(ts, 99)) * 1e3}
                              for k, (v, ts) in runs.items()}}
    del ins_d, del_d
    L.pool_trim()

    out = {
        "metric": "edge updates/sec (R-MAT 2^20 V, 16M E, batch 1e5); p99 batch latency; HBM GB/s",
        "value": value,
        "unit": "edge updates/s",
        "n_gpus": 1,
        "steps": len(seq),
        "warmup": args.warmup,
        "ms_per_step": r["wall"] * 1e3 / len(seq),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded R-MAT generator)",
        "config": {"workload": f"rmat{args.scale}_{args.edges}E_half_bulk_stream_insert_then_delete_batch{B}",
                   "vertices": V, "edges": args.edges, "batch": B, "segment_slots": args.segment_slots,
                   "timed": f"the whole stream: {len(ins_b)} insert + {len(del_b)} delete batches from a fresh bulk "
                            f"load (warm-up batches ran on another instance)",
                   "density": ("lower 0.42 / upper 1.0 (reference)" if not args.density else
                               "lower leaf,root / upper leaf,root = " + args.density),
                   "memory": "growth arrays for the stream's peak reserved in the device pool before timing "
                             "(lpma_reserve_capacity); resizes still rebuild in the timed batches",
                   "parallelism": "single GPU"},
        "p50_batch_ms": float(np.percentile(lat, 50)),
        "p99_batch_ms": float(np.percentile(lat, 99)),
        "mean_batch_ms": float(lat.mean()),
        "event_span_share": float(lat.sum() / 1e3 / r["wall"]),
        "rebalance_alg_GBps": reb_bytes / (reb_ms / 1e3) / 1e9 if reb_ms > 0 else 0.0,
        "phase_ms_total": phases,
        "roofline": {"bound": "hbm", "kernel": name, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "alg_bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_s * 1e6,
                     "launches": int(launches)},
        "cpu_baseline": cpu,
        "by_batch_size": sweep,
        "windows_per_level": windows,
        "rebuilds": int(st.rebuilds),
        "final_state": {"edges": r["edges"], "capacity": r["capacity"], "tree_height": r["height"]},
    }
    for leg in [x for x in args.legs.split(",") if x]:
        t = time.time()
        try:
            out[leg] = LEGS[leg](args)
        except Exception as e:  # a leg that fails leaves the headline intact
            out[leg] = {"error": repr(e)}
        log(f"leg {leg}: {time.time() - t:.1f}s")
    return out


def batched_leg(args):
    """The batched form against the plain form: 64 PMAs, 5 warm-up and 95
    timed rounds, the autotuned configuration."""
    import copy
    a = copy.copy(args)
    a.steps, a.warmup, a.forms, a.tune, a.graph = 95, 5, "both", 1, 0
    o = batched_measure(a)
    keep = ("value", "plain_value", "speedup_vs_plain", "round_p50_ms", "round_p99_ms", "plain_round_p50_ms",
            "plain_round_p99_ms", "graph_rounds_value", "graph_round_p50_ms", "graph_round_p99_ms")
    leg = {k: o[k] for k in keep}
    leg["unit"] = "edge updates/s"
    leg["config"] = {k: o["config"][k] for k in ("workload", "graphs", "initial_edges", "updates_timed",
                                                  "segment_slots", "density_leaf_root")}
    leg["p99_not_above_plain"] = o["round_p99_ms"] <= o["plain_round_p99_ms"]
    return leg


def served_leg(args):
    """The request server against the one-request-at-a-time loop at the top
    offered rate of the sweep (16000 requests/s for 1 s)."""
    import copy
    a = copy.copy(args)
    a.rates, a.duration, a.served_modes = "16000", 1.0, "served,plain"
    o = served_measure(a)
    top, base = o["sweep"]["served"][-1], o["sweep"]["plain"][-1]
    return {"value": o["value"], "unit": "edge updates/s", "plain_value": base["sustained_updates_per_s"],
            "speedup_vs_plain": o["speedup_vs_plain"], "update_p50_ms": top.get("update_p50_ms"),
            "update_p99_ms": top.get("update_p99_ms"), "plain_update_p50_ms": base.get("update_p50_ms"),
            "plain_update_p99_ms": base.get("update_p99_ms"), "read_p99_ms": top.get("read_p99_ms"),
            "bfs_p99_ms": top.get("bfs_p99_ms"),
            "p99_not_above_plain": (top.get("update_p99_ms") or 0) <= (base.get("update_p99_ms") or 0),
            "config": {"workload": o["config"]["workload"], "edges": o["config"]["edges"],
                       "offered_req_per_s": 16000, "duration_s": 1.0}}


LEGS = {"batched": batched_leg, "served": served_leg}


if __name__ == "__main__":
    main()
