"""Autotuner of the batched form: segment size, density bounds and the
workgroup shape of the single-segment rebalance for a group of PMAs.

The reference fixes 32-slot segments and flat 0.42 / 1.0 density bounds
(rpma.cuh:76-79, :1769); a group of many PMAs with uneven batches is best
served by other choices, and the best one depends on the graphs and the
batch mix.  The tuner measures instead of modelling: each candidate builds
the group from the same initial graphs, runs the first rounds of the
workload, and is scored by edge updates per second over those rounds.  The
search is coordinate descent (segment size, then density, then shape), so
it costs a handful of short trials.  Any choice gives the same edge sets;
only the layout and the speed differ.
"""
import ctypes
import time

import numpy as np

from . import Group, PMA, lib, to_device, LpmaError

SEGMENT_SLOTS = (32, 64, 128, 256, 512, 1024)
# (lower_leaf, lower_root, upper_leaf, upper_root)
DENSITIES = (
    (0.42, 0.42, 1.0, 1.0),    # the reference's flat bounds
    (0.25, 0.25, 1.0, 1.0),    # shrink later
    (0.42, 0.42, 0.9, 0.9),    # leave slack in every window
    (0.08, 0.30, 1.0, 0.75),   # interpolated leaf-to-root bounds (classic PMA shape)
)
SLOTS_PER_LANE = (0, 1, 2, 8)  # 0 = the library default (4)


class Trial:
    def __init__(self, segment_slots, density, slots_per_lane, value, seconds):
        self.segment_slots = segment_slots
        self.density = density
        self.slots_per_lane = slots_per_lane
        self.value = value
        self.seconds = seconds

    def as_dict(self):
        return {"segment_slots": self.segment_slots, "density": list(self.density),
                "seg_slots_per_lane": self.slots_per_lane, "updates_per_s": self.value}


def build_group(initial_dev, Vs, segment_slots, density, slots_per_lane=0, device=None):
    """PMAs bulk-loaded from device key tensors, and their group."""
    import torch
    ll, lr, ul, ur = density
    members = []
    for V, keys in zip(Vs, initial_dev):
        p = PMA(V, segment_slots=segment_slots, density_lower=lr, density_upper=ur, density_lower_leaf=ll,
                density_upper_leaf=ul, device=device)
        p.bulk_load(keys)
        members.append(p)
    g = Group(members)
    g.set_launch(slots_per_lane)
    torch.cuda.synchronize()
    return members, g


def round_args(ins_dev, del_dev, r):
    n = len(ins_dev)
    pi = (ctypes.c_void_p * n)()
    ni = (ctypes.c_uint64 * n)()
    pd = (ctypes.c_void_p * n)()
    nd = (ctypes.c_uint64 * n)()
    for i in range(n):
        a, d = ins_dev[i][r], del_dev[i][r]
        ni[i], nd[i] = a.numel(), d.numel()
        pi[i] = a.data_ptr() if a.numel() else None
        pd[i] = d.data_ptr() if d.numel() else None
    return pi, ni, pd, nd


def run_trial(initial_dev, Vs, ins_dev, del_dev, rounds, segment_slots, density, slots_per_lane):
    """Edge updates/s of the group over `rounds` (after one untimed round)."""
    import torch
    t = time.time()
    members, g = build_group(initial_dev, Vs, segment_slots, density, slots_per_lane)
    try:
        L = lib()
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        args = [round_args(ins_dev, del_dev, r) for r in rounds]
        rc = L.lpma_group_update(g._h, *args[0], stream)
        if rc:
            raise LpmaError(rc, "trial round")
        torch.cuda.synchronize()
        n = 0
        t0 = time.perf_counter()
        for a in args[1:]:
            rc = L.lpma_group_update(g._h, *a, stream)
            if rc:
                raise LpmaError(rc, "trial round")
            n += sum(a[1]) + sum(a[3])
        torch.cuda.synchronize()
        value = n / (time.perf_counter() - t0)
    finally:
        g.close()
        for p in members:
            p.close()
    return Trial(segment_slots, density, slots_per_lane, value, time.time() - t)


def autotune(initial, Vs, ins, dels, rounds=4, segment_slots=SEGMENT_SLOTS, densities=DENSITIES,
             slots_per_lane=SLOTS_PER_LANE, log=None):
    """Best (segment_slots, density, slots_per_lane) for a group.

    initial[i]: the initial keys of graph i (numpy uint64 or device tensor);
    ins[i][r] / dels[i][r]: its batches of round r (the first `rounds` are
    used; they need not be the rounds the tuned group will run).
    Returns (best Trial, list of every Trial).
    """
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    initial_dev = [k if hasattr(k, "data_ptr") else to_device(np.asarray(k, dtype=np.uint64), dev) for k in initial]
    ins_dev = [[to_device(np.asarray(b, dtype=np.uint64), dev) for b in ins[i][:rounds]] for i in range(len(Vs))]
    del_dev = [[to_device(np.asarray(b, dtype=np.uint64), dev) for b in dels[i][:rounds]] for i in range(len(Vs))]
    rr = list(range(rounds))
    trials = []

    def trial(s, d, k):
        for t in trials:
            if (t.segment_slots, t.density, t.slots_per_lane) == (s, d, k):
                return t
        t = run_trial(initial_dev, Vs, ins_dev, del_dev, rr, s, d, k)
        trials.append(t)
        if log:
            log(f"tune: S={s} density={d} slots/lane={k or 'default'}: {t.value:.3e} updates/s ({t.seconds:.1f}s)")
        return t

    best = max((trial(s, densities[0], 0) for s in segment_slots), key=lambda t: t.value)
    best = max((trial(best.segment_slots, d, 0) for d in densities), key=lambda t: t.value)
    best = max((trial(best.segment_slots, best.density, k) for k in slots_per_lane), key=lambda t: t.value)
    return best, trials
