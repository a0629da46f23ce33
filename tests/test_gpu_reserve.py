"""Growth without device allocation (``lpma_reserve_capacity``).

A PMA bulk-loaded small and grown through several heights by insert
batches: after reserve_capacity for its peak and reserve for its batch size,
the updates (resizes included) take every array from what the handle holds,
so the bytes the device pool holds from hipMalloc stay the same, and the
result equals the oracle's set; deleting back down to the bulk-loaded set
(shrinks, root rebuilds) allocates nothing either (rpma.cuh:1396-1415 is the
reference's resize, which allocates its levels when it happens).
"""
import numpy as np
import pytest
import torch

import lpma_amd as L
import oracle
from gpu_util import say

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", [1, 2])
def test_reserved_growth_allocates_nothing(seed):
    V = 1 << 12
    keys = L.shuffle(L.gen_uniform(V, 90000, seed), seed)
    bulk, ins = keys[:10000], keys[10000:]
    say(f"reserve_capacity: growth from {bulk.size} to {keys.size} edges, seed {seed}")
    p = L.PMA(V)
    p.bulk_load(bulk)
    h0 = int(p.view().tree_height)
    batch = 8000
    p.reserve_capacity(bulk.size + ins.size)
    p.reserve(batch)
    torch.cuda.synchronize()
    held = L.pool_reserved_bytes()
    ref = oracle.apply(np.empty(0, dtype=np.uint64), bulk)
    for i in range(0, ins.size, batch):
        p.update(ins[i:i + batch], 1)
        ref = oracle.apply(ref, ins[i:i + batch])
    torch.cuda.synchronize()
    assert int(p.view().tree_height) >= h0 + 2, "the stream must resize through two heights"
    assert L.pool_reserved_bytes() == held, "a reserved resize allocated device memory"
    check(p, ref, V)
    torch.cuda.synchronize()
    held = L.pool_reserved_bytes()  # the export and check above take scratch of their own
    # back down to the bulk-loaded set: shrinks and root rebuilds at heights
    # held before allocate nothing either
    rng = np.random.default_rng(seed)
    dels = rng.permutation(ins)
    trace = []
    for i in range(0, dels.size, batch):
        p.update(dels[i:i + batch], 0)
        ref = oracle.apply(ref, (), dels[i:i + batch])
        torch.cuda.synchronize()
        trace.append((int(p.view().tree_height), L.pool_reserved_bytes() - held))
    assert int(p.view().tree_height) < h0 + 2, "the deletes must shrink the PMA"
    assert L.pool_reserved_bytes() == held, f"a reserved shrink allocated device memory: h0 {h0}, {trace}"
    check(p, ref, V)
    p.close()


def check(p, ref, V):
    ek, ro = p.export()
    assert np.array_equal(L.to_host_u64(ek), ref)
    assert np.array_equal(ro.cpu().numpy().view(np.uint32), oracle.row_offsets(ref, V))
    assert p.check() == 0
