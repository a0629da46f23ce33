"""The allocation-free export (lpma_export_compacted_noalloc), through the C
ABI: its output equals lpma_export_compacted's on two calls in a row, the
second on a smaller structure (the static scratch keeps the first call's
contents), it makes no allocation (an armed allocation failure stays
unfired), and a structure beyond its scratch is refused with
LPMA_ERR_TOO_LARGE."""
import ctypes

import numpy as np
import pytest
import torch

import lpma_amd as L
import oracle
from gpu_util import keys_of, say

pytestmark = pytest.mark.gpu

TOO_LARGE = 5


def exports(pma):
    """(keys, offsets) from the allocating export and from the allocation-free one."""
    lib = L.lib()
    v = pma.view()
    n = max(int(v.edges), 1)
    out = []
    for fn in (lib.lpma_export_compacted, lib.lpma_export_compacted_noalloc):
        keys = torch.full((n,), -7, dtype=torch.int64, device=pma.device)
        ro = torch.full((pma.num_vertices + 1,), -7, dtype=torch.int32, device=pma.device)
        if fn is lib.lpma_export_compacted_noalloc:
            L.inject_alloc_failure(1)  # any allocation from here on would fail
        rc = fn(pma._h, ctypes.c_void_p(keys.data_ptr()), ctypes.c_void_p(ro.data_ptr()),
                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        if fn is lib.lpma_export_compacted_noalloc:
            assert L.inject_alloc_failure(0) == 1, "the allocation-free export allocated"
        assert rc == 0, rc
        out.append((L.to_host_u64(keys[: int(v.edges)]), ro.cpu().numpy().view(np.uint32).copy()))
    return out


def test_noalloc_export_equals_export_twice_then_smaller():
    say("allocation-free export")
    rng = np.random.default_rng(71)
    # first call: 2^20 vertices, a few million edges (2^19 segments)
    V1 = 1 << 20
    k1 = keys_of(rng.integers(0, V1, 3_000_000), rng.integers(0, V1, 3_000_000))
    # second call: a smaller structure, whose scan covers only a prefix of
    # what the first call left in the scratch
    V2 = 5000
    k2 = keys_of(rng.integers(0, V2, 40_000), rng.integers(0, V2, 40_000))
    for V, keys in ((V1, k1), (V2, k2)):
        with L.PMA(V) as pma:
            pma.bulk_load(keys)
            ref = oracle.sort_unique(keys)
            (ka, ra), (kb, rb) = exports(pma)
            assert np.array_equal(kb, ka) and np.array_equal(rb, ra)
            assert np.array_equal(kb, ref)
            assert np.array_equal(rb, oracle.row_offsets(ref, V))
            say(f"  V={V}: {ref.size} edges, {int(pma.view().tree_height)} levels, equal")


def test_noalloc_export_refuses_beyond_its_scratch():
    # 2-slot segments over 2^23 row markers: 2^22 segments, beyond 2^21
    with L.PMA(1 << 23, segment_slots=2) as pma:
        assert (1 << int(pma.view().tree_height)) > (1 << 21)
        lib = L.lib()
        rc = lib.lpma_export_compacted_noalloc(pma._h, None, None, None)
        assert rc == TOO_LARGE
