"""GPU parity across every path that meets a count pyramid left owing.

A lone PMA's small batches leave their pyramid's chunk levels to the next
batch's sort kernels.  Every other path must sum them first: a batch too
large for that sort, a graph replay, a group run.  After each batch the
slot layout is compared with the oracle's layout model (window choices
depend on every pyramid count) and the edges with the oracle's sorted
merge.
"""
import numpy as np
import pytest

import lpma_amd as L
import oracle
from gpu_util import Triple, keys_of, say

pytestmark = pytest.mark.gpu


def rand_keys(rng, V, n):
    return keys_of(rng.integers(0, V, n), rng.integers(0, V, n))


def test_owed_pyramid_paid_on_every_path():
    say("owed pyramid levels across paths")
    rng = np.random.default_rng(41)
    V = 20000
    t = Triple(V)
    other = L.PMA(V)
    other_ref = np.empty(0, dtype=np.uint64)

    def small():
        ins = rand_keys(rng, V, int(rng.integers(100, 20000)))
        dels = t.ref[rng.integers(0, t.ref.size, t.ref.size // 20)] if t.ref.size else np.empty(0, np.uint64)
        t.mixed(ins, dels)
        t.compare()

    for _ in range(3):
        small()
    # more keys than the sort that carries the owed levels takes
    t.insert(rand_keys(rng, V, 150000))
    t.compare()
    small()
    # graph replays
    t.pma.set_graph(True)
    small()
    small()
    t.pma.set_graph(False)
    small()
    # a group run with the handle as a member
    g = L.Group([t.pma, other])
    ins = [rand_keys(rng, V, 30000), rand_keys(rng, V, 5000)]
    dels = [t.ref[rng.integers(0, t.ref.size, 3000)], np.empty(0, dtype=np.uint64)]
    g.update(ins, dels)
    t.model.update_mixed(ins[0], dels[0])
    t.ref = oracle.apply(t.ref, ins[0], dels[0])
    other_ref = oracle.apply(other_ref, ins[1], dels[1])
    t.compare()
    assert np.array_equal(L.to_host_u64(other.export()[0]), other_ref)
    small()
    small()
    # deletes down to a shrink, then growth again
    t.delete(t.ref[rng.permutation(t.ref.size)[: t.ref.size * 3 // 4]])
    t.compare()
    small()
    g.close()
    t.close()
    other.close()
