"""Failure paths of an update: running out of device memory.

The reference prints a failed allocation and carries on (cErr,
rpma.cuh:13-19) and silently skips a window too large for its scratch
(rpma.cuh:1248-1251).  Here an update that cannot get its memory returns
LPMA_ERR_OUT_OF_MEMORY and leaves every PMA it touched unchanged, slot for
slot.  Two ways to run out are tested: a pool budget below what a resize
needs, and an injected failure of each allocation of an update in turn
(``lpma_inject_alloc_failure``): a lone PMA that resizes, a lone PMA whose
batch takes the multi-workgroup window path (its single-segment windows are
queued before the status is read), and a group in which one member resizes
while the others rebalance windows, with eager and graph rounds.  After
every failure the same update is applied again and must equal the oracle.
"""
import numpy as np
import pytest
import torch

import lpma_amd as L
import oracle
from gpu_util import keys_of, say

pytestmark = pytest.mark.gpu

OOM = 3


def rand_keys(rng, V, n):
    return keys_of(rng.integers(0, V, n), rng.integers(0, V, n))


def snap(pma):
    """Everything an update may change: version, geometry, live count, the
    gapped slot array and the marker row offsets."""
    torch.cuda.synchronize()
    v = pma.view()
    gk, gro = pma.copy_view()
    return ((int(v.version), int(v.capacity), int(v.live_keys), int(v.tree_height)), L.to_host_u64(gk),
            gro.cpu().numpy().copy())


def same(a, b):
    return a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


def check_oracle(pma, ref, V):
    torch.cuda.synchronize()
    ek, ro = pma.export()
    assert np.array_equal(L.to_host_u64(ek), ref)
    assert np.array_equal(ro.cpu().numpy().view(np.uint32), oracle.row_offsets(ref, V))
    assert pma.check() == 0


class Case:
    """PMAs in a known state, the update under test, and the oracle's sets
    before and after it."""

    def __init__(self, specs, group=False, graph=False):
        """specs: (V, bulk keys, inserts, deletes) per PMA."""
        self.specs = specs
        self.pmas = []
        self.before = []
        self.after = []
        for V, bulk, ins, dels in specs:
            p = L.PMA(V)
            p.bulk_load(bulk)
            self.pmas.append(p)
            b = oracle.apply(np.empty(0, dtype=np.uint64), bulk)
            self.before.append(b)
            self.after.append(oracle.apply(b, ins, dels))
        self.group = None
        if group:
            self.group = L.Group(self.pmas)
            self.group.set_graph(graph)
        torch.cuda.synchronize()

    def apply(self):
        if self.group is not None:
            self.group.update([s[2] for s in self.specs], [s[3] for s in self.specs])
        else:
            for p, s in zip(self.pmas, self.specs):
                p.update_mixed(s[2], s[3])

    def snaps(self):
        return [snap(p) for p in self.pmas]

    def verify(self):
        for p, s, ref in zip(self.pmas, self.specs, self.after):
            check_oracle(p, ref, s[0])

    def close(self):
        if self.group is not None:
            self.group.close()
        for p in self.pmas:
            p.close()


def sweep_failures(make, limit=600):
    """For k = 1, 2, ...: a fresh case, the k-th allocation of its update made
    to fail.  A failed update must leave every PMA unchanged and succeed when
    applied again; a successful one must equal the oracle.  Ends at the first
    k beyond the update's allocations.  Returns (failures, points)."""
    failures = 0
    for k in range(1, limit):
        case = make()
        before = case.snaps()
        L.inject_alloc_failure(k)
        code = 0
        try:
            case.apply()
        except L.LpmaError as e:
            code = e.code
        remaining = L.inject_alloc_failure(0)
        if code:
            assert code == OOM, (k, code)
            assert remaining == 0, k
            after = case.snaps()
            for i, (a, b) in enumerate(zip(before, after)):
                assert same(a, b), f"allocation {k} failed but PMA {i} changed"
            failures += 1
            case.apply()
        case.verify()
        case.close()
        if remaining:
            return failures, k - 1
    raise AssertionError("the injection still fired after %d allocations" % limit)


def growth_spec(seed):
    rng = np.random.default_rng(seed)
    V = 1 << 12
    keys = L.shuffle(L.gen_uniform(V, 60000, seed), seed)
    return (V, keys[:20000], keys[20000:], np.empty(0, dtype=np.uint64)), rng


def hub_spec(seed):
    rng = np.random.default_rng(seed)
    V = 1 << 14
    bulk = rand_keys(rng, V, 150000)
    hub = keys_of(np.full(12000, 9000), rng.choice(V, 12000, replace=False))
    ins = np.concatenate([hub, rand_keys(rng, V, 3000)])
    return (V, bulk, ins, bulk[:500].copy())


def small_spec(seed):
    rng = np.random.default_rng(seed)
    V = 1 << 12
    bulk = rand_keys(rng, V, 30000)
    return (V, bulk, rand_keys(rng, V, 2000), bulk[100:600].copy())


def test_injected_failure_lone_resize():
    say("failure injection: lone PMA that resizes")
    spec, _ = growth_spec(1)
    fails, points = sweep_failures(lambda: Case([spec]))
    assert fails >= 1 and points >= 1, (fails, points)


def test_injected_failure_lone_large_windows():
    say("failure injection: lone PMA with multi-workgroup windows")
    spec = hub_spec(2)
    case = Case([spec])
    case.pmas[0].reset_stats()
    case.apply()
    assert case.pmas[0].stats().alg_bytes_large > 0, "the batch must take the multi-workgroup path"
    case.verify()
    case.close()
    sweep_failures(lambda: Case([spec]))


@pytest.mark.parametrize("graph", [False, True])
def test_injected_failure_group(graph):
    say(f"failure injection: group, graph={graph}")
    g0, _ = growth_spec(3)
    specs = [g0, hub_spec(4), small_spec(5)]
    fails, points = sweep_failures(lambda: Case(specs, group=True, graph=graph))
    assert fails >= 1, (fails, points)


def test_pool_budget_exhausted_by_resize():
    say("pool budget below a resize")
    spec, _ = growth_spec(6)
    case = Case([spec])
    before = case.snaps()
    L.pool_trim()
    try:
        L.set_pool_budget(L.pool_reserved_bytes() + 4096)
        with pytest.raises(L.LpmaError) as e:
            case.apply()
        assert e.value.code == OOM
    finally:
        L.set_pool_budget(0)
    assert all(same(a, b) for a, b in zip(before, case.snaps()))
    case.apply()
    case.verify()
    case.close()
