"""GPU parity of the batched form: a group of independent PMAs updated by
one pipeline run per round.

Every member must end exactly as if it had been updated alone: after each
round its gapped slot array, tree height and marker row offsets equal those
of a twin PMA that received the same batch through ``update_mixed``, and its
exported edges and CSR offsets equal the oracle's sorted merge.
"""
import numpy as np
import pytest

import lpma_amd as L
import oracle
from gpu_util import keys_of, say

pytestmark = pytest.mark.gpu


def rand_keys(rng, V, n):
    return keys_of(rng.integers(0, V, n), rng.integers(0, V, n))


class Fleet:
    """Group members, their solo twins and the oracle's key sets."""

    def __init__(self, Vs, segs=None, lower=0.42, upper=1.0, graph=True):
        """graph: replay every round as a captured graph (the tests cover
        both modes; eager rounds are the library default)."""
        segs = segs or [32] * len(Vs)
        self.Vs = list(Vs)
        self.members = [L.PMA(V, segment_slots=s, density_lower=lower, density_upper=upper) for V, s in zip(Vs, segs)]
        self.twins = [L.PMA(V, segment_slots=s, density_lower=lower, density_upper=upper) for V, s in zip(Vs, segs)]
        self.refs = [np.empty(0, dtype=np.uint64) for _ in Vs]
        self.group = L.Group(self.members)
        self.group.set_graph(graph)

    def round(self, ins, dels):
        self.group.update(ins, dels)
        for i in range(len(self.Vs)):
            a = ins[i] if ins[i] is not None else np.empty(0, dtype=np.uint64)
            d = dels[i] if dels[i] is not None else np.empty(0, dtype=np.uint64)
            # a round is one update call for every member, empty or not
            self.twins[i].update_mixed(a, d)
            self.refs[i] = oracle.apply(self.refs[i], a, d)

    def compare(self, light=()):
        """Members in `light` (too big to copy to the host) are checked
        through their edge count and edge queries of the oracle's keys."""
        import torch
        torch.cuda.synchronize()
        for i, (m, t) in enumerate(zip(self.members, self.twins)):
            if i in light:
                assert m.edges == self.refs[i].size, i
                if self.refs[i].size:
                    assert bool(m.edge_query(self.refs[i]).cpu().numpy().all()), i
                assert int(m.view().tree_height) == int(t.view().tree_height), i
                continue
            vm, vt = m.view(), t.view()
            assert int(vm.tree_height) == int(vt.tree_height), (i, int(vm.tree_height), int(vt.tree_height))
            assert int(vm.live_keys) == int(vt.live_keys), i
            assert int(vm.version) == int(vt.version), i
            gk, gro = m.copy_view()
            tk, tro = t.copy_view()
            gk, tk = L.to_host_u64(gk), L.to_host_u64(tk)
            bad = np.nonzero(gk != tk)[0]
            assert bad.size == 0, f"member {i}: layout differs from the solo update at slots {bad[:8]}"
            assert np.array_equal(gro.cpu().numpy(), tro.cpu().numpy()), i
            ek, ro = m.export()
            ek = L.to_host_u64(ek)
            assert np.array_equal(ek, self.refs[i]), i
            assert np.array_equal(ro.cpu().numpy().view(np.uint32), oracle.row_offsets(self.refs[i], self.Vs[i])), i
            assert m.check() == 0, i

    def close(self):
        self.group.close()
        for p in self.members + self.twins:
            p.close()


def mixed_round(rng, fleet, scale, p_empty=0.2):
    ins, dels = [], []
    for i, V in enumerate(fleet.Vs):
        if rng.random() < p_empty:
            ins.append(None)
            dels.append(None)
            continue
        n = int(rng.lognormal(np.log(scale), 1.0)) + 1
        ni = int(0.7 * n) + 1
        ins.append(rand_keys(rng, V, ni))
        live = fleet.refs[i]
        nd = n - ni
        d = live[rng.integers(0, live.size, nd)] if live.size and nd > 0 else np.empty(0, dtype=np.uint64)
        dels.append(np.concatenate([d, rand_keys(rng, V, max(1, nd // 10))]))
    return ins, dels


def test_group_matches_solo_updates():
    say("group of 12, mixed rounds")
    rng = np.random.default_rng(21)
    Vs = [int(v) for v in rng.integers(16, 5000, 12)]
    f = Fleet(Vs)
    assert f.group.partitions == 1
    for r in range(12):
        ins, dels = mixed_round(rng, f, 300)
        f.round(ins, dels)
        f.compare()
    f.close()


def test_group_resizes_and_root_rebuilds_in_one_round():
    """Small members take batches many times their size (resizes and root
    rebalances) in the same rounds as members that only take window
    rebalances, then shrink again."""
    say("group with resizes")
    rng = np.random.default_rng(22)
    Vs = [8, 40, 3000, 100, 20000, 2]
    f = Fleet(Vs)
    for r in range(8):
        ins = [rand_keys(rng, V, int(rng.integers(1, 4 * V * (1 if V > 1000 else 8)))) for V in Vs]
        f.round(ins, [None] * len(Vs))
        f.compare()
    for r in range(6):
        dels = [f.refs[i][rng.integers(0, max(1, f.refs[i].size), int(0.5 * f.refs[i].size) + 1)]
                if f.refs[i].size else None for i in range(len(Vs))]
        f.round([None] * len(Vs), dels)
        f.compare()
    f.close()


def test_group_large_windows():
    """Skewed batches into a few rows force windows beyond the LDS classes
    while other members take small windows."""
    say("group with large windows")
    rng = np.random.default_rng(23)
    Vs = [1 << 14, 1 << 12, 300]
    f = Fleet(Vs)
    f.round([rand_keys(rng, V, 20 * V) for V in Vs], [None] * 3)
    f.compare()
    for r in range(4):
        hot = rng.integers(0, 8)
        ins = [keys_of(np.full(30000, hot), rng.integers(0, Vs[0], 30000)),
               rand_keys(rng, Vs[1], 200), rand_keys(rng, Vs[2], 50)]
        f.round(ins, [None] * 3)
        f.compare()
    f.close()


def test_group_mixed_segment_sizes_partition():
    say("group with mixed segment sizes")
    rng = np.random.default_rng(24)
    Vs = [500, 900, 1200, 64, 4000]
    segs = [32, 128, 32, 1024, 8]
    f = Fleet(Vs, segs)
    assert f.group.partitions == 4
    for r in range(6):
        ins, dels = mixed_round(rng, f, 400, p_empty=0.1)
        f.round(ins, dels)
        f.compare()
    f.close()


def test_group_invalid_key_changes_nothing():
    say("group invalid key")
    rng = np.random.default_rng(25)
    Vs = [100, 200, 300]
    f = Fleet(Vs)
    f.round([rand_keys(rng, V, 500) for V in Vs], [None] * 3)
    f.compare()
    before = [L.to_host_u64(m.copy_view()[0]) for m in f.members]
    vers = [int(m.view().version) for m in f.members]
    bad = [rand_keys(rng, 100, 50), keys_of([250], [5]), rand_keys(rng, 300, 50)]  # src 250 >= 200
    with pytest.raises(L.LpmaError) as e:
        f.group.update(bad, [None] * 3)
    assert e.value.code == 2
    for m, b, v in zip(f.members, before, vers):
        assert np.array_equal(L.to_host_u64(m.copy_view()[0]), b)
        assert int(m.view().version) == v
    f.round([rand_keys(rng, V, 100) for V in Vs], [None] * 3)
    f.compare()
    f.close()


def test_group_of_one_equals_plain():
    say("group of one")
    rng = np.random.default_rng(26)
    f = Fleet([5000])
    for r in range(6):
        ins, dels = mixed_round(rng, f, 2000, p_empty=0.0)
        f.round(ins, dels)
        f.compare()
    f.close()


def test_group_many_members_wide_ids():
    """65 members need 7 id bits; wide vertex ids leave fewer bits, so the
    members split into several partitions."""
    say("group of 65")
    rng = np.random.default_rng(27)
    Vs = [int(v) for v in rng.integers(2, 400, 64)] + [1 << 29]
    f = Fleet(Vs)
    assert f.group.partitions == 2
    for r in range(4):
        ins, dels = mixed_round(rng, f, 40, p_empty=0.3)
        f.round(ins, dels)
        f.compare(light=(64,))
    f.close()


def test_group_eager_rounds_match_solo_updates():
    """The same rounds without graph replay (stream-ordered launches with one
    synchronisation after the election)."""
    say("group eager rounds")
    rng = np.random.default_rng(28)
    Vs = [int(v) for v in rng.integers(16, 3000, 9)]
    f = Fleet(Vs, graph=False)
    for r in range(8):
        ins, dels = mixed_round(rng, f, 500)
        f.round(ins, dels)
        f.compare()
    f.close()


def test_plain_handle_graph_replays():
    """A handle in graph mode against an eager twin: batch sizes that grow
    (capacity changes and captures again), resizes, mixed batches, an
    invalid batch between replays, and empty batches."""
    say("plain handle graph replays")
    rng = np.random.default_rng(29)
    V = 3000
    g = L.PMA(V)
    g.set_graph(True)
    e = L.PMA(V)
    ref = np.empty(0, dtype=np.uint64)
    sizes = [10, 100, 5000, 20, 40000, 300, 120000, 7, 2000, 60000]
    for k, n in enumerate(sizes):
        ins = rand_keys(rng, V, n)
        dels = ref[rng.integers(0, ref.size, n // 3)] if ref.size else np.empty(0, dtype=np.uint64)
        g.update_mixed(ins, dels)
        e.update_mixed(ins, dels)
        ref = oracle.apply(ref, ins, dels)
        if k == 4:
            with pytest.raises(L.LpmaError):
                g.insert(keys_of([V + 1], [0]))
            with pytest.raises(L.LpmaError):
                e.insert(keys_of([V + 1], [0]))
        if k == 6:
            g.update_mixed(np.empty(0, dtype=np.uint64), np.empty(0, dtype=np.uint64))
            e.update_mixed(np.empty(0, dtype=np.uint64), np.empty(0, dtype=np.uint64))
        gk, gro = g.copy_view()
        ek, ero = e.copy_view()
        assert int(g.view().tree_height) == int(e.view().tree_height)
        assert np.array_equal(L.to_host_u64(gk), L.to_host_u64(ek)), k
        assert np.array_equal(gro.cpu().numpy(), ero.cpu().numpy()), k
        assert np.array_equal(L.to_host_u64(g.export()[0]), ref), k
        assert g.check() == 0
    p = g.profile()
    assert p.graph_launches >= len(sizes) and p.graph_captures >= 2
    g.close()
    e.close()


def test_group_reserve_captures_once():
    """After reserve(max round), rounds up to that size replay cached graphs
    (no capture), also across member resizes, and still match solo updates."""
    say("group reserve")
    rng = np.random.default_rng(30)
    Vs = [int(v) for v in rng.integers(8, 2000, 10)]
    f = Fleet(Vs)
    f.group.reserve(60000)
    f.group.set_profiling(True)
    f.group.profile()
    for r in range(10):
        ins, dels = mixed_round(rng, f, 800 * (r + 1), p_empty=0.1)
        tot = sum(a.size for a in ins if a is not None) + sum(d.size for d in dels if d is not None)
        if tot > 60000:
            continue
        f.round(ins, dels)
        f.compare()
    p = f.group.profile()
    assert p.graph_captures == 0 and p.graph_launches > 0
    assert f.group.stats().rebuilds > 0  # members were resized between replays
    f.close()


@pytest.mark.parametrize("spl", [1, 2, 8, 64])
def test_group_launch_shapes(spl):
    """Every single-segment rebalance shape gives the solo layout."""
    say(f"group launch shape {spl}")
    rng = np.random.default_rng(31)
    Vs = [int(v) for v in rng.integers(16, 3000, 6)]
    f = Fleet(Vs, segs=[64] * 6, graph=False)
    f.group.set_launch(spl)
    for r in range(5):
        ins, dels = mixed_round(rng, f, 400)
        f.round(ins, dels)
        f.compare()
    f.close()


def test_autotuner_smoke():
    """The tuner runs its trials on a tiny fleet and returns a valid choice."""
    from lpma_amd.tune import autotune
    say("autotuner")
    rng = np.random.default_rng(32)
    Vs = [256, 1024, 4096]
    initial = [np.unique(rand_keys(rng, V, 4 * V)) for V in Vs]
    ins = [[rand_keys(rng, V, 500) for _ in range(3)] for V in Vs]
    dels = [[initial[i][rng.integers(0, initial[i].size, 200)] for _ in range(3)] for i in range(3)]
    best, trials = autotune(initial, Vs, ins, dels, rounds=3, segment_slots=(32, 128), slots_per_lane=(0, 2))
    assert best.segment_slots in (32, 128) and best.value > 0
    assert len(trials) == 2 + 3 + 1
