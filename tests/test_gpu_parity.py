"""GPU parity of the batched edge-update path against the oracle.

After every batch: the exported edge keys and CSR offsets equal the oracle's
sorted merge (the contract of test_update_correctness_rpma.cu:47-67 and
test_csr.cu:14-39, extended to deletes), the gapped slot array, tree height
and marker row offsets equal the oracle's PMA model slot for slot, and the
on-device invariant checker reports nothing.
"""
import numpy as np
import pytest

import lpma_amd as L
import oracle
from gpu_util import Triple, keys_of, say

pytestmark = pytest.mark.gpu


def rand_keys(rng, V, n):
    return keys_of(rng.integers(0, V, n), rng.integers(0, V, n))


@pytest.mark.parametrize("V", [1, 2, 31, 32, 33, 1000, 70000])
def test_create_markers(V):
    say(f"create V={V}")
    t = Triple(V)
    t.compare()
    # to_csr convention with markers: row_offset[v] = E(<v) + v
    ro, col = t.pma.to_csr()
    ro = ro.cpu().numpy().view(np.uint32)
    assert np.array_equal(ro, np.arange(V + 1, dtype=np.uint32))
    t.close()


def test_small_inserts_with_duplicates():
    say("small inserts")
    rng = np.random.default_rng(1)
    t = Triple(64)
    for b in range(25):
        k = rand_keys(rng, 64, int(rng.integers(1, 200)))
        k = np.concatenate([k, k[: k.size // 3]])  # in-batch duplicates
        t.insert(k)
        t.compare()
    t.close()


def test_deletes_and_mixed():
    say("deletes and mixed")
    rng = np.random.default_rng(2)
    t = Triple(128)
    t.insert(rand_keys(rng, 128, 3000))
    t.compare()
    for b in range(20):
        live = t.ref
        d = live[rng.integers(0, live.size, 150)] if live.size else live
        d = np.concatenate([d, rand_keys(rng, 128, 50)])  # deletes of absent keys
        if b % 2:
            t.delete(d)
        else:
            ins = rand_keys(rng, 128, 200)
            ins = np.concatenate([ins, d[:20]])  # same key deleted and inserted
            t.mixed(ins, d)
        t.compare()
    # delete everything
    t.delete(t.ref.copy())
    t.compare()
    assert t.pma.edges == 0
    t.close()


@pytest.mark.parametrize("S", [2, 4, 64, 256])
def test_segment_sizes(S):
    say(f"segment size {S}")
    rng = np.random.default_rng(10 + S)
    t = Triple(300, segment_slots=S)
    for b in range(8):
        t.insert(rand_keys(rng, 300, 400))
        t.compare()
    for b in range(6):
        t.delete(t.ref[rng.integers(0, t.ref.size, 300)])
        t.compare()
    t.close()


def test_interpolated_density():
    say("interpolated density")
    rng = np.random.default_rng(5)
    t = Triple(500, lower=0.3, upper=0.8, lower_leaf=0.08, upper_leaf=1.0)
    for b in range(10):
        t.insert(rand_keys(rng, 500, 700))
        t.compare()
    for b in range(10):
        t.delete(t.ref[rng.integers(0, t.ref.size, 500)])
        t.compare()
    t.close()


def test_invalid_keys_leave_structure_unchanged():
    say("invalid keys")
    rng = np.random.default_rng(3)
    t = Triple(100)
    t.insert(rand_keys(rng, 100, 500))
    before = L.to_host_u64(t.pma.copy_view()[0])
    for bad in (keys_of([100], [0]), keys_of([0], [100]), keys_of([5], [0xFFFFFFFF]),
                np.array([L.KEY_NONE], dtype=np.uint64)):
        # one sort tile, then the invalid key in a later tile of a batch of several
        for head, tail in ((20, 0), (3000, 2000)):
            k = np.concatenate([rand_keys(rng, 100, head), bad, rand_keys(rng, 100, tail)])
            with pytest.raises(L.LpmaError) as e:
                t.pma.insert(k)
            assert e.value.code == 2
            with pytest.raises(L.LpmaError):
                t.pma.delete(k)
    after = L.to_host_u64(t.pma.copy_view()[0])
    assert np.array_equal(before, after)
    t.compare()
    t.close()


def test_empty_and_noop_batches():
    say("empty / no-op batches")
    rng = np.random.default_rng(4)
    t = Triple(50)
    t.insert(rand_keys(rng, 50, 300))
    v0 = t.pma.view().version
    t.insert(np.empty(0, dtype=np.uint64))
    t.insert(t.ref[:10].copy())       # insert-present
    t.delete(keys_of([1, 2], [3, 3]) if False else np.empty(0, dtype=np.uint64))
    absent = np.setdiff1d(rand_keys(rng, 50, 100), t.ref)
    t.delete(absent)                  # delete-absent
    assert t.pma.view().version > v0
    t.compare()
    t.close()


def test_hub_vertex_large_windows():
    """A hub vertex receiving many edges at once forces windows beyond the
    LDS classes (the multi-workgroup path) without a full rebuild."""
    say("hub vertex")
    rng = np.random.default_rng(6)
    V = 1 << 14
    t = Triple(V)
    t.insert(rand_keys(rng, V, 200000))
    t.compare()
    for hub, n in ((7, 6000), (9000, 12000), (16000, 3000)):
        dst = rng.choice(V, n, replace=False)
        t.insert(keys_of(np.full(n, hub), dst))
        t.compare()
    s = t.pma.stats()
    assert s.alg_bytes_large > 0
    t.delete(keys_of(np.full(6000, 7), np.arange(6000)))
    t.compare()
    t.close()


def test_uniform_config_shape_reduced():
    """Config-1 shape at reduced size: bulk load, 10 insert batches of absent
    edges, 10 delete batches drawn from the live set, checked after each."""
    say("config-1 shape, reduced")
    V = 1 << 12
    t = Triple(V)
    base = L.gen_uniform(V, 60000, 1)
    t.insert(base)
    t.compare()
    for b in range(10):
        t.insert(L.gen_absent(V, t.ref, 1000, 100 + b))
        t.compare()
    rng = np.random.default_rng(7)
    for b in range(10):
        t.delete(t.ref[rng.choice(t.ref.size, 1000, replace=False)])
        t.compare()
    t.close()


def test_rmat_stream_reduced():
    """Config-2 shape at reduced size: R-MAT, shuffled, half bulk-loaded, the
    rest streamed as insert batches, then equal-volume deletes."""
    say("rmat stream, reduced")
    scale = 13
    E = 1 << 17
    keys = L.shuffle(L.gen_rmat(scale, E, seed=2), 12)
    t = Triple(1 << scale)
    t.insert(keys[: E // 2])
    t.compare()
    rest = keys[E // 2:]
    B = 4000
    for i in range(0, rest.size, B):
        t.insert(rest[i:i + B])
        t.compare(layout=(i // B) % 4 == 0)
    rng = np.random.default_rng(8)
    order = rng.permutation(t.ref.size)[: E // 2]
    dels = t.ref[order]
    for i in range(0, dels.size, B):
        t.delete(dels[i:i + B])
        t.compare(layout=(i // B) % 4 == 0)
    t.close()


def test_queries():
    say("edge and neighbor queries")
    rng = np.random.default_rng(9)
    V = 2000
    t = Triple(V, model=False)
    t.insert(rand_keys(rng, V, 40000))
    t.delete(t.ref[rng.integers(0, t.ref.size, 5000)])
    t.compare()
    q = np.concatenate([t.ref[rng.integers(0, t.ref.size, 3000)], rand_keys(rng, V, 3000),
                        keys_of([0, 5], [0xFFFFFFFF, 0xFFFFFFFF])])
    res = t.pma.edge_query(q).cpu().numpy()
    want = np.isin(q, t.ref).astype(np.int32)
    assert np.array_equal(res, want)
    import torch
    srcs = torch.tensor(list(range(0, V, 7)) + [V + 5], dtype=torch.int32)
    offs, nb = t.pma.neighbors(srcs)
    offs = offs.cpu().numpy()
    nb = nb.cpu().numpy().view(np.uint32)
    src_of = (t.ref >> np.uint64(32)).astype(np.int64)
    for i, s in enumerate(srcs.tolist()):
        want = (t.ref[src_of == s] & np.uint64(0xFFFFFFFF)).astype(np.uint32)
        assert np.array_equal(nb[offs[i]:offs[i + 1]], want)
    # reference CSR convention (to_csr_rpma): row_offset[v] = E(<v) + v
    ro, col = t.pma.to_csr()
    ro = ro.cpu().numpy().view(np.uint32)
    std = oracle.row_offsets(t.ref, V)
    assert np.array_equal(ro, std + np.arange(V + 1, dtype=np.uint32))
    col = col.cpu().numpy().view(np.uint32)
    for r in (0, 1, V // 2, V - 1):
        assert col[ro[r + 1] - 1] == r  # marker slot holds its row id
    t.close()


@pytest.mark.parametrize("n", [9000, 300000, 700000, 3000000, 5000000, 9000000])
def test_batch_sizes_cover_sort_paths(n):
    """Batches that exercise every sort path: the sample-split sort (up to
    2^17 keys), the one-kernel-per-digit LSD sort (up to 2^22), and the
    histogram / scan / scatter passes above (8.4M keys take their 16-key
    tiles)."""
    say(f"batch of {n}")
    rng = np.random.default_rng(n)
    V = 1 << 16
    t = Triple(V, model=n <= 300000)
    k = rand_keys(rng, V, n)
    t.insert(np.concatenate([k, k[: n // 10]]))
    t.compare()
    t.mixed(rand_keys(rng, V, n // 2), t.ref[rng.integers(0, t.ref.size, n // 2)])
    t.compare()
    t.close()


def test_uniform_2p16_one_million_edges():
    """Uniform graph at full size: V = 2^16, 10^6 distinct uniform edges (seed 1,
    no self loops), then 10 insert batches of 10^4 absent edges and 10 delete
    batches of 10^4 live edges, compared with the oracle after every batch."""
    say("uniform 2^16 V, 10^6 edges, full size")
    V = 1 << 16
    t = Triple(V, model=False)
    t.insert(L.gen_uniform(V, 1_000_000, 1))
    t.compare(layout=False)
    for b in range(10):
        t.insert(L.gen_absent(V, t.ref, 10_000, 100 + b))
        t.compare(layout=False)
    rng = np.random.default_rng(11)
    for b in range(10):
        t.delete(t.ref[rng.choice(t.ref.size, 10_000, replace=False)])
        t.compare(layout=False)
    assert t.pma.edges == 1_000_000
    t.close()


def net_batch(batches):
    """The single mixed batch equal to applying `batches` in order: per key,
    the last request touching it decides (inside a request, inserts after
    deletes)."""
    keys, order = [], []
    for r, (ins, dels) in enumerate(batches):
        for arr, op in ((dels, 0), (ins, 1)):
            arr = np.asarray(arr, dtype=np.uint64)
            keys.append(arr)
            order.append(np.full(arr.size, 2 * r + op, dtype=np.int64))
    keys = np.concatenate(keys)
    order = np.concatenate(order)
    idx = np.lexsort((order, keys))
    k, o = keys[idx], order[idx]
    last = np.r_[k[1:] != k[:-1], True]
    k, o = k[last], o[last]
    return k[(o & 1) == 1], k[(o & 1) == 0]


@pytest.mark.parametrize("nreq", [2, 7, 40])
def test_update_sequence_equals_sequential_application(nreq):
    """A merged sequence of mixed batches with overlapping keys: the layout
    equals the model after the net batch, and the key set equals the
    batches applied one by one on a twin."""
    say(f"update sequence of {nreq}")
    rng = np.random.default_rng(40 + nreq)
    V = 300
    t = Triple(V)
    t.insert(rand_keys(rng, V, 4000))
    twin = L.PMA(V)
    twin.insert(t.ref)
    for rep in range(4):
        pool = np.concatenate([t.ref[rng.integers(0, t.ref.size, 200)], rand_keys(rng, V, 200)])
        batches = []
        for r in range(nreq):
            ins = pool[rng.integers(0, pool.size, int(rng.integers(0, 120)))]
            dels = pool[rng.integers(0, pool.size, int(rng.integers(0, 120)))]
            batches.append((ins, dels))
        t.pma.update_sequence(batches)
        for ins, dels in batches:
            twin.update_mixed(ins, dels)
        ni, nd = net_batch(batches)
        t.model.update_mixed(ni, nd)
        t.ref = oracle.apply(t.ref, ni, nd)
        t.compare()
        assert np.array_equal(L.to_host_u64(twin.export()[0]), t.ref)
    twin.close()
    t.close()


@pytest.mark.parametrize("n", [1, 2, 1023, 1024, 1025, 4097, 65536, 131072, 131073])
def test_sample_split_sort_sizes(n):
    """Batch sizes at the sample-split sort's tile and bucket boundaries and
    just past its limit (2^17 + 1 takes the LSD path)."""
    say(f"sample-split sizes n={n}")
    rng = np.random.default_rng(1000 + n)
    V = 1 << 14
    t = Triple(V, model=n <= 65536)
    t.insert(rand_keys(rng, V, 50000))
    t.insert(rand_keys(rng, V, n))
    t.compare()
    t.mixed(rand_keys(rng, V, n), t.ref[rng.integers(0, t.ref.size, n)])
    t.compare()
    t.close()


def test_sample_split_sort_skew():
    """Skewed batches: most keys on one hub (a sample split still cuts the
    hub's keys into buckets), one key repeated 10^5 times (one bucket, no
    bits vary), and 3 * 10^4 inserts with 3 * 10^4 deletes of one key (an
    oversized bucket whose keys differ in the operation bit only: the
    passes through global memory)."""
    say("sample-split skew")
    rng = np.random.default_rng(77)
    V = 1 << 15
    t = Triple(V)
    t.insert(rand_keys(rng, V, 40000))
    t.compare()
    hub = keys_of(np.full(70000, 123), rng.integers(0, V, 70000))
    t.insert(np.concatenate([hub, rand_keys(rng, V, 30000)]))
    t.compare()
    k = keys_of([777], [4242])
    t.insert(np.repeat(k, 100000))
    t.compare()
    t.mixed(np.repeat(k, 30000), np.repeat(k, 30000))
    t.compare()
    t.mixed(np.repeat(t.ref[:5], 8000), np.repeat(t.ref[5:10], 8000))
    t.compare()
    # an oversized bucket over many bits: 6000 copies each of 3 keys of
    # one row beside 50000 spread keys
    few = keys_of([9, 9, 9], [1, 1 << 14, (1 << 15) - 1])
    t.mixed(np.concatenate([np.repeat(few, 6000), rand_keys(rng, V, 50000)]), np.repeat(few[:1], 7000))
    t.compare()
    t.close()
