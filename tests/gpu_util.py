"""Helpers shared by the GPU parity tests: drive the HIP store, the oracle's
set contract and the oracle's PMA model with the same batches, and compare
after every batch."""
import sys

import numpy as np

import lpma_amd as L
import oracle

U32 = np.uint64(32)


def keys_of(src, dst):
    return (np.asarray(src, dtype=np.uint64) << U32) | np.asarray(dst, dtype=np.uint64)


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def with_markers(ref, ro, V):
    """The live keys get_data_rpma returns (rpma.cuh:1735-1761): the edges
    and one row marker r<<32|0xFFFFFFFF per vertex, in key order (a row's
    marker follows its edges, at the row's end offset)."""
    markers = (np.arange(V, dtype=np.uint64) << U32) | np.uint64(0xFFFFFFFF)
    return np.insert(ref, ro[1:].astype(np.int64), markers)


def check_get_data(pma, ref, ro, V):
    got = L.to_host_u64(pma.get_data())
    want = with_markers(ref, ro, V)
    assert got.size == want.size, (got.size, want.size)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, f"get_data differs at {bad[:5]}: got {got[bad[:5]]}, want {want[bad[:5]]}"


class Triple:
    """The HIP store, the sorted-merge oracle and the PMA layout model, kept in
    lock step.  `model=False` skips the layout model (large cases)."""

    def __init__(self, V, segment_slots=32, lower=0.42, upper=1.0, lower_leaf=None, upper_leaf=None, model=True):
        self.V = V
        ll = lower if lower_leaf is None else lower_leaf
        ul = upper if upper_leaf is None else upper_leaf
        self.pma = L.PMA(V, segment_slots=segment_slots, density_lower=lower, density_upper=upper,
                         density_lower_leaf=ll, density_upper_leaf=ul)
        self.model = oracle.PmaModel(V, segment_slots, ll, lower, ul, upper) if model else None
        self.ref = np.empty(0, dtype=np.uint64)

    def insert(self, keys):
        keys = np.asarray(keys, dtype=np.uint64)
        self.pma.insert(keys)
        if self.model is not None:
            self.model.insert(keys)
        self.ref = oracle.apply(self.ref, keys, ())

    def delete(self, keys):
        keys = np.asarray(keys, dtype=np.uint64)
        self.pma.delete(keys)
        if self.model is not None:
            self.model.delete(keys)
        self.ref = oracle.apply(self.ref, (), keys)

    def mixed(self, ins, dels):
        ins = np.asarray(ins, dtype=np.uint64)
        dels = np.asarray(dels, dtype=np.uint64)
        self.pma.update_mixed(ins, dels)
        if self.model is not None:
            self.model.update_mixed(ins, dels)
        self.ref = oracle.apply(self.ref, ins, dels)

    def compare(self, layout=True):
        import torch
        torch.cuda.synchronize()
        v = self.pma.view()
        assert int(v.edges) == self.ref.size, (int(v.edges), self.ref.size)
        ek, ro = self.pma.export()
        ek = L.to_host_u64(ek)
        ro = ro.cpu().numpy().view(np.uint32)
        assert ek.size == self.ref.size
        bad = np.nonzero(ek != self.ref)[0]
        assert bad.size == 0, f"first mismatch at {bad[:5]}: got {ek[bad[:5]]}, want {self.ref[bad[:5]]}"
        want_ro = oracle.row_offsets(self.ref, self.V)
        assert np.array_equal(ro, want_ro)
        check_get_data(self.pma, self.ref, want_ro, self.V)
        if layout and self.model is not None:
            gk, gro = self.pma.copy_view()
            gk = L.to_host_u64(gk)
            gro = gro.cpu().numpy().view(np.uint32)
            mk = self.model.keys()
            assert int(v.tree_height) == self.model.height, (int(v.tree_height), self.model.height)
            assert gk.size == mk.size
            bad = np.nonzero(gk != mk)[0]
            assert bad.size == 0, f"layout differs from the model at slots {bad[:8]}"
            assert np.array_equal(gro, self.model.row_offsets())
        assert self.pma.check() == 0

    def close(self):
        self.pma.close()
        if self.model is not None:
            self.model.close()
