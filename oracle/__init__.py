"""CPU oracle for the batched edge-update path (test infrastructure only).

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this package.  The product (``lpma_amd``) never does.

``liboracle.so`` is built from ``oracle/oracle.c`` by ``oracle/Makefile``.
Two parts (the C file's header gives the reference citations):

* ``apply`` / ``row_offsets``: the set contract.  After every batch the live
  keys must equal ``sort(unique((S - D) | I))`` and the row offsets must count
  keys per source vertex.
* ``PmaModel``: the packed memory array in logical slot space, restating the
  reference's per-level update loop; the GPU's gapped slot array must match it
  slot for slot.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

KEY_NONE = 0xFFFFFFFFFFFFFFFF

_u64p = ctypes.POINTER(ctypes.c_uint64)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def build():
    """Compile liboracle.so with gcc (no GPU toolchain needed)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.or_sort_unique.argtypes = [_u64p, ctypes.c_size_t]
        L.or_sort_unique.restype = ctypes.c_size_t
        L.or_apply.argtypes = [_u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t,
                               _u64p, ctypes.c_int]
        L.or_apply.restype = ctypes.c_size_t
        L.or_row_offsets.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_uint32, _u32p]
        L.or_row_offsets.restype = ctypes.c_int
        L.or_row_offsets_mt.argtypes = [_u64p, ctypes.c_size_t, ctypes.c_uint32, _u32p, ctypes.c_int]
        L.or_row_offsets_mt.restype = ctypes.c_int
        L.or_is_sorted_unique.argtypes = [_u64p, ctypes.c_size_t]
        L.or_is_sorted_unique.restype = ctypes.c_int
        L.or_pma_create.argtypes = [ctypes.c_uint32, ctypes.c_uint32, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double]
        L.or_pma_create.restype = ctypes.c_void_p
        L.or_pma_destroy.argtypes = [ctypes.c_void_p]
        L.or_pma_update.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_size_t, ctypes.c_int]
        L.or_pma_update.restype = ctypes.c_int
        L.or_pma_update_mixed.argtypes = [ctypes.c_void_p, _u64p, ctypes.c_size_t, _u64p, ctypes.c_size_t]
        L.or_pma_update_mixed.restype = ctypes.c_int
        for name, rt in (("or_pma_capacity", ctypes.c_uint64), ("or_pma_height", ctypes.c_uint32),
                         ("or_pma_live", ctypes.c_uint64), ("or_pma_rebuilds", ctypes.c_uint64)):
            getattr(L, name).argtypes = [ctypes.c_void_p]
            getattr(L, name).restype = rt
        L.or_pma_keys.argtypes = [ctypes.c_void_p]
        L.or_pma_keys.restype = _u64p
        L.or_pma_sep.argtypes = [ctypes.c_void_p]
        L.or_pma_sep.restype = _u64p
        L.or_pma_cnt.argtypes = [ctypes.c_void_p]
        L.or_pma_cnt.restype = _u32p
        L.or_pma_row_offsets.argtypes = [ctypes.c_void_p]
        L.or_pma_row_offsets.restype = _u32p
        L.or_pma_windows_at_level.argtypes = [ctypes.c_void_p, ctypes.c_uint32]
        L.or_pma_windows_at_level.restype = ctypes.c_uint64
        L.or_pma_export.argtypes = [ctypes.c_void_p, _u64p]
        L.or_pma_export.restype = ctypes.c_uint64
        L.or_bfs.argtypes = [_u32p, _u32p, ctypes.c_uint32, ctypes.c_uint32, _u32p]
        L.or_bfs.restype = ctypes.c_uint32
        L.or_pagerank.argtypes = [_u32p, _u32p, ctypes.c_uint32, ctypes.c_double, ctypes.c_uint32,
                                  ctypes.POINTER(ctypes.c_double)]
        L.or_pagerank.restype = ctypes.c_double
        _lib = L
    return _lib


def _p64(a):
    return a.ctypes.data_as(_u64p)


def _as_u64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64))


def sort_unique(a):
    a = _as_u64(a).copy()
    n = lib().or_sort_unique(_p64(a), a.size)
    return a[:n]


def apply(S, ins=(), dels=(), threads=1, out=None):
    """(S \\ dels) | ins as a sorted unique uint64 array.  S must be sorted unique.
    `out` (optional, not overlapping S): a contiguous uint64 buffer of at least
    S.size + len(ins) + 1 entries to merge into; the result is a view of it."""
    S = _as_u64(S)
    I = _as_u64(ins)
    D = _as_u64(dels)
    need = S.size + I.size + 1
    if out is None:
        out = np.empty(need, dtype=np.uint64)
    elif out.dtype != np.uint64 or not out.flags.c_contiguous or out.size < need or np.may_share_memory(out, S):
        raise ValueError("out must be a contiguous uint64 buffer of S.size + len(ins) + 1 entries apart from S")
    n = lib().or_apply(_p64(S), S.size, _p64(I), I.size, _p64(D), D.size, _p64(out), int(threads))
    return out[:n]


def row_offsets(keys, V, threads=1):
    """Standard CSR offsets: ro[v] = #edge keys with src < v (row markers skipped)."""
    keys = _as_u64(keys)
    ro = np.empty(int(V) + 1, dtype=np.uint32)
    if threads > 1:
        r = lib().or_row_offsets_mt(_p64(keys), keys.size, int(V), ro.ctypes.data_as(_u32p), int(threads))
    else:
        r = lib().or_row_offsets(_p64(keys), keys.size, int(V), ro.ctypes.data_as(_u32p))
    if r != 0:
        raise ValueError("key with src >= V")
    return ro


def csr_of(keys, V):
    """(row offsets, dst ids) of a sorted edge-key array."""
    keys = _as_u64(keys)
    ro = row_offsets(keys, V)
    col = np.ascontiguousarray((keys & np.uint64(0xFFFFFFFF)).astype(np.uint32))
    return ro, col


def bfs(keys, V, src):
    """BFS levels (0xFFFFFFFF unreached) over the sorted edge keys."""
    ro, col = csr_of(keys, V)
    lv = np.empty(int(V), dtype=np.uint32)
    lib().or_bfs(ro.ctypes.data_as(_u32p), col.ctypes.data_as(_u32p), int(V), int(src), lv.ctypes.data_as(_u32p))
    return lv


def pagerank(keys, V, alpha=0.85, iters=50):
    """FP64 PageRank after exactly `iters` iterations (L1-normalised), and the
    L1 change of the last iteration."""
    ro, col = csr_of(keys, V)
    pr = np.empty(int(V), dtype=np.float64)
    d = lib().or_pagerank(ro.ctypes.data_as(_u32p), col.ctypes.data_as(_u32p), int(V), float(alpha), int(iters),
                          pr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return pr, d


def is_sorted_unique(keys):
    keys = _as_u64(keys)
    return bool(lib().or_is_sorted_unique(_p64(keys), keys.size))


def marker_key(r):
    return (int(r) << 32) | 0xFFFFFFFF


class PmaModel:
    """CPU packed-memory-array model in logical slot space (see oracle.c)."""

    def __init__(self, V, segment_slots=32, lower_leaf=0.42, lower_root=0.42, upper_leaf=1.0, upper_root=1.0):
        self._L = lib()
        self.V = int(V)
        self._h = self._L.or_pma_create(self.V, int(segment_slots), lower_leaf, lower_root, upper_leaf, upper_root)
        if not self._h:
            raise ValueError("bad PMA model configuration")

    def close(self):
        if self._h:
            self._L.or_pma_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def update(self, keys, op):
        keys = _as_u64(keys)
        r = self._L.or_pma_update(self._h, _p64(keys), keys.size, 1 if op else 0)
        if r != 0:
            raise ValueError("invalid key")

    def insert(self, keys):
        self.update(keys, 1)

    def delete(self, keys):
        self.update(keys, 0)

    def update_mixed(self, ins, dels):
        I = _as_u64(ins)
        D = _as_u64(dels)
        r = self._L.or_pma_update_mixed(self._h, _p64(I), I.size, _p64(D), D.size)
        if r != 0:
            raise ValueError("invalid key")

    @property
    def capacity(self):
        return int(self._L.or_pma_capacity(self._h))

    @property
    def height(self):
        return int(self._L.or_pma_height(self._h))

    @property
    def live(self):
        return int(self._L.or_pma_live(self._h))

    @property
    def rebuilds(self):
        return int(self._L.or_pma_rebuilds(self._h))

    def windows_at_level(self, level):
        return int(self._L.or_pma_windows_at_level(self._h, level))

    def keys(self):
        return np.ctypeslib.as_array(self._L.or_pma_keys(self._h), shape=(self.capacity,)).copy()

    def cnt(self):
        return np.ctypeslib.as_array(self._L.or_pma_cnt(self._h), shape=(1 << self.height,)).copy()

    def sep(self):
        return np.ctypeslib.as_array(self._L.or_pma_sep(self._h), shape=(1 << self.height,)).copy()

    def row_offsets(self):
        return np.ctypeslib.as_array(self._L.or_pma_row_offsets(self._h), shape=(self.V + 1,)).copy()

    def export(self):
        out = np.empty(self.live + 1, dtype=np.uint64)
        n = self._L.or_pma_export(self._h, _p64(out))
        return out[:n]
