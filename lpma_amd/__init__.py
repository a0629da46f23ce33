"""MI355X-native batched edge-update store (LPMA's update path) -- Python host mirror.

Everything runs in ``liblpma.so`` (hand-written HIP kernels for gfx950 behind
the C ABI declared in ``include/lpma.h``); this module only binds that ABI with
ctypes and uses torch for device buffers and streams.  There is no CPU
fallback: without the built library every call raises.

The reference is a header-only CUDA C++ library with free functions over a
struct (rpma.cuh, gpma.cuh).  Its names are mirrored here:

    ==========================  =========================================
    reference (rpma.cuh)        here
    ==========================  =========================================
    init_csr_rpma :1807         ``init_csr_rpma(V)`` / ``PMA(V)``
    update_rpma :1549           ``update_rpma(pma, keys, op)`` / ``PMA.update``
    get_data_rpma :1735         ``get_data_rpma(pma)`` / ``PMA.get_data``
    to_csr_rpma :2249           ``to_csr_rpma(pma)`` / ``PMA.to_csr``
    edge_query_rpma :2167       ``edge_query_rpma(pma, keys)``
    query_rpma :1974            ``query_rpma(pma, srcs)`` (no time window)
    destroy_rpma :1720          ``destroy_rpma(pma)`` / ``PMA.close``
    ==========================  =========================================

Keys are ``src << 32 | dst`` held in int64 tensors (the bit pattern of the
uint64 key; KEY_NONE reads as -1).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblpma.so")

KEY_NONE = 0xFFFFFFFFFFFFFFFF
OP_DELETE = 0
OP_INSERT = 1

_STATUS = {
    0: "ok",
    1: "invalid argument",
    2: "invalid key (src or dst >= num_vertices)",
    3: "out of device memory or pool budget",
    4: "HIP runtime error",
    5: "batch or capacity too large",
    6: "device code not available",
}


class LpmaError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"{what}: lpma status {code} ({_STATUS.get(code, 'unknown')})")


class Config(ctypes.Structure):
    _fields_ = [("segment_slots", ctypes.c_uint32), ("density_lower_leaf", ctypes.c_double),
                ("density_lower_root", ctypes.c_double), ("density_upper_leaf", ctypes.c_double),
                ("density_upper_root", ctypes.c_double), ("device", ctypes.c_int)]


class View(ctypes.Structure):
    _fields_ = [("keys", ctypes.c_void_p), ("row_offsets", ctypes.c_void_p), ("capacity", ctypes.c_uint64),
                ("live_keys", ctypes.c_uint64), ("edges", ctypes.c_uint64), ("num_vertices", ctypes.c_uint32),
                ("segment_slots", ctypes.c_uint32), ("tree_height", ctypes.c_uint32), ("version", ctypes.c_uint64)]


class Stats(ctypes.Structure):
    _fields_ = [("updates_calls", ctypes.c_uint64), ("effective_updates", ctypes.c_uint64),
                ("windows_per_level", ctypes.c_uint64 * 32), ("slots_per_level", ctypes.c_uint64 * 32),
                ("rebuilds", ctypes.c_uint64), ("rebuild_slots", ctypes.c_uint64), ("marker_moves", ctypes.c_uint64),
                ("alg_bytes_wave", ctypes.c_uint64), ("alg_bytes_block", ctypes.c_uint64),
                ("alg_bytes_large", ctypes.c_uint64), ("alg_bytes_seg", ctypes.c_uint64)]


class Profile(ctypes.Structure):
    _fields_ = [("sort_ms", ctypes.c_double), ("locate_ms", ctypes.c_double), ("elect_ms", ctypes.c_double),
                ("rebalance_wave_ms", ctypes.c_double), ("rebalance_block_ms", ctypes.c_double),
                ("large_ms", ctypes.c_double), ("rebuild_ms", ctypes.c_double), ("pyramid_ms", ctypes.c_double),
                ("rebalance_seg_ms", ctypes.c_double), ("launches_wave", ctypes.c_uint64),
                ("launches_block", ctypes.c_uint64), ("launches_seg", ctypes.c_uint64), ("graph_ms", ctypes.c_double),
                ("graph_launches", ctypes.c_uint64), ("graph_captures", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("bfs_workers", ctypes.c_uint32), ("hbm_budget_bytes", ctypes.c_uint64),
                ("max_merge_keys", ctypes.c_uint64), ("max_merge_requests", ctypes.c_uint32),
                ("max_merge_vertices", ctypes.c_uint64), ("snapshot_max_age_us", ctypes.c_uint64),
                ("search_workgroups", ctypes.c_uint32)]


class ServerResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("kind", ctypes.c_int), ("applied_updates", ctypes.c_uint64),
                ("version", ctypes.c_uint64), ("queue_us", ctypes.c_double), ("latency_us", ctypes.c_double),
                ("count", ctypes.c_uint64), ("offsets", ctypes.POINTER(ctypes.c_uint32)),
                ("values", ctypes.POINTER(ctypes.c_uint32)), ("depth", ctypes.c_uint32),
                ("edges_scanned", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64 * 3), ("completed", ctypes.c_uint64 * 3),
                ("update_runs", ctypes.c_uint64), ("read_runs", ctypes.c_uint64), ("snapshots", ctypes.c_uint64),
                ("max_concurrent_bfs", ctypes.c_uint64), ("admission_waits", ctypes.c_uint64),
                ("budget_bytes", ctypes.c_uint64), ("in_use_bytes", ctypes.c_uint64), ("peak_bytes", ctypes.c_uint64),
                ("applied_updates", ctypes.c_uint64), ("update_us", ctypes.c_double), ("read_us", ctypes.c_double),
                ("dispatch_us", ctypes.c_double), ("stage_us", ctypes.c_double), ("pipeline_us", ctypes.c_double),
                ("finish_us", ctypes.c_double)]


REQ_UPDATE, REQ_READ, REQ_BFS = 0, 1, 2

_vp = ctypes.c_void_p
_u64 = ctypes.c_uint64
_u32 = ctypes.c_uint32
_int = ctypes.c_int

# (name, restype, argtypes) for every entry point of include/lpma.h
SIGNATURES = [
    ("lpma_config_default", None, [ctypes.POINTER(Config)]),
    ("lpma_create", _int, [ctypes.POINTER(_vp), _u32, ctypes.POINTER(Config), _vp]),
    ("lpma_destroy", _int, [_vp]),
    ("lpma_update", _int, [_vp, _vp, _u64, _int, _vp]),
    ("lpma_update_mixed", _int, [_vp, _vp, _u64, _vp, _u64, _vp]),
    ("lpma_bulk_load", _int, [_vp, _vp, _u64, _vp]),
    ("lpma_update_sequence", _int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_u64), ctypes.POINTER(_vp),
                                    ctypes.POINTER(_u64), _u32, _vp]),
    ("lpma_get_view", _int, [_vp, ctypes.POINTER(View)]),
    ("lpma_copy_view", _int, [_vp, _vp, _vp, _vp]),
    ("lpma_export_compacted", _int, [_vp, _vp, _vp, _vp]),
    ("lpma_export_compacted_noalloc", _int, [_vp, _vp, _vp, _vp]),
    ("lpma_export_data", _int, [_vp, _vp, _vp]),
    ("lpma_to_csr", _int, [_vp, _vp, _vp, _vp]),
    ("lpma_edge_query", _int, [_vp, _vp, _u64, _vp, _vp]),
    ("lpma_neighbor_count", _int, [_vp, _vp, _u64, _vp, _vp]),
    ("lpma_neighbor_gather", _int, [_vp, _vp, _u64, _vp, _vp, _vp]),
    ("lpma_check", _int, [_vp, ctypes.POINTER(_u64), _vp]),
    ("lpma_bfs", _int, [_vp, _u32, _vp, ctypes.POINTER(_u32), ctypes.POINTER(_u64), _vp]),
    ("lpma_pagerank", _int, [_vp, ctypes.c_float, ctypes.c_float, _u32, _vp, ctypes.POINTER(_u32),
                             ctypes.POINTER(ctypes.c_float), _vp]),
    ("lpma_get_stats", _int, [_vp, ctypes.POINTER(Stats)]),
    ("lpma_reset_stats", _int, [_vp]),
    ("lpma_set_profiling", _int, [_vp, _int]),
    ("lpma_get_profile", _int, [_vp, ctypes.POINTER(Profile)]),
    ("lpma_set_graph", _int, [_vp, _int]),
    ("lpma_reserve", _int, [_vp, _u64, _vp]),
    ("lpma_reserve_capacity", _int, [_vp, _u64, _vp]),
    ("lpma_set_launch", _int, [_vp, _u32]),
    ("lpma_set_pool_budget", _int, [_u64]),
    ("lpma_pool_trim", _int, []),
    ("lpma_pool_reserved_bytes", _u64, []),
    ("lpma_inject_alloc_failure", _int, [_u64, ctypes.POINTER(_u64)]),
    ("lpma_strerror", ctypes.c_char_p, [_int]),
    ("lpma_server_config_default", None, [ctypes.POINTER(ServerConfig)]),
    ("lpma_server_create", _int, [ctypes.POINTER(_vp), _vp, ctypes.POINTER(ServerConfig)]),
    ("lpma_server_destroy", _int, [_vp]),
    ("lpma_server_submit_update", _int, [_vp, _vp, _u64, _vp, _u64, ctypes.POINTER(_u64)]),
    ("lpma_server_submit_read", _int, [_vp, _vp, _u64, ctypes.POINTER(_u64)]),
    ("lpma_server_submit_bfs", _int, [_vp, _u32, _int, ctypes.POINTER(_u64)]),
    ("lpma_server_wait", _int, [_vp, _u64, ctypes.POINTER(ServerResult)]),
    ("lpma_server_poll", _int, [_vp, _u64, ctypes.POINTER(_int), ctypes.POINTER(ServerResult)]),
    ("lpma_server_release", _int, [_vp, _u64]),
    ("lpma_server_get_stats", _int, [_vp, ctypes.POINTER(ServerStats)]),
    ("lpma_snapshot_bytes", _u64, [_vp]),
    ("lpma_snapshot_create", _int, [ctypes.POINTER(_vp), _vp, _vp]),
    ("lpma_snapshot_destroy", _int, [_vp]),
    ("lpma_snapshot_get_view", _int, [_vp, ctypes.POINTER(View)]),
    ("lpma_snapshot_neighbor_count", _int, [_vp, _vp, _u64, _vp, _vp]),
    ("lpma_snapshot_neighbor_gather", _int, [_vp, _vp, _u64, _vp, _vp, _vp]),
    ("lpma_snapshot_edge_query", _int, [_vp, _vp, _u64, _vp, _vp]),
    ("lpma_snapshot_bfs", _int, [_vp, _u32, _vp, ctypes.POINTER(_u32), ctypes.POINTER(_u64), _vp]),
    ("lpma_snapshot_set_search_width", _int, [_vp, _u32]),
    ("lpma_group_create", _int, [ctypes.POINTER(_vp), ctypes.POINTER(_vp), _u32]),
    ("lpma_group_destroy", _int, [_vp]),
    ("lpma_group_update", _int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_u64), ctypes.POINTER(_vp),
                                 ctypes.POINTER(_u64), _vp]),
    ("lpma_group_partitions", _u32, [_vp]),
    ("lpma_group_get_stats", _int, [_vp, ctypes.POINTER(Stats)]),
    ("lpma_group_set_profiling", _int, [_vp, _int]),
    ("lpma_group_set_graph", _int, [_vp, _int]),
    ("lpma_group_reserve", _int, [_vp, _u64, _vp]),
    ("lpma_group_set_launch", _int, [_vp, _u32]),
    ("lpma_group_get_profile", _int, [_vp, ctypes.POINTER(Profile)]),
    ("lpma_gen_uniform", _int, [_u32, _u64, _u64, _vp]),
    ("lpma_gen_rmat", _int, [_u32, _u64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u64, _vp]),
    ("lpma_gen_rmat_samples", _int, [_u32, _u64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u64, _vp]),
    ("lpma_shuffle", _int, [_vp, _u64, _u64]),
    ("lpma_gen_rmat_device", _int, [_u32, _u64, ctypes.c_double, ctypes.c_double, ctypes.c_double, _u64, _vp, _vp]),
    ("lpma_gen_absent", _int, [_u32, _vp, _u64, _u64, _u64, _vp]),
    ("lpma_read_edgelist", _int, [ctypes.c_char_p, ctypes.POINTER(_u32), ctypes.POINTER(_u64), _u64, _vp]),
    ("lpma_mt64_fill", _int, [_u64, _u64, _vp]),
]

_lib = None


def lib():
    """Load liblpma.so; raises when it has not been built (no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m lpma_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(code, what):
    if code != 0:
        raise LpmaError(code, what)


# ----------------------------------------------------------------- inputs

def _np_u64(n):
    return np.empty(int(n), dtype=np.uint64)


def gen_uniform(V, n, seed):
    out = _np_u64(n)
    _check(lib().lpma_gen_uniform(V, n, seed, out.ctypes.data), "gen_uniform")
    return out


def gen_rmat(scale, n, a=0.57, b=0.19, c=0.19, seed=2):
    out = _np_u64(n)
    _check(lib().lpma_gen_rmat(scale, n, a, b, c, seed, out.ctypes.data), "gen_rmat")
    return out


def gen_rmat_samples(scale, n, a=0.57, b=0.19, c=0.19, seed=2):
    """R-MAT edges drawn with replacement (repeats kept, no self loops)."""
    out = _np_u64(n)
    _check(lib().lpma_gen_rmat_samples(scale, n, a, b, c, seed, out.ctypes.data), "gen_rmat_samples")
    return out


def gen_rmat_device(scale, n, a=0.57, b=0.19, c=0.19, seed=4, device=None, stream=None):
    """R-MAT samples drawn on the GPU (int64 tensor holding the uint64 keys)."""
    torch = _torch()
    dev = torch.device("cuda", torch.cuda.current_device() if device is None else int(device))
    out = torch.empty(max(int(n), 1), dtype=torch.int64, device=dev)
    _check(lib().lpma_gen_rmat_device(scale, int(n), a, b, c, seed, out.data_ptr(), _stream_ptr(stream)),
           "gen_rmat_device")
    return out[: int(n)]


def shuffle(keys, seed):
    keys = np.ascontiguousarray(keys, dtype=np.uint64)
    _check(lib().lpma_shuffle(keys.ctypes.data, keys.size, seed), "shuffle")
    return keys


def gen_absent(V, present_sorted, n, seed):
    p = np.ascontiguousarray(present_sorted, dtype=np.uint64)
    out = _np_u64(n)
    _check(lib().lpma_gen_absent(V, p.ctypes.data, p.size, n, seed, out.ctypes.data), "gen_absent")
    return out


def read_edgelist(path, max_edges=0):
    V = _u32(0)
    E = _u64(0)
    _check(lib().lpma_read_edgelist(path.encode(), ctypes.byref(V), ctypes.byref(E), max_edges, None), "edgelist")
    out = _np_u64(E.value)
    _check(lib().lpma_read_edgelist(path.encode(), ctypes.byref(V), ctypes.byref(E), max_edges, out.ctypes.data),
           "edgelist")
    return V.value, out


def mt64(seed, n):
    out = _np_u64(n)
    _check(lib().lpma_mt64_fill(seed, n, out.ctypes.data), "mt64")
    return out


# ------------------------------------------------------------ device side

def set_pool_budget(nbytes):
    """Caps the bytes the device pool holds from hipMalloc (0 = unlimited)."""
    _check(lib().lpma_set_pool_budget(int(nbytes)), "set_pool_budget")


def pool_reserved_bytes():
    return int(lib().lpma_pool_reserved_bytes())


def pool_trim():
    _check(lib().lpma_pool_trim(), "pool_trim")


def inject_alloc_failure(nth):
    """The nth allocation from now and every later one fail (0 disables).
    Returns the countdown of the injection this call replaced (0: it fired,
    or none was set)."""
    rem = _u64(0)
    _check(lib().lpma_inject_alloc_failure(int(nth), ctypes.byref(rem)), "inject_alloc_failure")
    return int(rem.value)


def _torch():
    import torch
    return torch


def to_device(keys, device="cuda"):
    """uint64 numpy keys -> int64 device tensor with the same bits."""
    torch = _torch()
    a = np.ascontiguousarray(keys, dtype=np.uint64).view(np.int64)
    return torch.from_numpy(a).to(device)


def to_host_u64(t):
    return t.detach().cpu().numpy().view(np.uint64)


def _stream_ptr(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class PMA:
    """One packed-memory-array edge store on a GPU (a handle of the C ABI)."""

    def __init__(self, num_vertices, segment_slots=32, density_lower=0.42, density_upper=1.0,
                 density_lower_leaf=None, density_upper_leaf=None, device=None, stream=None):
        torch = _torch()
        L = lib()
        cfg = Config()
        L.lpma_config_default(ctypes.byref(cfg))
        cfg.segment_slots = segment_slots
        cfg.density_lower_root = density_lower
        cfg.density_upper_root = density_upper
        cfg.density_lower_leaf = density_lower if density_lower_leaf is None else density_lower_leaf
        cfg.density_upper_leaf = density_upper if density_upper_leaf is None else density_upper_leaf
        if device is None:
            device = torch.cuda.current_device()
        cfg.device = int(device)
        self.device = torch.device("cuda", int(device))
        self.stream = stream
        self.num_vertices = int(num_vertices)
        h = ctypes.c_void_p()
        _check(L.lpma_create(ctypes.byref(h), self.num_vertices, ctypes.byref(cfg), _stream_ptr(stream)), "create")
        self._h = h

    # -- lifecycle
    def close(self):
        if getattr(self, "_h", None):
            lib().lpma_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- updates
    def _dev_keys(self, keys):
        torch = _torch()
        if isinstance(keys, np.ndarray) or isinstance(keys, (list, tuple)):
            keys = to_device(np.asarray(keys, dtype=np.uint64), self.device)
        if keys.dtype != torch.int64 or not keys.is_cuda:
            raise TypeError("keys must be an int64 device tensor (bit pattern of src<<32|dst)")
        return keys.contiguous()

    def update(self, keys, op):
        k = self._dev_keys(keys)
        _check(lib().lpma_update(self._h, k.data_ptr(), k.numel(), int(op), _stream_ptr(self.stream)), "update")

    def insert(self, keys):
        self.update(keys, OP_INSERT)

    def delete(self, keys):
        self.update(keys, OP_DELETE)

    def update_mixed(self, ins, dels):
        i = self._dev_keys(ins)
        d = self._dev_keys(dels)
        _check(lib().lpma_update_mixed(self._h, i.data_ptr(), i.numel(), d.data_ptr(), d.numel(),
                                       _stream_ptr(self.stream)), "update_mixed")

    def update_sequence(self, batches):
        """Applies [(ins, dels), ...] in order as one pipeline run."""
        n = len(batches)
        keep = []
        pi = (ctypes.c_void_p * max(n, 1))()
        ni = (_u64 * max(n, 1))()
        pd = (ctypes.c_void_p * max(n, 1))()
        nd = (_u64 * max(n, 1))()
        for i, (a, d) in enumerate(batches):
            for arr, p_, n_ in ((a, pi, ni), (d, pd, nd)):
                if arr is None or len(arr) == 0:
                    continue
                t = self._dev_keys(arr)
                keep.append(t)
                p_[i] = t.data_ptr()
                n_[i] = t.numel()
        _check(lib().lpma_update_sequence(self._h, pi, ni, pd, nd, n, _stream_ptr(self.stream)), "update_sequence")

    def bulk_load(self, keys):
        k = self._dev_keys(keys)
        _check(lib().lpma_bulk_load(self._h, k.data_ptr(), k.numel(), _stream_ptr(self.stream)), "bulk_load")

    # -- views
    def view(self):
        v = View()
        _check(lib().lpma_get_view(self._h, ctypes.byref(v)), "view")
        return v

    @property
    def edges(self):
        return int(self.view().edges)

    def copy_view(self):
        """(gapped keys, marker row offsets) as new device tensors."""
        torch = _torch()
        v = self.view()
        keys = torch.empty(int(v.capacity), dtype=torch.int64, device=self.device)
        ro = torch.empty(self.num_vertices + 1, dtype=torch.int32, device=self.device)
        _check(lib().lpma_copy_view(self._h, keys.data_ptr(), ro.data_ptr(), _stream_ptr(self.stream)), "copy_view")
        return keys, ro

    def export(self):
        """(edge keys, standard CSR row offsets) as device tensors -- the parity surface."""
        torch = _torch()
        v = self.view()
        keys = torch.empty(max(int(v.edges), 1), dtype=torch.int64, device=self.device)
        ro = torch.empty(self.num_vertices + 1, dtype=torch.int32, device=self.device)
        _check(lib().lpma_export_compacted(self._h, keys.data_ptr(), ro.data_ptr(), _stream_ptr(self.stream)),
               "export")
        return keys[: int(v.edges)], ro

    def row_offsets(self):
        """Standard CSR row offsets alone (export_compacted without the keys)."""
        torch = _torch()
        ro = torch.empty(self.num_vertices + 1, dtype=torch.int32, device=self.device)
        _check(lib().lpma_export_compacted(self._h, None, ro.data_ptr(), _stream_ptr(self.stream)), "row_offsets")
        return ro

    def get_data(self):
        torch = _torch()
        v = self.view()
        keys = torch.empty(int(v.live_keys), dtype=torch.int64, device=self.device)
        _check(lib().lpma_export_data(self._h, keys.data_ptr(), _stream_ptr(self.stream)), "get_data")
        return keys

    def to_csr(self):
        torch = _torch()
        v = self.view()
        ro = torch.empty(self.num_vertices + 1, dtype=torch.int32, device=self.device)
        col = torch.empty(int(v.live_keys), dtype=torch.int32, device=self.device)
        _check(lib().lpma_to_csr(self._h, ro.data_ptr(), col.data_ptr(), _stream_ptr(self.stream)), "to_csr")
        return ro, col

    def edge_query(self, keys):
        torch = _torch()
        k = self._dev_keys(keys)
        res = torch.empty(max(k.numel(), 1), dtype=torch.int32, device=self.device)
        _check(lib().lpma_edge_query(self._h, k.data_ptr(), k.numel(), res.data_ptr(), _stream_ptr(self.stream)),
               "edge_query")
        return res[: k.numel()]

    def neighbors(self, srcs):
        """(offsets[n+1], neighbor dst ids) for a device int32 tensor of vertex ids."""
        torch = _torch()
        srcs = srcs.to(self.device, torch.int32).contiguous()
        n = srcs.numel()
        ln = torch.empty(max(n, 1), dtype=torch.int32, device=self.device)
        _check(lib().lpma_neighbor_count(self._h, srcs.data_ptr(), n, ln.data_ptr(), _stream_ptr(self.stream)),
               "neighbor_count")
        offs = torch.zeros(n + 1, dtype=torch.int64, device=self.device)
        if n:
            offs[1:] = torch.cumsum(ln[:n].to(torch.int64), 0)
        total = int(offs[-1].item())
        offs32 = offs.to(torch.int32)
        out = torch.empty(max(total, 1), dtype=torch.int32, device=self.device)
        _check(lib().lpma_neighbor_gather(self._h, srcs.data_ptr(), n, offs32.data_ptr(), out.data_ptr(),
                                          _stream_ptr(self.stream)), "neighbor_gather")
        return offs32, out[:total]

    def bfs(self, src):
        """(levels int32 device tensor, depth, live edges scanned); unreached = -1."""
        torch = _torch()
        lv = torch.empty(self.num_vertices, dtype=torch.int32, device=self.device)
        depth = _u32(0)
        scanned = _u64(0)
        _check(lib().lpma_bfs(self._h, int(src), lv.data_ptr(), ctypes.byref(depth), ctypes.byref(scanned),
                              _stream_ptr(self.stream)), "bfs")
        return lv, int(depth.value), int(scanned.value)

    def pagerank(self, damping=0.85, tolerance=1e-6, max_iter=50):
        """(rank float32 device tensor, iterations, last L1 change)."""
        torch = _torch()
        pr = torch.empty(self.num_vertices, dtype=torch.float32, device=self.device)
        it = _u32(0)
        dl = ctypes.c_float(0)
        _check(lib().lpma_pagerank(self._h, float(damping), float(tolerance), int(max_iter), pr.data_ptr(),
                                   ctypes.byref(it), ctypes.byref(dl), _stream_ptr(self.stream)), "pagerank")
        return pr, int(it.value), float(dl.value)

    def check(self):
        v = _u64(0)
        _check(lib().lpma_check(self._h, ctypes.byref(v), _stream_ptr(self.stream)), "check")
        return int(v.value)

    def stats(self):
        s = Stats()
        _check(lib().lpma_get_stats(self._h, ctypes.byref(s)), "stats")
        return s

    def reset_stats(self):
        _check(lib().lpma_reset_stats(self._h), "reset_stats")

    def set_profiling(self, on=True):
        _check(lib().lpma_set_profiling(self._h, 1 if on else 0), "set_profiling")

    def profile(self):
        p = Profile()
        _check(lib().lpma_get_profile(self._h, ctypes.byref(p)), "profile")
        return p

    def set_graph(self, on=True):
        _check(lib().lpma_set_graph(self._h, 1 if on else 0), "set_graph")

    def reserve(self, max_batch):
        _check(lib().lpma_reserve(self._h, int(max_batch), _stream_ptr(self.stream)), "reserve")

    def reserve_capacity(self, max_edges):
        """Allocate now what resizes up to max_edges live edges will need
        (cached in the device pool), so those resize batches make no device
        allocation."""
        _check(lib().lpma_reserve_capacity(self._h, int(max_edges), _stream_ptr(self.stream)), "reserve_capacity")

    def set_launch(self, seg_slots_per_lane=0):
        _check(lib().lpma_set_launch(self._h, int(seg_slots_per_lane)), "set_launch")


class Snapshot:
    """Read-only copy of a PMA's gapped view at one version (reads on it see
    that version while the PMA keeps changing)."""

    def __init__(self, pma):
        h = ctypes.c_void_p()
        _check(lib().lpma_snapshot_create(ctypes.byref(h), pma._h, _stream_ptr(pma.stream)), "snapshot_create")
        self._h = h
        self.device = pma.device
        self.stream = pma.stream
        self.num_vertices = pma.num_vertices

    def close(self):
        if getattr(self, "_h", None):
            lib().lpma_snapshot_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def view(self):
        v = View()
        _check(lib().lpma_snapshot_get_view(self._h, ctypes.byref(v)), "snapshot_view")
        return v

    def neighbors(self, srcs):
        torch = _torch()
        srcs = srcs.to(self.device, torch.int32).contiguous()
        n = srcs.numel()
        ln = torch.empty(max(n, 1), dtype=torch.int32, device=self.device)
        _check(lib().lpma_snapshot_neighbor_count(self._h, srcs.data_ptr(), n, ln.data_ptr(),
                                                  _stream_ptr(self.stream)), "snapshot_neighbor_count")
        offs = torch.zeros(n + 1, dtype=torch.int64, device=self.device)
        if n:
            offs[1:] = torch.cumsum(ln[:n].to(torch.int64), 0)
        total = int(offs[-1].item())
        offs32 = offs.to(torch.int32)
        out = torch.empty(max(total, 1), dtype=torch.int32, device=self.device)
        _check(lib().lpma_snapshot_neighbor_gather(self._h, srcs.data_ptr(), n, offs32.data_ptr(), out.data_ptr(),
                                                   _stream_ptr(self.stream)), "snapshot_neighbor_gather")
        return offs32, out[:total]

    def bfs(self, src):
        torch = _torch()
        lv = torch.empty(self.num_vertices, dtype=torch.int32, device=self.device)
        depth = _u32(0)
        scanned = _u64(0)
        _check(lib().lpma_snapshot_bfs(self._h, int(src), lv.data_ptr(), ctypes.byref(depth), ctypes.byref(scanned),
                                       _stream_ptr(self.stream)), "snapshot_bfs")
        return lv, int(depth.value), int(scanned.value)


class Group:
    """Independent PMAs updated by one pipeline run per round (the batched
    form; the reference has no counterpart).  Each member ends exactly as if
    it had been updated alone with ``update_mixed``."""

    def __init__(self, members, stream=None):
        self.members = list(members)
        n = len(self.members)
        hs = (ctypes.c_void_p * n)(*[m._h for m in self.members])
        h = ctypes.c_void_p()
        _check(lib().lpma_group_create(ctypes.byref(h), hs, n), "group_create")
        self._h = h
        self.stream = stream

    def close(self):
        if getattr(self, "_h", None):
            lib().lpma_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def partitions(self):
        return int(lib().lpma_group_partitions(self._h))

    def update(self, ins, dels):
        """One round: ins[i] / dels[i] (int64 device tensors, numpy arrays or
        None) for member i."""
        n = len(self.members)
        keep = []
        pi = (ctypes.c_void_p * n)()
        ni = (_u64 * n)()
        pd = (ctypes.c_void_p * n)()
        nd = (_u64 * n)()
        for i, m in enumerate(self.members):
            for arr, p_, n_ in ((ins[i] if ins is not None else None, pi, ni),
                                (dels[i] if dels is not None else None, pd, nd)):
                if arr is None or len(arr) == 0:
                    continue
                t = m._dev_keys(arr)
                keep.append(t)
                p_[i] = t.data_ptr()
                n_[i] = t.numel()
        _check(lib().lpma_group_update(self._h, pi, ni, pd, nd, _stream_ptr(self.stream)), "group_update")

    def stats(self):
        s = Stats()
        _check(lib().lpma_group_get_stats(self._h, ctypes.byref(s)), "group_stats")
        return s

    def set_profiling(self, on=True):
        _check(lib().lpma_group_set_profiling(self._h, 1 if on else 0), "group_set_profiling")

    def set_graph(self, on=True):
        _check(lib().lpma_group_set_graph(self._h, 1 if on else 0), "group_set_graph")

    def reserve(self, max_batch):
        _check(lib().lpma_group_reserve(self._h, int(max_batch), _stream_ptr(self.stream)), "group_reserve")

    def set_launch(self, seg_slots_per_lane=0):
        _check(lib().lpma_group_set_launch(self._h, int(seg_slots_per_lane)), "group_set_launch")

    def profile(self):
        p = Profile()
        _check(lib().lpma_group_get_profile(self._h, ctypes.byref(p)), "group_profile")
        return p


class Server:
    """The served form over one PMA: submit update batches, neighbor reads
    and BFS requests from any thread; results carry the number of update
    requests applied when they were served."""

    def __init__(self, pma, bfs_workers=2, hbm_budget_bytes=0, max_merge_keys=1 << 24, max_merge_requests=4096,
                 max_merge_vertices=1 << 20, snapshot_max_age_us=50000, search_workgroups=256):
        cfg = ServerConfig()
        lib().lpma_server_config_default(ctypes.byref(cfg))
        cfg.bfs_workers = bfs_workers
        cfg.hbm_budget_bytes = hbm_budget_bytes
        cfg.max_merge_keys = max_merge_keys
        cfg.max_merge_requests = max_merge_requests
        cfg.max_merge_vertices = max_merge_vertices
        cfg.snapshot_max_age_us = snapshot_max_age_us
        cfg.search_workgroups = search_workgroups
        h = ctypes.c_void_p()
        _check(lib().lpma_server_create(ctypes.byref(h), pma._h, ctypes.byref(cfg)), "server_create")
        self._h = h
        self.pma = pma

    def close(self):
        if getattr(self, "_h", None):
            lib().lpma_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit_update(self, ins=None, dels=None):
        a = np.ascontiguousarray(ins if ins is not None else np.empty(0), dtype=np.uint64)
        d = np.ascontiguousarray(dels if dels is not None else np.empty(0), dtype=np.uint64)
        t = _u64(0)
        _check(lib().lpma_server_submit_update(self._h, a.ctypes.data if a.size else None, a.size,
                                               d.ctypes.data if d.size else None, d.size, ctypes.byref(t)),
               "submit_update")
        return t.value

    def submit_read(self, vertices):
        v = np.ascontiguousarray(vertices, dtype=np.uint32)
        t = _u64(0)
        _check(lib().lpma_server_submit_read(self._h, v.ctypes.data if v.size else None, v.size, ctypes.byref(t)),
               "submit_read")
        return t.value

    def submit_bfs(self, root, levels=False):
        t = _u64(0)
        _check(lib().lpma_server_submit_bfs(self._h, int(root), 1 if levels else 0, ctypes.byref(t)), "submit_bfs")
        return t.value

    def wait(self, ticket, release=True):
        r = ServerResult()
        _check(lib().lpma_server_wait(self._h, int(ticket), ctypes.byref(r)), "server_wait")
        out = {"status": r.status, "kind": r.kind, "applied": int(r.applied_updates), "version": int(r.version),
               "queue_us": r.queue_us, "latency_us": r.latency_us, "depth": int(r.depth),
               "edges_scanned": int(r.edges_scanned), "offsets": None, "values": None}
        if r.kind == REQ_READ and r.status == 0:
            k = int(r.count)
            offs = np.ctypeslib.as_array(r.offsets, shape=(k + 1,)).copy() if r.offsets else np.zeros(k + 1, np.uint32)
            n = int(offs[-1]) if k + 1 else 0
            out["offsets"] = offs
            out["values"] = np.ctypeslib.as_array(r.values, shape=(n,)).copy() if n else np.empty(0, np.uint32)
        elif r.kind == REQ_BFS and r.status == 0 and r.count:
            out["values"] = np.ctypeslib.as_array(r.values, shape=(int(r.count),)).copy()
        if release:
            lib().lpma_server_release(self._h, int(ticket))
        return out

    def stats(self):
        s = ServerStats()
        _check(lib().lpma_server_get_stats(self._h, ctypes.byref(s)), "server_stats")
        return {f: (list(getattr(s, f)) if f in ("submitted", "completed") else
                    (float(getattr(s, f)) if f.endswith("_us") else int(getattr(s, f))))
                for f, _ in ServerStats._fields_}


# ------------------------------------------- reference-named free functions

def init_csr_rpma(num_vertices, **kw):
    return PMA(num_vertices, **kw)


def update_rpma(pma, keys, op=OP_INSERT):
    pma.update(keys, op)


def get_data_rpma(pma):
    return pma.get_data()


def to_csr_rpma(pma):
    return pma.to_csr()


def edge_query_rpma(pma, keys):
    return pma.edge_query(keys)


def query_rpma(pma, srcs):
    return pma.neighbors(srcs)


def destroy_rpma(pma):
    pma.close()
