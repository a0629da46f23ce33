// lpma.cpp -- host side of the C ABI: handle, device pool allocator, and
// the update orchestrator that replaces update_rpma (rpma.cuh:1549-1718).
//
// One update call = one stream-ordered kernel sequence with exactly one
// host synchronisation (after the density check, to learn whether the batch
// is a plain rebalance, a root rebalance or a resize), instead of the
// reference's per-level host loop with a device sync after every step.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <mutex>
#include <new>
#include <unordered_map>
#include <vector>

#include "../../include/lpma.h"
#include "common.hpp"
#include "kernels.hpp"

using namespace pma;

// ------------------------------------------------------------ device pool
//
// Size-classed cache of device blocks per device (regrowth reuses freed
// level arrays and scratch instead of hipMalloc/hipFree per batch, the
// reference's alloc_or_del_levels, rpma.cuh:1396-1415).  Blocks up to 64 MiB
// round to a power of two, larger ones to 16 MiB multiples.  A byte budget
// caps what the pool holds from hipMalloc.

namespace {

struct Pool {
    std::mutex mu;
    std::map<size_t, std::vector<void *>> free_blocks;
    std::unordered_map<void *, size_t> live;
    size_t reserved = 0;
    size_t cached = 0;

    static size_t round(size_t b) {
        if (b < 256) b = 256;
        if (b <= (64ull << 20)) {
            size_t r = 256;
            while (r < b) r <<= 1;
            return r;
        }
        size_t q = 16ull << 20;
        return (b + q - 1) / q * q;
    }

    void trim_locked() {
        for (auto &kv : free_blocks)
            for (void *p : kv.second) {
                (void)hipFree(p);
                reserved -= kv.first;
                cached -= kv.first;
            }
        free_blocks.clear();
    }
};

std::mutex g_pools_mu;
std::map<int, Pool *> g_pools;
size_t g_budget = 0;  // 0 = unlimited

// Fault injection for the failure-path tests: when positive, a countdown of
// allocations (pool blocks, cached or new, and the pinned host buffers of a
// run all count); the one that brings it to zero fails and so does every
// later one (-1) until the injection is reset, as when memory runs out.
std::atomic<long long> g_fail_countdown{0};

bool injected_failure() {
    long long c = g_fail_countdown.load(std::memory_order_relaxed);
    while (c > 0) {
        if (g_fail_countdown.compare_exchange_weak(c, c == 1 ? -1 : c - 1, std::memory_order_relaxed))
            return c == 1;
    }
    return c < 0;
}

Pool &pool_for(int dev) {
    std::lock_guard<std::mutex> g(g_pools_mu);
    auto it = g_pools.find(dev);
    if (it != g_pools.end()) return *it->second;
    Pool *p = new Pool();
    g_pools[dev] = p;
    return *p;
}

int pool_alloc(int dev, void **out, size_t bytes) {
    if (injected_failure()) return LPMA_ERR_OUT_OF_MEMORY;
    Pool &P = pool_for(dev);
    std::lock_guard<std::mutex> g(P.mu);
    size_t r = Pool::round(bytes);
    auto it = P.free_blocks.find(r);
    if (it != P.free_blocks.end() && !it->second.empty()) {
        *out = it->second.back();
        it->second.pop_back();
        P.cached -= r;
        P.live[*out] = r;
        return LPMA_OK;
    }
    if (g_budget && P.reserved + r > g_budget) P.trim_locked();
    if (g_budget && P.reserved + r > g_budget) return LPMA_ERR_OUT_OF_MEMORY;
    void *p = nullptr;
    if (hipMalloc(&p, r) != hipSuccess) {
        (void)hipGetLastError();
        P.trim_locked();
        if (hipMalloc(&p, r) != hipSuccess) {
            (void)hipGetLastError();
            return LPMA_ERR_OUT_OF_MEMORY;
        }
    }
    P.reserved += r;
    P.live[p] = r;
    *out = p;
    return LPMA_OK;
}

void pool_free(int dev, void *p) {
    if (!p) return;
    Pool &P = pool_for(dev);
    std::lock_guard<std::mutex> g(P.mu);
    auto it = P.live.find(p);
    if (it == P.live.end()) return;
    size_t r = it->second;
    P.live.erase(it);
    P.free_blocks[r].push_back(p);
    P.cached += r;
}

// growable device buffer from the pool
template <typename T>
struct DBuf {
    T *p = nullptr;
    u64 n = 0;  // capacity in elements
};

}  // namespace

// ---------------------------------------------------------------- pipeline
//
// Batch scratch, run status and timing of one update pipeline.  A handle owns
// one for its own updates (a group of one PMA); a group owns one per
// partition of its members.

enum Phase { PH_SORT, PH_LOCATE, PH_ELECT, PH_WAVE, PH_BLOCK, PH_LARGE, PH_REBUILD, PH_PYR, PH_SEG, PH_GRAPH, PH_NUM };

struct Pipeline {
    int device = 0;
    hipStream_t cur_stream = nullptr;  // stream of the call in progress

    DBuf<u64> pk0, pk1, ukey, useg, ckey, cseg, merged, lmbase;
    DBuf<u32> sort_scr, rlevel, gpre, scan_scr, lchunk, lhdr, lvmask, wfirst;
    DBuf<uint4> rwin;
    DBuf<i32> pd, part;
    DBuf<Work> work[NUM_CLASSES];
    DBuf<PmaEntry> tab;
    DBuf<BatchSeg> segs;
    DBuf<u64> out;  // Status, then one PmaStat per PMA

    // pinned host mirrors
    PmaEntry *h_tab = nullptr;
    u32 h_tab_n = 0;
    std::vector<PmaEntry> tab_shadow;  // what the device table holds
    const PmaEntry *tab_shadow_at = nullptr;
    BatchSeg *h_segs = nullptr;
    u32 h_segs_n = 0;
    unsigned char *h_out = nullptr;
    u64 h_out_bytes = 0;

    u64 *d_stats = nullptr;
    u64 rounds = 0, effective = 0, rebuilds = 0, rebuild_slots = 0;

    int seg_slots_log2 = -1;  // slots per lane of the single-segment rebalance (-1: default)

    // graph replays: one captured graph per round shape
    bool graph_on = false;
    hipStream_t cap_stream = nullptr;
    struct GraphEntry {
        hipGraphExec_t exec;
        std::vector<u64> key;  // everything the captured graph has baked in
    };
    std::map<u64, GraphEntry> graphs;  // by batch capacity
    DBuf<RoundHdr> rhdr;
    RoundHdr *h_rhdr = nullptr;
    u64 graph_captures = 0, graph_launches = 0;

    bool profiling = false;
    struct Ev {
        hipEvent_t a, b;
        int phase;
    };
    std::vector<Ev> events;
    std::vector<hipEvent_t> ev_pool;
    double phase_ms[PH_NUM] = {0};
    u64 launches_wave = 0, launches_block = 0, launches_seg = 0;

    hipEvent_t ev_status = nullptr;  // the round's status has reached the host
};

static constexpr u64 STATUS_BYTES = (sizeof(Status) + 15) / 16 * 16;

// ------------------------------------------------------------------ handle

// the arrays of one height
struct Geometry {
    DBuf<u64> keys, sep, owner;
    DBuf<u32> cnt, pyr;
};

struct lpma {
    int device = 0;
    hipStream_t cur_stream = nullptr;  // stream of the call in progress
    lpma_config_t cfg;
    u32 V = 0, vbits = 1, log2S = 5;
    u64 live = 0;
    u64 version = 0;
    // the count pyramid's chunk levels are owed: the last batch changed
    // segment counts and left them to the next batch's first kernels (a
    // lone PMA's early path); any other path sums them before its batch
    bool pyr_owed = false;
    u32 last_epoch = 0;  // epoch of the last run that updated this PMA
    DevState st;
    Density dn;

    DBuf<u64> keys;
    DBuf<u32> cnt, pyr, row_off;
    DBuf<u64> owner;
    DBuf<u64> sep;
    // arrays held for rebuilds, by height (lpma_reserve_capacity): a
    // rebuild takes its new arrays from here and, while keep_spares is set,
    // leaves its old ones here
    std::map<u32, std::vector<Geometry>> spares;
    bool keep_spares = false;

    Pipeline pl;  // plain updates, views and analytics scratch
    // analytics scratch
    DBuf<u32> aq0, apoff, apcb;
    DBuf<float> af0, af1, af2, ared, apart;
    DBuf<u64> apairs;
    u32 *d_err = nullptr;

    // host counters
    u64 calls = 0, effective = 0, rebuilds = 0, rebuild_slots = 0;
};

// A group of PMAs updated together: one pipeline run per partition (members
// with the same segment size whose ids fit above their sort keys).
struct lpma_group {
    int device = 0;
    std::vector<lpma *> members;
    std::vector<std::vector<u32>> parts;
    std::vector<Pipeline *> pls;
    u64 calls = 0;
};

// A read-only copy of a PMA's gapped view (keys, counts, separators, row
// offsets) at one version: readers use it while writers go on.
struct lpma_snapshot {
    int device = 0;
    DevState st;
    DBuf<u64> keys, sep;
    DBuf<u32> cnt, row_off;
    u64 version = 0, live = 0;
    u32 V = 0, log2S = 5;
    hipEvent_t ready = nullptr;  // the copies are done
    u32 search_grid = 0;         // workgroups of a search's expansion (0: the full width)
};

namespace {

std::atomic<u32> g_epoch{0};

struct DevGuard {
    int prev = -1;
    explicit DevGuard(int dev) {
        (void)hipGetDevice(&prev);
        if (prev != dev) (void)hipSetDevice(dev);
    }
    ~DevGuard() {
        int cur;
        (void)hipGetDevice(&cur);
        if (prev >= 0 && cur != prev) (void)hipSetDevice(prev);
    }
};

template <typename T>
int ensure_on(int device, hipStream_t sync_st, DBuf<T> &b, u64 n) {
    if (b.n >= n && b.p) return LPMA_OK;
    u64 want = std::max<u64>(n, b.n + b.n / 2);
    if (want < 64) want = 64;
    void *p = nullptr;
    int r = pool_alloc(device, &p, want * sizeof(T));
    if (r != LPMA_OK) {
        // retry exactly
        r = pool_alloc(device, &p, n * sizeof(T));
        if (r != LPMA_OK) return r;
        want = n;
    }
    if (b.p) {
        // queued kernels may still read the old block; it must not reach
        // another stream through the shared pool before they finish
        (void)hipStreamSynchronize(sync_st);
        pool_free(device, b.p);
    }
    b.p = (T *)p;
    b.n = want;
    return LPMA_OK;
}

template <typename T>
int ensure(lpma *h, DBuf<T> &b, u64 n) { return ensure_on(h->device, h->cur_stream, b, n); }
template <typename T>
int ensure(Pipeline &pl, DBuf<T> &b, u64 n) { return ensure_on(pl.device, pl.cur_stream, b, n); }

template <typename T>
void release_on(int device, DBuf<T> &b) {
    pool_free(device, b.p);
    b.p = nullptr;
    b.n = 0;
}
template <typename T>
void release(lpma *h, DBuf<T> &b) { release_on(h->device, b); }

// pinned host buffer of at least n elements (contents not kept)
template <typename T>
int ensure_host(T *&p, u32 &cap, u64 n, hipStream_t sync_st) {
    if (p && cap >= n) return LPMA_OK;
    if (injected_failure()) return LPMA_ERR_OUT_OF_MEMORY;
    u64 want = std::max<u64>(n, 2ull * cap);
    if (want < 8) want = 8;
    if (p) {
        (void)hipStreamSynchronize(sync_st);
        (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
    if (hipHostMalloc((void **)&p, want * sizeof(T), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
        return LPMA_ERR_OUT_OF_MEMORY;
    }
    cap = (u32)want;
    return LPMA_OK;
}

int pipeline_init(Pipeline &pl, int device, hipStream_t st) {
    pl.device = device;
    pl.cur_stream = st;
    if (hipMalloc(&pl.d_stats, ST_ROWS * ST_NUM * sizeof(u64)) != hipSuccess) {
        (void)hipGetLastError();
        pl.d_stats = nullptr;
        return LPMA_ERR_OUT_OF_MEMORY;
    }
    if (fill_u64(pl.d_stats, (u64)ST_ROWS * ST_NUM, 0, st) != hipSuccess) {
        (void)hipGetLastError();
        return LPMA_ERR_DEVICE;
    }
    return LPMA_OK;
}

void collect_events(Pipeline &pl) {
    for (auto &e : pl.events) {
        (void)hipEventSynchronize(e.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e.a, e.b);
        pl.phase_ms[e.phase] += ms;
        pl.ev_pool.push_back(e.a);
        pl.ev_pool.push_back(e.b);
    }
    pl.events.clear();
}

void pipeline_free(Pipeline &pl) {
    collect_events(pl);
    for (auto e : pl.ev_pool) (void)hipEventDestroy(e);
    pl.ev_pool.clear();
    if (pl.ev_status) (void)hipEventDestroy(pl.ev_status);
    pl.ev_status = nullptr;
    int d = pl.device;
    release_on(d, pl.pk0); release_on(d, pl.pk1); release_on(d, pl.ukey); release_on(d, pl.useg);
    release_on(d, pl.ckey); release_on(d, pl.cseg); release_on(d, pl.merged); release_on(d, pl.lmbase);
    release_on(d, pl.sort_scr); release_on(d, pl.rlevel); release_on(d, pl.rwin); release_on(d, pl.gpre); release_on(d, pl.scan_scr);
    release_on(d, pl.lchunk); release_on(d, pl.lhdr); release_on(d, pl.pd); release_on(d, pl.part);
    release_on(d, pl.lvmask); release_on(d, pl.wfirst);
    for (int c = 0; c < NUM_CLASSES; ++c) release_on(d, pl.work[c]);
    release_on(d, pl.tab); release_on(d, pl.segs); release_on(d, pl.out); release_on(d, pl.rhdr);
    for (auto &kv : pl.graphs) (void)hipGraphExecDestroy(kv.second.exec);
    pl.graphs.clear();
    if (pl.cap_stream) (void)hipStreamDestroy(pl.cap_stream);
    pl.cap_stream = nullptr;
    if (pl.h_rhdr) (void)hipHostFree(pl.h_rhdr);
    pl.h_rhdr = nullptr;
    if (pl.h_tab) (void)hipHostFree(pl.h_tab);
    if (pl.h_segs) (void)hipHostFree(pl.h_segs);
    if (pl.h_out) (void)hipHostFree(pl.h_out);
    pl.h_tab = nullptr; pl.h_segs = nullptr; pl.h_out = nullptr;
    if (pl.d_stats) (void)hipFree(pl.d_stats);
    pl.d_stats = nullptr;
}

hipEvent_t take_event(Pipeline &pl) {
    if (!pl.ev_pool.empty()) {
        hipEvent_t e = pl.ev_pool.back();
        pl.ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

struct PhaseTimer {
    Pipeline &pl;
    hipStream_t st;
    int phase;
    hipEvent_t a = nullptr;
    PhaseTimer(Pipeline &pl_, hipStream_t s, int p) : pl(pl_), st(s), phase(p) {
        if (pl.profiling) {
            a = take_event(pl);
            (void)hipEventRecord(a, st);
        }
    }
    ~PhaseTimer() {
        if (a) {
            hipEvent_t b = take_event(pl);
            (void)hipEventRecord(b, st);
            pl.events.push_back({a, b, phase});
        }
    }
};

// Events that a launch records itself (hipExtLaunchKernelGGL): the phase
// time is then the kernel's execution, without the ~1.7 us that a pair of
// separate event records adds around it.  Null events when not profiling.
struct LaunchTimer {
    Pipeline &pl;
    int phase;
    hipEvent_t a = nullptr, b = nullptr;
    LaunchTimer(Pipeline &pl_, int p) : pl(pl_), phase(p) {
        if (pl.profiling) {
            a = take_event(pl);
            b = take_event(pl);
        }
    }
    ~LaunchTimer() {
        if (a) pl.events.push_back({a, b, phase});
    }
};

// Density table, rpma.cuh:136-168 (see lpma_config_t)
void compute_density(lpma *h) {
    const lpma_config_t &c = h->cfg;
    const u32 H = h->st.H;
    for (u32 l = 0; l < 32; ++l) {
        double len = (double)((u64)c.segment_slots << l);
        double t = (H == 0 || l >= H) ? 0.0 : (double)(H - l) / (double)H;
        double lo = c.density_lower_root + (c.density_lower_leaf - c.density_lower_root) * t;
        double up = c.density_upper_root + (c.density_upper_leaf - c.density_upper_root) * t;
        h->dn.lower[l] = (u64)(lo * len);
        h->dn.upper[l] = (u64)(up * len);
    }
    h->dn.root_shrink_below = H > 0 ? (u64)(c.density_upper_root * (double)((u64)c.segment_slots << (H - 1))) : 0;
}

// smallest height whose root window holds m keys (resize_rpma, rpma.cuh:1341-1344)
u32 fit_height(const lpma *h, u64 m) {
    u32 H = 0;
    while ((u64)(h->cfg.density_upper_root * (double)((u64)h->cfg.segment_slots << H)) < m) {
        ++H;
        if (H > 40) break;
    }
    return H;
}

bool height_ok(const lpma *h, u32 H) { return H <= 31 && ((u64)h->cfg.segment_slots << H) <= (1ull << 32); }

// arrays for height H; the caller spreads keys into every slot and segment
int alloc_geometry(lpma *h, u32 H, Geometry &g, hipStream_t st) {
    u64 nseg = 1ull << H;
    u64 cap = nseg << h->log2S;
    int r;
    auto spare = h->spares.find(H);
    if (spare != h->spares.end() && !spare->second.empty()) {
        g = spare->second.back();
        spare->second.pop_back();
    } else if ((r = ensure(h, g.keys, cap)) || (r = ensure(h, g.cnt, nseg)) || (r = ensure(h, g.pyr, nseg)) ||
               (r = ensure(h, g.sep, nseg)) || (r = ensure(h, g.owner, 2 * nseg))) {
        release(h, g.keys); release(h, g.cnt); release(h, g.pyr); release(h, g.sep); release(h, g.owner);
        return r;
    }
    // keys, counts, separators and the pyramid are written whole by the
    // spread that follows every allocation (k_pull, k_meta, the pyramid
    // rebuild); the window claims must start clear
    if (fill_u64(g.owner.p, 2 * nseg, 0, st) != hipSuccess) {
        (void)hipGetLastError();
        release(h, g.keys); release(h, g.cnt); release(h, g.pyr); release(h, g.sep); release(h, g.owner);
        return LPMA_ERR_DEVICE;
    }
    return LPMA_OK;
}

void release_geometry(lpma *h, Geometry &g) {
    release(h, g.keys); release(h, g.cnt); release(h, g.pyr); release(h, g.sep); release(h, g.owner);
}

// installs g as h's arrays (g is left empty)
void install_geometry(lpma *h, u32 H, Geometry &g) {
    h->pyr_owed = false;  // new arrays come with their pyramid
    if (h->keep_spares && h->keys.p) {
        Geometry old;
        old.keys = h->keys; old.cnt = h->cnt; old.pyr = h->pyr; old.sep = h->sep; old.owner = h->owner;
        h->spares[h->st.H].push_back(old);
    } else {
        release(h, h->keys); release(h, h->cnt); release(h, h->pyr); release(h, h->sep); release(h, h->owner);
    }
    h->keys = g.keys; h->cnt = g.cnt; h->pyr = g.pyr; h->sep = g.sep; h->owner = g.owner;
    g = Geometry();
    h->st.keys = h->keys.p;
    h->st.cnt = h->cnt.p;
    h->st.pyr = h->pyr.p;
    h->st.sep = h->sep.p;
    h->st.owner = h->owner.p;
    h->st.H = H;
    compute_density(h);
}

DevState state_with(const lpma *h, u32 H, const Geometry &g) {
    DevState d = h->st;
    d.keys = g.keys.p; d.cnt = g.cnt.p; d.pyr = g.pyr.p; d.sep = g.sep.p; d.owner = g.owner.p;
    d.H = H;
    return d;
}

Status *status_of(Pipeline &pl) { return (Status *)pl.out.p; }
PmaStat *pstat_of(Pipeline &pl) { return (PmaStat *)((unsigned char *)pl.out.p + STATUS_BYTES); }

BatchBufs batch_bufs(Pipeline &pl) {
    BatchBufs b;
    b.ukey = pl.ukey.p;
    b.useg = pl.useg.p;
    b.ckey = pl.ckey.p;
    b.cseg = pl.cseg.p;
    b.pd = pl.pd.p;
    b.rlevel = pl.rlevel.p;
    b.rwin = pl.rwin.p;
    b.part = pl.part.p;
    for (int c = 0; c < NUM_CLASSES; ++c) b.work[c] = pl.work[c].p;
    b.lchunk = pl.lchunk.p;
    b.lmbase = pl.lmbase.p;
    b.lhdr = pl.lhdr.p;
    b.wfirst = pl.wfirst.p;
    b.status = status_of(pl);
    return b;
}

int ensure_batch(Pipeline &pl, u64 n, u32 npma, u32 nsegs, u64 runs) {
    int r;
    u64 nb = locate_blocks(n);
    u64 out_words = (STATUS_BYTES + (u64)npma * sizeof(PmaStat) + 7) / 8;
    if ((r = ensure(pl, pl.pk0, n)) || (r = ensure(pl, pl.pk1, n)) ||
        (r = ensure(pl, pl.sort_scr, sort_scratch_words(n))) || (r = ensure(pl, pl.ukey, n)) ||
        (r = ensure(pl, pl.useg, n)) || (r = ensure(pl, pl.ckey, n)) || (r = ensure(pl, pl.cseg, n)) ||
        (r = ensure(pl, pl.pd, n + 1)) || (r = ensure(pl, pl.rlevel, n)) || (r = ensure(pl, pl.rwin, n)) || (r = ensure(pl, pl.part, 2 * nb + 2)) ||
        (r = ensure(pl, pl.tab, npma)) || (r = ensure(pl, pl.segs, std::max<u32>(nsegs, INLINE_SEGS))) ||
        (r = ensure(pl, pl.lvmask, (u64)npma * LV_REPL)) || (r = ensure(pl, pl.wfirst, WFIRST_N + 1)) ||
        (r = ensure(pl, pl.out, out_words)))
        return r;
    for (int c = 0; c < NUM_CLASSES; ++c)
        if ((r = ensure(pl, pl.work[c], runs))) return r;
    if ((r = ensure_host(pl.h_tab, pl.h_tab_n, npma, pl.cur_stream)) ||
        (r = ensure_host(pl.h_segs, pl.h_segs_n, nsegs, pl.cur_stream)))
        return r;
    if (!pl.h_out || pl.h_out_bytes < out_words * 8) {
        if (injected_failure()) return LPMA_ERR_OUT_OF_MEMORY;
        if (pl.h_out) {
            (void)hipStreamSynchronize(pl.cur_stream);
            (void)hipHostFree(pl.h_out);
            pl.h_out = nullptr;
        }
        u64 want = std::max<u64>(out_words * 8, 2 * pl.h_out_bytes);
        if (hipHostMalloc((void **)&pl.h_out, want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            pl.h_out = nullptr;
            pl.h_out_bytes = 0;
            return LPMA_ERR_OUT_OF_MEMORY;
        }
        pl.h_out_bytes = want;
    }
    return LPMA_OK;
}

#define HIPCHECK(x)                                    \
    do {                                               \
        hipError_t e_ = (x);                           \
        if (e_ != hipSuccess) {                        \
            (void)hipGetLastError();                   \
            return LPMA_ERR_DEVICE;                    \
        }                                              \
    } while (0)

#define CHECKED(x)                  \
    do {                            \
        int r_ = (x);               \
        if (r_ != LPMA_OK) return r_; \
    } while (0)

// Spread m sorted keys from `merged` over the fresh arrays g of height H and
// install them (create and resize); g is released on failure.
// the count pyramid of new arrays g (keys and counts written), then g
// replaces h's arrays; g is released on failure
int pyramid_and_install(lpma *h, u32 H, Geometry &g, hipStream_t st) {
    // old arrays may still be read by queued kernels: wait before they
    // return to the pool
    if (rebuild_pyramid(state_with(h, H, g), true, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
        (void)hipGetLastError();
        release_geometry(h, g);
        return LPMA_ERR_DEVICE;
    }
    install_geometry(h, H, g);
    return LPMA_OK;
}

int spread_and_install(lpma *h, u32 H, Geometry &g, const u64 *merged, u64 m, hipStream_t st) {
    if (rebuild_spread(state_with(h, H, g), merged, m, st) != hipSuccess) {
        (void)hipGetLastError();
        release_geometry(h, g);
        return LPMA_ERR_DEVICE;
    }
    return pyramid_and_install(h, H, g, st);
}

// One PMA's share of a pipeline run.
// One request of a member: inserts and deletes (deletes first).
struct Req {
    const u64 *ins;
    u64 nins;
    const u64 *del;
    u64 ndel;
};

// One PMA's share of a pipeline run: its requests, applied in order.
struct Member {
    lpma *h;
    const Req *reqs;
    u32 nreq;
};

u32 bits_for(u64 n) {  // bits to hold the values 0 .. n-1
    u32 b = 0;
    while ((1ull << b) < n) ++b;
    return b;
}

u32 next_epoch() {
    u32 e = ++g_epoch;
    if (e == 0) e = ++g_epoch;
    return e;
}

// Epoch of a new round; clears a member's claims after the counter wrapped.
int begin_epoch(const Member *mem, u32 npma, u32 &epoch, hipStream_t st) {
    epoch = next_epoch();
    for (u32 p = 0; p < npma; ++p) {
        lpma *h = mem[p].h;
        if (epoch <= h->last_epoch) HIPCHECK(fill_u64(h->owner.p, 2 * (1ull << h->st.H), 0, st));
        h->last_epoch = epoch;
    }
    return LPMA_OK;
}

// Device table of the members (uploaded only when it changed); returns the
// workgroups of the group pyramid launch.
// the chunk levels a member's last batch left owing (the batch-begin kernel
// then sums the levels above them)
hipError_t pay_owed_pyramids(const Member *mem, u32 npma, hipStream_t st) {
    for (u32 p = 0; p < npma; ++p) {
        lpma *h = mem[p].h;
        if (!h->pyr_owed) continue;
        const hipError_t e = rebuild_pyramid(h->st, false, st);
        if (e != hipSuccess) return e;
        h->pyr_owed = false;
    }
    return hipSuccess;
}

int upload_table(Pipeline &pl, const Member *mem, u32 npma, u32 &pyr_blocks, hipStream_t st) {
    pyr_blocks = 0;
    for (u32 p = 0; p < npma; ++p) {
        PmaEntry &e = pl.h_tab[p];
        memset(&e, 0, sizeof(e));
        e.d = mem[p].h->st;
        e.dn = mem[p].h->dn;
        e.pyr_blk0 = pyr_blocks;
        e.pyr_nblk = pyramid_blocks(e.d.H);
        pyr_blocks += e.pyr_nblk;
    }
    if (pl.tab_shadow_at != pl.tab.p || pl.tab_shadow.size() != npma ||
        memcmp(pl.tab_shadow.data(), pl.h_tab, npma * sizeof(PmaEntry)) != 0) {
        HIPCHECK(hipMemcpyAsync(pl.tab.p, pl.h_tab, npma * sizeof(PmaEntry), hipMemcpyHostToDevice, st));
        pl.tab_shadow.assign(pl.h_tab, pl.h_tab + npma);
        pl.tab_shadow_at = pl.tab.p;
    }
    return LPMA_OK;
}

// After the round's status is on the host: the window classes still to be
// launched, the pyramid, root rebuilds / resizes of the flagged members, then
// the bookkeeping.  `windows_done` tells whether the window rebalances and
// the pyramid have already been queued (graph replays); `early` that the
// single-segment and wave classes were queued before the status was read
// (a lone PMA: they only touch a PMA without a root rebuild or resize, whose
// windows the election drops, and their scratch was reserved beforehand),
// and that the pyramid's chunk levels are left owing to the next batch.
//
// Every fallible allocation of the round comes before its first change of a
// PMA, so a round that fails with LPMA_ERR_OUT_OF_MEMORY changes nothing.  A
// graph replay over several PMAs elects no windows at all when one of them
// needs a rebuild (Group::defer_rebuild); the election then runs again here,
// after the allocations.
int finish_round(Pipeline &pl, const Member *mem, u32 npma, const Group &g_in, u32 pyr_blocks, bool windows_done,
                 hipStream_t st, bool early = false) {
    Status s = *(const Status *)pl.h_out;
    const PmaStat *ps = (const PmaStat *)(pl.h_out + STATUS_BYTES);
    if (s.flags & FLAG_INVALID_KEY) return LPMA_ERR_INVALID_KEY;
    if (s.flags & FLAG_SORT_STALL) return LPMA_ERR_DEVICE;
    const u64 n_eff = s.n_eff;
    if (n_eff == 0) {
        for (u32 p = 0; p < npma; ++p) mem[p].h->version++;
        return LPMA_OK;
    }

    // new live counts and heights; every root rebuild / resize is checked
    // before any kernel changes a PMA
    std::vector<u64> Mnew(npma);
    std::vector<u32> newH(npma);
    u64 m_sum = 0;
    bool any_rebuild = false;
    for (u32 p = 0; p < npma; ++p) {
        lpma *h = mem[p].h;
        Mnew[p] = (u64)((i64)h->live + (ps[p].pd1 - ps[p].pd0));
        m_sum += Mnew[p];
        newH[p] = h->st.H;
        if (ps[p].flags & (FLAG_ROOT | FLAG_RESIZE)) {
            any_rebuild = true;
            if (ps[p].flags & FLAG_RESIZE) newH[p] = fit_height(h, Mnew[p]);
            if (!height_ok(h, newH[p])) return LPMA_ERR_TOO_LARGE;
        }
    }
    const bool reelect = windows_done && any_rebuild && g_in.defer_rebuild;
    if (reelect) windows_done = false;
    Group g = g_in;
    g.defer_rebuild = 0;

    // allocations: scratch of the large windows (bounded by the runs when the
    // election is still to come), and the new arrays of every member that
    // is rebuilt (a resize, or a root rebalance at the same height)
    std::vector<Geometry> geo(npma);
    auto undo = [&](int r) {
        for (u32 p = 0; p < npma; ++p) release_geometry(mem[p].h, geo[p]);
        return r;
    };
    {
        int r = LPMA_OK;
        const u64 nlarge = reelect ? n_eff : (windows_done ? 0 : s.nwork[CLASS_LARGE]);
        if (nlarge) {
            if (!(r = ensure(pl, pl.merged, m_sum + 1)) && !(r = ensure(pl, pl.lchunk, nlarge + 1)) &&
                !(r = ensure(pl, pl.lmbase, nlarge)))
                r = ensure(pl, pl.lhdr, 4);
        }
        for (u32 p = 0; !r && any_rebuild && p < npma; ++p)
            if (ps[p].flags & (FLAG_ROOT | FLAG_RESIZE)) {
                r = ensure(pl, pl.gpre, rebuild_index_words(mem[p].h->st));
                if (!r) r = alloc_geometry(mem[p].h, newH[p], geo[p], st);
            }
        if (r) return undo(r);
    }
    BatchBufs bb = batch_bufs(pl);
#define HIPUNDO(x)                           \
    do {                                     \
        if ((x) != hipSuccess) {             \
            (void)hipGetLastError();         \
            return undo(LPMA_ERR_DEVICE);    \
        }                                    \
    } while (0)

    if (reelect) {
        // the graph's election stood down: clear its (empty) work counts and
        // elect again, now that nothing can fail for lack of memory
        HIPUNDO(reset_work_counts(bb.status, st));
        HIPUNDO(elect_windows(g, n_eff, bb, st));
        HIPUNDO(hipMemcpyAsync(pl.h_out, pl.out.p, sizeof(Status), hipMemcpyDeviceToHost, st));
        HIPUNDO(hipStreamSynchronize(st));
        s = *(const Status *)pl.h_out;
    }

    if (!windows_done) {
        if (s.nwork[CLASS_SEG] && !early) {
            LaunchTimer t(pl, PH_SEG);
            HIPUNDO(rebalance_seg(g, bb, s.nwork[CLASS_SEG], st, t.a, t.b));
            pl.launches_seg++;
        }
        if (s.nwork[CLASS_WAVE] && !early) {
            LaunchTimer t(pl, PH_WAVE);
            HIPUNDO(rebalance_lds(g, bb, CLASS_WAVE, std::min<u32>(s.nwork[CLASS_WAVE], 65535), st, t.a, t.b));
            pl.launches_wave++;
        }
        if (s.nwork[CLASS_BLOCK]) {
            LaunchTimer t(pl, PH_BLOCK);
            HIPUNDO(rebalance_lds(g, bb, CLASS_BLOCK, std::min<u32>(s.nwork[CLASS_BLOCK], 65535), st, t.a, t.b));
            pl.launches_block++;
        }
        if (s.nwork[CLASS_LARGE]) {
            PhaseTimer t(pl, st, PH_LARGE);
            HIPUNDO(rebalance_large(g, bb, pl.merged.p, 1024, st));
        }
        if (early) {
            // a lone PMA's next batch sums the chunk levels in its sort
            // kernels (or pays them first on another path)
            mem[0].h->pyr_owed = true;
        } else {
            PhaseTimer t(pl, st, PH_PYR);
            HIPUNDO(group_pyramid(g, pyr_blocks, st));
        }
    }

    if (any_rebuild) {
        // root rebalance or resize: every live key and every update of the
        // PMA go straight to their slots in the new arrays, which then
        // replace the old ones (rebuild_rpma, rpma.cuh:1417-1472)
        PhaseTimer t(pl, st, PH_REBUILD);
        for (u32 p = 0; p < npma; ++p) {
            if (!(ps[p].flags & (FLAG_ROOT | FLAG_RESIZE))) continue;
            lpma *h = mem[p].h;
            HIPUNDO(rebuild_scatter(h->st, state_with(h, newH[p], geo[p]), p, ps[p].u0, ps[p].u1, bb, Mnew[p],
                                    pl.gpre.p, st));
            const int r = pyramid_and_install(h, newH[p], geo[p], st);
            if (r) return undo(r);
            h->rebuilds++;
            h->rebuild_slots += (u64)1 << (h->st.H + h->log2S);
            pl.rebuilds++;
            pl.rebuild_slots += (u64)1 << (h->st.H + h->log2S);
        }
    }
#undef HIPUNDO
    for (u32 p = 0; p < npma; ++p) {
        lpma *h = mem[p].h;
        h->live = Mnew[p];
        h->effective += ps[p].u1 - ps[p].u0;
        h->version++;
    }
    pl.effective += n_eff;
    return LPMA_OK;
}

int run_pipeline_graph(Pipeline &pl, const Member *mem, u32 npma, u64 n, u32 pshift, u32 hi_bit, hipStream_t st);

// One update round over the members (all on pl's device, same segment size,
// ids fitting above the sort keys): sort + dedup + locate + density check +
// election over every member at once, one host synchronisation, then the
// rebalance classes, the pyramid, and root rebuilds / resizes per member.
int run_pipeline(Pipeline &pl, const Member *mem, u32 npma, hipStream_t st) {
    pl.cur_stream = st;
    u64 n = 0, nseg_total = 0;
    bool mixed = false;
    u32 vmax = 1, nsegs = 0, maxreq = 0;
    for (u32 p = 0; p < npma; ++p) {
        lpma *h = mem[p].h;
        h->cur_stream = st;
        h->calls++;
        bool any_ins = false, any_del = false;
        for (u32 r = 0; r < mem[p].nreq; ++r) {
            const Req &q = mem[p].reqs[r];
            n += q.nins + q.ndel;
            any_ins |= q.nins > 0;
            any_del |= q.ndel > 0;
            nsegs += (q.nins ? 1 : 0) + (q.ndel ? 1 : 0);
        }
        mixed |= any_ins && any_del;
        maxreq = std::max(maxreq, mem[p].nreq);
        vmax = std::max(vmax, h->vbits);
        nseg_total += 1ull << h->st.H;
    }
    pl.rounds++;
    if (n == 0) {
        for (u32 p = 0; p < npma; ++p) mem[p].h->version++;
        return LPMA_OK;
    }
    if (n > 0x7FFFFFF0ull) return LPMA_ERR_TOO_LARGE;
    const u32 sbits = bits_for(maxreq);
    const u32 pshift = 2 * vmax + 1 + sbits;
    const u32 hi_bit = pshift + bits_for(npma);
    if (hi_bit > 64) return LPMA_ERR_TOO_LARGE;
    if (pl.graph_on && maxreq <= 1) return run_pipeline_graph(pl, mem, npma, n, pshift, hi_bit, st);
    CHECKED(ensure_batch(pl, n, npma, nsegs, std::min<u64>(n, nseg_total)));
    // A lone PMA's single-segment and wave windows are queued before the
    // host reads the status (below): the large windows that may follow
    // them get their scratch for the worst case now (their merged keys are
    // at most the live keys plus the batch's inserts), or the status is
    // read first.
    bool early = npma == 1;
    if (early) {
        u64 nins = 0;
        for (u32 r = 0; r < mem[0].nreq; ++r) nins += mem[0].reqs[r].nins;
        if (ensure(pl, pl.merged, mem[0].h->live + nins + 1) || ensure(pl, pl.lchunk, n + 1) ||
            ensure(pl, pl.lmbase, n) || ensure(pl, pl.lhdr, 4))
            early = false;
    }

    u32 epoch, pyr_blocks;
    CHECKED(begin_epoch(mem, npma, epoch, st));
    CHECKED(upload_table(pl, mem, npma, pyr_blocks, st));

    // batch sources: each request's deletes (op 0), then its inserts (op 1)
    InlineSegs in;
    memset(&in, 0, sizeof(in));
    {
        u64 off = 0;
        u32 k = 0;
        for (u32 p = 0; p < npma; ++p) {
            lpma *h = mem[p].h;
            for (u32 r = 0; r < mem[p].nreq; ++r) {
                const Req &q = mem[p].reqs[r];
                for (int op = 0; op < 2; ++op) {
                    const u64 *keys = op ? q.ins : q.del;
                    u64 cnt = op ? q.nins : q.ndel;
                    if (!cnt) continue;
                    BatchSeg &b = pl.h_segs[k++];
                    memset(&b, 0, sizeof(b));
                    b.keys = keys;
                    b.n = cnt;
                    b.off = off;
                    b.pid = p;
                    b.op = (u32)op;
                    b.V = h->V;
                    b.vbits = h->vbits;
                    b.seq = r;
                    off += cnt;
                }
            }
        }
        if (nsegs <= INLINE_SEGS) {
            for (u32 k2 = 0; k2 < nsegs; ++k2) in.s[k2] = pl.h_segs[k2];
            in.n = nsegs;
        } else {
            HIPCHECK(hipMemcpyAsync(pl.segs.p, pl.h_segs, nsegs * sizeof(BatchSeg), hipMemcpyHostToDevice, st));
        }
    }

    Group g;
    g.tab = pl.tab.p;
    g.pstat = pstat_of(pl);
    g.lvmask = pl.lvmask.p;
    g.stats = pl.d_stats;
    g.npma = npma;
    g.pshift = pshift;
    g.sbits = sbits;
    g.epoch = epoch;
    g.log2S = mem[0].h->log2S;
    g.seg_log2G = seg_lanes_log2(g.log2S, pl.seg_slots_log2);
    g.rh = nullptr;
    g.inline_one = npma == 1 ? 1u : 0u;
    g.defer_rebuild = 0;
    g.one = mem[0].h->st;
    BatchBufs bb = batch_bufs(pl);
    // bit 0 of a packed key is the operation (with the request order above
    // it): sorted only when a member mixes them or merges requests
    const u32 lo_bit = (mixed || sbits) ? 0 : 1;
    u64 *pk;
    {
        PhaseTimer t(pl, st, PH_SORT);
        if (npma == 1 && nsegs <= INLINE_SEGS && sort_carries_first(n)) {
            lpma *h0 = mem[0].h;
            const u32 chunks = h0->pyr_owed ? pyramid_blocks(h0->st.H) : 0u;
            pk = pack_and_sort_first(g, bb.status, bb.wfirst, in, n, pshift, sbits, hi_bit, pl.pk0.p, pl.pk1.p,
                                     pl.ukey.p, pl.sort_scr.p, chunks, st);
            h0->pyr_owed = false;
        } else {
            HIPCHECK(pay_owed_pyramids(mem, npma, st));
            HIPCHECK(batch_begin(g, bb.status, in, pl.segs.p, pl.sort_scr.p, sort_zero_words(), bb.wfirst, st));
            pk = pack_and_sort(pl.segs.p, nsegs, n, nullptr, pshift, sbits, lo_bit, hi_bit, pl.pk0.p, pl.pk1.p,
                               pl.ukey.p, pl.sort_scr.p, &bb.status->flags, st);
        }
        HIPCHECK(hipGetLastError());
    }
    {
        PhaseTimer t(pl, st, PH_LOCATE);
        HIPCHECK(locate_and_compact(g, pk, n, bb, st));
        HIPCHECK(density_check(g, n, bb, st));
    }
    {
        // window election before the single host synchronisation, so the
        // host learns the exact work of every rebalance class and sizes each
        // grid to it (idle workgroups still cost dispatch time)
        PhaseTimer t(pl, st, PH_ELECT);
        HIPCHECK(elect_windows(g, n, bb, st));
    }
    const u64 out_bytes = STATUS_BYTES + (u64)npma * sizeof(PmaStat);
    // A lone PMA's single-segment and wave windows are queued before the
    // host reads the status, with grids sized for the batch and the window
    // counts read on the device, so they run while the host waits for the
    // status and queues the rest: a PMA with a root rebuild or resize has no
    // windows, so a round that fails after the status changes nothing.
    HIPCHECK(hipMemcpyAsync(pl.h_out, pl.out.p, out_bytes, hipMemcpyDeviceToHost, st));
    if (early) {
        if (!pl.ev_status) HIPCHECK(hipEventCreateWithFlags(&pl.ev_status, hipEventDisableTiming));
        HIPCHECK(hipEventRecord(pl.ev_status, st));
        {
            LaunchTimer t(pl, PH_SEG);
            HIPCHECK(rebalance_seg_wave_bound(g, bb, n, st, t.a, t.b));
            pl.launches_seg++;
        }
        HIPCHECK(hipEventSynchronize(pl.ev_status));
    } else {
        HIPCHECK(hipStreamSynchronize(st));
    }
    return finish_round(pl, mem, npma, g, pyr_blocks, false, st, early);
}

// Graph replay of a round: the host writes the batch descriptors and the
// round header into pinned memory and launches one captured graph (copies
// in, sort, locate, density check, election, every rebalance class and the
// pyramid with fixed grids, status copy out), then synchronises once at the
// end.  Grids are sized for a batch capacity; the kernels read the batch
// size, epoch and scratch pointer from the copied header and the members'
// geometry from the device table, so a graph stays valid across rounds
// until the capacity or a batch buffer changes.

// batch capacity of n elements: the next of 4096 * {1, 1.25, 1.5, 1.75} * 2^k,
// so the grids exceed the batch by at most a quarter
u64 graph_capacity(u64 n) {
    for (u64 b = 4096;; b <<= 1)
        for (u64 q = 4; q < 8; ++q)
            if (b * q / 4 >= n) return b * q / 4;
}

static constexpr u32 GRAPH_PYR_GRID = 2048;

int prepare_graph_buffers(Pipeline &pl, u64 cap, u32 npma, u64 merged_keys) {
    CHECKED(ensure_batch(pl, cap, npma, 2 * npma, cap));
    // merged scratch of the large windows: at most every live key of the
    // group after the round
    CHECKED(ensure(pl, pl.merged, merged_keys + 1));
    CHECKED(ensure(pl, pl.lchunk, cap + 1));
    CHECKED(ensure(pl, pl.lmbase, cap));
    CHECKED(ensure(pl, pl.lhdr, 4));
    CHECKED(ensure(pl, pl.rhdr, 1));
    if (!pl.h_rhdr) {
        if (injected_failure()) return LPMA_ERR_OUT_OF_MEMORY;
        if (hipHostMalloc((void **)&pl.h_rhdr, sizeof(RoundHdr), hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            pl.h_rhdr = nullptr;
            return LPMA_ERR_OUT_OF_MEMORY;
        }
    }
    if (!pl.cap_stream) HIPCHECK(hipStreamCreateWithFlags(&pl.cap_stream, hipStreamNonBlocking));
    return LPMA_OK;
}

Group graph_group(Pipeline &pl, u32 npma, u32 pshift, u32 log2S) {
    Group g;
    g.tab = pl.tab.p;
    g.pstat = pstat_of(pl);
    g.lvmask = pl.lvmask.p;
    g.stats = pl.d_stats;
    g.npma = npma;
    g.pshift = pshift;
    g.sbits = 0;
    g.epoch = 0;
    g.log2S = log2S;
    g.seg_log2G = seg_lanes_log2(log2S, pl.seg_slots_log2);
    g.rh = pl.rhdr.p;
    // a captured graph outlives resizes: geometry always from the table
    g.inline_one = 0;
    // several PMAs: no windows in a round where one of them rebuilds (its
    // allocations come first, then the election runs again)
    g.defer_rebuild = npma > 1 ? 1u : 0u;
    memset(&g.one, 0, sizeof(g.one));
    return g;
}

// The graph of capacity `cap`, captured now unless a valid one is cached.
int graph_for(Pipeline &pl, u64 cap, u32 npma, u32 pshift, u32 hi_bit, u32 log2S, hipGraphExec_t *out,
              hipStream_t st) {
    const u32 nsegs = 2 * npma;
    std::vector<u64> key = {cap, npma, pshift, hi_bit, log2S, (u64)(i64)pl.seg_slots_log2,
                            (u64)pl.pk0.p, (u64)pl.pk1.p, (u64)pl.sort_scr.p, (u64)pl.ukey.p, (u64)pl.useg.p,
                            (u64)pl.ckey.p, (u64)pl.cseg.p, (u64)pl.pd.p, (u64)pl.rlevel.p, (u64)pl.rwin.p, (u64)pl.part.p,
                            (u64)pl.lchunk.p, (u64)pl.lmbase.p, (u64)pl.lhdr.p, (u64)pl.out.p, (u64)pl.tab.p,
                            (u64)pl.segs.p, (u64)pl.rhdr.p, (u64)pl.d_stats, (u64)pl.h_out, (u64)pl.h_segs,
                            (u64)pl.h_rhdr, (u64)pl.lvmask.p, (u64)pl.wfirst.p};
    for (int c = 0; c < NUM_CLASSES; ++c) key.push_back((u64)pl.work[c].p);
    auto it = pl.graphs.find(cap);
    if (it != pl.graphs.end() && it->second.key != key) {
        HIPCHECK(hipGraphExecDestroy(it->second.exec));
        pl.graphs.erase(it);
        it = pl.graphs.end();
    }
    if (it == pl.graphs.end()) {
        const Group g = graph_group(pl, npma, pshift, log2S);
        const u64 out_bytes = STATUS_BYTES + (u64)npma * sizeof(PmaStat);
        // the capture stream must not start before the work queued on st
        HIPCHECK(hipStreamSynchronize(st));
        hipStream_t cs = pl.cap_stream;
        HIPCHECK(hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed));
        BatchBufs bb = batch_bufs(pl);
        InlineSegs none;
        memset(&none, 0, sizeof(none));
        bool ok = hipMemcpyAsync(pl.rhdr.p, pl.h_rhdr, sizeof(RoundHdr), hipMemcpyHostToDevice, cs) == hipSuccess &&
                  hipMemcpyAsync(pl.segs.p, pl.h_segs, nsegs * sizeof(BatchSeg), hipMemcpyHostToDevice, cs) ==
                      hipSuccess &&
                  batch_begin(g, bb.status, none, pl.segs.p, pl.sort_scr.p, sort_zero_words(), bb.wfirst, cs) == hipSuccess;
        if (ok) {
            u64 *pk = pack_and_sort(pl.segs.p, nsegs, cap, &pl.rhdr.p->n, pshift, 0, 0, hi_bit, pl.pk0.p, pl.pk1.p,
                                    pl.ukey.p, pl.sort_scr.p, &bb.status->flags, cs);
            ok = hipGetLastError() == hipSuccess && locate_and_compact(g, pk, cap, bb, cs) == hipSuccess &&
                 density_check(g, cap, bb, cs) == hipSuccess && elect_windows(g, cap, bb, cs) == hipSuccess &&
                 rebalance_fixed(g, bb, (u32)cap, nullptr, cs) == hipSuccess &&
                 group_pyramid(g, GRAPH_PYR_GRID, cs) == hipSuccess &&
                 hipMemcpyAsync(pl.h_out, pl.out.p, out_bytes, hipMemcpyDeviceToHost, cs) == hipSuccess;
        }
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(cs, &graph);
        if (!ok || e != hipSuccess || !graph) {
            (void)hipGetLastError();
            if (graph) (void)hipGraphDestroy(graph);
            return LPMA_ERR_DEVICE;
        }
        hipGraphExec_t exec = nullptr;
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return LPMA_ERR_DEVICE;
        }
        it = pl.graphs.emplace(cap, Pipeline::GraphEntry{exec, key}).first;
        pl.graph_captures++;
    }
    *out = it->second.exec;
    return LPMA_OK;
}

int run_pipeline_graph(Pipeline &pl, const Member *mem, u32 npma, u64 n, u32 pshift, u32 hi_bit, hipStream_t st) {
    const u64 cap = graph_capacity(n);
    u64 live_sum = 0;
    for (u32 p = 0; p < npma; ++p) live_sum += mem[p].h->live;
    CHECKED(prepare_graph_buffers(pl, cap, npma, live_sum + n));
    HIPCHECK(pay_owed_pyramids(mem, npma, st));

    u32 epoch, pyr_blocks;
    CHECKED(begin_epoch(mem, npma, epoch, st));
    CHECKED(upload_table(pl, mem, npma, pyr_blocks, st));

    // fixed descriptor layout: member p's deletes at 2p, inserts at 2p+1
    u64 off = 0;
    for (u32 p = 0; p < npma; ++p) {
        lpma *h = mem[p].h;
        const Req none{nullptr, 0, nullptr, 0};
        const Req &q = mem[p].nreq ? mem[p].reqs[0] : none;
        for (int op = 0; op < 2; ++op) {
            BatchSeg &b = pl.h_segs[2 * p + op];
            memset(&b, 0, sizeof(b));
            b.keys = op ? q.ins : q.del;
            b.n = op ? q.nins : q.ndel;
            b.off = off;
            b.pid = p;
            b.op = (u32)op;
            b.V = h->V;
            b.vbits = h->vbits;
            off += b.n;
        }
    }
    const u32 log2S = mem[0].h->log2S;
    hipGraphExec_t exec;
    CHECKED(graph_for(pl, cap, npma, pshift, hi_bit, log2S, &exec, st));
    pl.h_rhdr->n = n;
    pl.h_rhdr->merged = pl.merged.p;
    pl.h_rhdr->epoch = epoch;
    pl.h_rhdr->pad_ = 0;
    {
        PhaseTimer t(pl, st, PH_GRAPH);
        HIPCHECK(hipGraphLaunch(exec, st));
    }
    pl.graph_launches++;
    HIPCHECK(hipStreamSynchronize(st));
    Group g = graph_group(pl, npma, pshift, log2S);
    g.epoch = epoch;  // a repeated election runs eagerly, outside the graph
    g.rh = nullptr;
    return finish_round(pl, mem, npma, g, pyr_blocks, true, st);
}

// Sizes the batch buffers for rounds of up to max_batch elements and, in
// graph mode, captures the graph of every capacity up to it now.
int reserve_pipeline(Pipeline &pl, lpma *const *members, u32 npma, u64 max_batch, hipStream_t st) {
    pl.cur_stream = st;
    u32 vmax = 1;
    u64 live_sum = 0, nseg_total = 0;
    for (u32 p = 0; p < npma; ++p) {
        vmax = std::max(vmax, members[p]->vbits);
        live_sum += members[p]->live;
        nseg_total += 1ull << members[p]->st.H;
    }
    if (max_batch == 0) return LPMA_OK;
    if (max_batch > 0x7FFFFFF0ull) return LPMA_ERR_TOO_LARGE;
    const u32 pshift = 2 * vmax + 1;
    const u32 hi_bit = pshift + bits_for(npma);
    if (hi_bit > 64) return LPMA_ERR_INVALID_ARGUMENT;
    if (!pl.graph_on) {
        // work lists for as many runs as keys (the segments may multiply
        // first), and the large-window scratch of a batch that size
        int r = ensure_batch(pl, max_batch, npma, 2 * npma, max_batch);
        if (!r && !(r = ensure(pl, pl.lchunk, max_batch + 1)) && !(r = ensure(pl, pl.lmbase, max_batch)))
            r = ensure(pl, pl.lhdr, 4);
        return r;
    }
    const u64 top = graph_capacity(max_batch);
    // headroom for the live keys to grow before the scratch is reallocated
    CHECKED(prepare_graph_buffers(pl, top, npma, live_sum + live_sum / 4 + max_batch));
    u32 pyr_blocks;
    Member m0{members[0], nullptr, 0};
    std::vector<Member> mem(npma, m0);
    for (u32 p = 0; p < npma; ++p) mem[p].h = members[p];
    CHECKED(upload_table(pl, mem.data(), npma, pyr_blocks, st));
    for (u64 c = 4096; c <= top; c = graph_capacity(c + 1)) {
        hipGraphExec_t exec;
        CHECKED(graph_for(pl, c, npma, pshift, hi_bit, members[0]->log2S, &exec, st));
    }
    HIPCHECK(hipStreamSynchronize(st));
    return LPMA_OK;
}

int run_update(lpma *h, const u64 *ins, u64 nins, const u64 *del, u64 ndel, hipStream_t st) {
    Req q{ins, nins, del, ndel};
    Member m{h, &q, 1};
    return run_pipeline(h->pl, &m, 1, st);
}

void fill_stats(const u64 *rows, u32 log2S, lpma_stats_t *out) {
    u64 s[ST_NUM] = {0};
    for (u32 r = 0; r < ST_ROWS; ++r)
        for (u32 i = 0; i < ST_NUM; ++i) s[i] += rows[(size_t)r * ST_NUM + i];
    for (int l = 0; l < 32; ++l) {
        out->windows_per_level[l] += s[ST_WINDOWS + l];
        out->slots_per_level[l] += s[ST_WINDOWS + l] << (l + log2S);
    }
    out->marker_moves += s[ST_MARKERS];
    out->alg_bytes_wave += s[ST_ALG_BYTES_WAVE];
    out->alg_bytes_block += s[ST_ALG_BYTES_BLOCK];
    out->alg_bytes_large += s[ST_ALG_BYTES_LARGE];
    out->alg_bytes_seg += s[ST_ALG_BYTES_SEG];
}

int read_stats(Pipeline &pl, u32 log2S, lpma_stats_t *out) {
    std::vector<u64> rows((size_t)ST_ROWS * ST_NUM);
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(rows.data(), pl.d_stats, rows.size() * sizeof(u64), hipMemcpyDeviceToHost));
    fill_stats(rows.data(), log2S, out);
    return LPMA_OK;
}

void add_profile(Pipeline &pl, lpma_profile_t *out) {
    collect_events(pl);
    out->sort_ms += pl.phase_ms[PH_SORT];
    out->locate_ms += pl.phase_ms[PH_LOCATE];
    out->elect_ms += pl.phase_ms[PH_ELECT];
    out->rebalance_wave_ms += pl.phase_ms[PH_WAVE];
    out->rebalance_block_ms += pl.phase_ms[PH_BLOCK];
    out->large_ms += pl.phase_ms[PH_LARGE];
    out->rebuild_ms += pl.phase_ms[PH_REBUILD];
    out->pyramid_ms += pl.phase_ms[PH_PYR];
    out->rebalance_seg_ms += pl.phase_ms[PH_SEG];
    out->launches_wave += pl.launches_wave;
    out->launches_block += pl.launches_block;
    out->launches_seg += pl.launches_seg;
    out->graph_ms += pl.phase_ms[PH_GRAPH];
    out->graph_launches += pl.graph_launches;
    out->graph_captures += pl.graph_captures;
    for (double &v : pl.phase_ms) v = 0;
    pl.launches_wave = pl.launches_block = pl.launches_seg = 0;
    pl.graph_launches = pl.graph_captures = 0;
}

}  // namespace

// -------------------------------------------------------------------- C ABI

extern "C" {

void lpma_config_default(lpma_config_t *cfg) {
    if (!cfg) return;
    cfg->segment_slots = 32;
    cfg->density_lower_leaf = 0.42;
    cfg->density_lower_root = 0.42;
    cfg->density_upper_leaf = 1.0;
    cfg->density_upper_root = 1.0;
    cfg->device = -1;
}

const char *lpma_strerror(int code) {
    switch (code) {
        case LPMA_OK: return "ok";
        case LPMA_ERR_INVALID_ARGUMENT: return "invalid argument";
        case LPMA_ERR_INVALID_KEY: return "invalid key (src or dst >= num_vertices)";
        case LPMA_ERR_OUT_OF_MEMORY: return "out of device memory or pool budget";
        case LPMA_ERR_DEVICE: return "HIP runtime error";
        case LPMA_ERR_TOO_LARGE: return "batch or capacity too large";
        case LPMA_ERR_NOT_BUILT: return "device code not available";
        default: return "unknown status";
    }
}

int lpma_create(lpma_t **out, uint32_t num_vertices, const lpma_config_t *cfg_in, void *stream) {
    if (!out) return LPMA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    lpma_config_t cfg;
    lpma_config_default(&cfg);
    if (cfg_in) cfg = *cfg_in;
    u32 S = cfg.segment_slots;
    if (num_vertices < 1 || num_vertices >= (1u << 31)) return LPMA_ERR_INVALID_ARGUMENT;
    if (S < 2 || S > 1024 || (S & (S - 1))) return LPMA_ERR_INVALID_ARGUMENT;
    if (!(cfg.density_lower_leaf >= 0 && cfg.density_lower_root >= 0 && cfg.density_upper_leaf <= 1.0 &&
          cfg.density_upper_root <= 1.0 && cfg.density_upper_root > 0.0 && cfg.density_upper_leaf > 0.0 &&
          cfg.density_lower_leaf <= cfg.density_upper_leaf && cfg.density_lower_root <= cfg.density_upper_root))
        return LPMA_ERR_INVALID_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return LPMA_ERR_DEVICE;
    }
    int dev = cfg.device;
    if (dev < 0) (void)hipGetDevice(&dev);
    if (dev >= ndev) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(dev);
    hipStream_t st = (hipStream_t)stream;

    lpma *h = new (std::nothrow) lpma();
    if (!h) return LPMA_ERR_OUT_OF_MEMORY;
    h->cur_stream = st;
    h->device = dev;
    h->cfg = cfg;
    h->cfg.device = dev;
    h->V = num_vertices;
    u32 vb = 1;
    while ((1ull << vb) < num_vertices) ++vb;
    h->vbits = vb;
    h->log2S = (u32)__builtin_ctz(S);
    memset(&h->st, 0, sizeof(h->st));
    h->st.log2S = h->log2S;
    h->st.V = h->V;
    h->st.vbits = h->vbits;

    int r = LPMA_OK;
    // the allocation-free export's static scratch on this device
    if (export_scratch_setup(dev) != hipSuccess) {
        (void)hipGetLastError();
        lpma_destroy(h);
        return LPMA_ERR_DEVICE;
    }
    if (hipMalloc(&h->d_err, 16) != hipSuccess) {
        (void)hipGetLastError();
        h->d_err = nullptr;
        lpma_destroy(h);
        return LPMA_ERR_OUT_OF_MEMORY;
    }
    if ((r = pipeline_init(h->pl, dev, st)) != LPMA_OK) {
        lpma_destroy(h);
        return r;
    }
    if ((r = ensure(h, h->row_off, (u64)num_vertices + 1)) != LPMA_OK) {
        lpma_destroy(h);
        return r;
    }
    if (fill_u32(h->row_off.p, (u64)num_vertices + 1, 0, st) != hipSuccess) {
        (void)hipGetLastError();
        lpma_destroy(h);
        return LPMA_ERR_DEVICE;
    }
    h->st.row_off = h->row_off.p;

    // one marker per row, spread at the smallest fitting height
    // (init_csr_rpma inserts them through update_rpma, rpma.cuh:1819-1823)
    u32 H = fit_height(h, num_vertices);
    if (!height_ok(h, H)) {
        lpma_destroy(h);
        return LPMA_ERR_TOO_LARGE;
    }
    if ((r = ensure(h->pl, h->pl.merged, (u64)num_vertices + 1)) != LPMA_OK) {
        lpma_destroy(h);
        return r;
    }
    if (make_markers(h->pl.merged.p, num_vertices, st) != hipSuccess) {
        (void)hipGetLastError();
        lpma_destroy(h);
        return LPMA_ERR_DEVICE;
    }
    h->st.H = H;
    compute_density(h);
    {
        Geometry g;
        if ((r = alloc_geometry(h, H, g, st)) != LPMA_OK ||
            (r = spread_and_install(h, H, g, h->pl.merged.p, num_vertices, st)) != LPMA_OK) {
            lpma_destroy(h);
            return r;
        }
    }
    h->live = num_vertices;
    if (hipStreamSynchronize(st) != hipSuccess) {
        (void)hipGetLastError();
        lpma_destroy(h);
        return LPMA_ERR_DEVICE;
    }
    *out = h;
    return LPMA_OK;
}

int lpma_destroy(lpma_t *h) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    (void)hipDeviceSynchronize();
    pipeline_free(h->pl);
    for (auto &kv : h->spares)
        for (Geometry &g : kv.second) release_geometry(h, g);
    release(h, h->keys); release(h, h->cnt); release(h, h->pyr); release(h, h->owner);
    release(h, h->row_off); release(h, h->sep);
    release(h, h->aq0); release(h, h->apoff); release(h, h->apcb);
    release(h, h->af0); release(h, h->af1); release(h, h->af2); release(h, h->ared); release(h, h->apart);
    release(h, h->apairs);
    if (h->d_err) (void)hipFree(h->d_err);
    delete h;
    return LPMA_OK;
}

int lpma_update(lpma_t *h, const uint64_t *d_keys, uint64_t n, int op, void *stream) {
    if (!h || (!d_keys && n) || (op != LPMA_OP_INSERT && op != LPMA_OP_DELETE)) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    if (op == LPMA_OP_INSERT) return run_update(h, d_keys, n, nullptr, 0, (hipStream_t)stream);
    return run_update(h, nullptr, 0, d_keys, n, (hipStream_t)stream);
}

int lpma_update_mixed(lpma_t *h, const uint64_t *d_ins, uint64_t n_ins, const uint64_t *d_del, uint64_t n_del,
                      void *stream) {
    if (!h || (!d_ins && n_ins) || (!d_del && n_del)) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    return run_update(h, d_ins, n_ins, d_del, n_del, (hipStream_t)stream);
}

int lpma_update_sequence(lpma_t *h, const uint64_t *const *d_ins, const uint64_t *n_ins,
                         const uint64_t *const *d_del, const uint64_t *n_del, uint32_t count, void *stream) {
    if (!h || (count && !n_ins && !n_del)) return LPMA_ERR_INVALID_ARGUMENT;
    std::vector<Req> reqs(count);
    for (u32 i = 0; i < count; ++i) {
        Req &q = reqs[i];
        q.nins = n_ins ? n_ins[i] : 0;
        q.ndel = n_del ? n_del[i] : 0;
        if ((q.nins && (!d_ins || !d_ins[i])) || (q.ndel && (!d_del || !d_del[i]))) return LPMA_ERR_INVALID_ARGUMENT;
        q.ins = q.nins ? d_ins[i] : nullptr;
        q.del = q.ndel ? d_del[i] : nullptr;
    }
    DevGuard guard(h->device);
    Member m{h, reqs.data(), count};
    return run_pipeline(h->pl, &m, 1, (hipStream_t)stream);
}

int lpma_bulk_load(lpma_t *h, const uint64_t *d_keys, uint64_t n, void *stream) {
    return lpma_update(h, d_keys, n, LPMA_OP_INSERT, stream);
}

int lpma_get_view(const lpma_t *h, lpma_view_t *out) {
    if (!h || !out) return LPMA_ERR_INVALID_ARGUMENT;
    out->keys = h->st.keys;
    out->row_offsets = h->st.row_off;
    out->capacity = (u64)1 << (h->st.H + h->log2S);
    out->live_keys = h->live;
    out->edges = h->live - h->V;
    out->num_vertices = h->V;
    out->segment_slots = 1u << h->log2S;
    out->tree_height = h->st.H;
    out->version = h->version;
    return LPMA_OK;
}

int lpma_copy_view(lpma_t *h, uint64_t *d_keys, uint32_t *d_row_offsets, void *stream) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    hipStream_t st = (hipStream_t)stream;
    if (d_keys) HIPCHECK(copy_device(d_keys, h->st.keys, ((u64)1 << (h->st.H + h->log2S)) * sizeof(u64), st));
    if (d_row_offsets) HIPCHECK(copy_device(d_row_offsets, h->st.row_off, ((u64)h->V + 1) * sizeof(u32), st));
    return LPMA_OK;
}

static int export_common(lpma_t *h, u64 *edges, u32 *row_std, u64 *all, u32 *csr_row, u32 *csr_col,
                         void *stream) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    hipStream_t st = (hipStream_t)stream;
    h->cur_stream = st;
    h->pl.cur_stream = st;
    u64 nseg = 1ull << h->st.H;
    CHECKED(ensure(h->pl, h->pl.gpre, nseg + 1));
    CHECKED(ensure(h->pl, h->pl.scan_scr, scan_scratch_words(nseg) + 8));
    HIPCHECK(export_compacted(h->st, h->pl.gpre.p, h->pl.scan_scr.p, edges, row_std, all, csr_row, csr_col, st));
    return LPMA_OK;
}

int lpma_export_compacted(lpma_t *h, uint64_t *d_edge_keys, uint32_t *d_row_offsets, void *stream) {
    return export_common(h, d_edge_keys, d_row_offsets, nullptr, nullptr, nullptr, stream);
}

int lpma_export_compacted_noalloc(lpma_t *h, uint64_t *d_edge_keys, uint32_t *d_row_offsets, void *stream) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    if ((1ull << h->st.H) > export_static_max_segments()) return LPMA_ERR_TOO_LARGE;
    DevGuard guard(h->device);
    HIPCHECK(export_compacted_static(h->device, h->st, d_edge_keys, d_row_offsets, (hipStream_t)stream));
    return LPMA_OK;
}

int lpma_export_data(lpma_t *h, uint64_t *d_keys, void *stream) {
    if (!d_keys) return LPMA_ERR_INVALID_ARGUMENT;
    return export_common(h, nullptr, nullptr, d_keys, nullptr, nullptr, stream);
}

int lpma_to_csr(lpma_t *h, uint32_t *d_row_offset, uint32_t *d_col_index, void *stream) {
    if (!d_row_offset || !d_col_index) return LPMA_ERR_INVALID_ARGUMENT;
    return export_common(h, nullptr, nullptr, nullptr, d_row_offset, d_col_index, stream);
}

int lpma_edge_query(lpma_t *h, const uint64_t *d_keys, uint64_t n, uint32_t *d_result, void *stream) {
    if (!h || ((!d_keys || !d_result) && n)) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    HIPCHECK(edge_query(h->st, d_keys, n, d_result, (hipStream_t)stream));
    return LPMA_OK;
}

int lpma_neighbor_count(lpma_t *h, const uint32_t *d_srcs, uint64_t n, uint32_t *d_len, void *stream) {
    if (!h || ((!d_srcs || !d_len) && n)) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    HIPCHECK(neighbor_count(h->st, d_srcs, n, d_len, (hipStream_t)stream));
    return LPMA_OK;
}

int lpma_neighbor_gather(lpma_t *h, const uint32_t *d_srcs, uint64_t n, const uint32_t *d_offsets, uint32_t *d_out,
                         void *stream) {
    if (!h || ((!d_srcs || !d_offsets || !d_out) && n)) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    HIPCHECK(neighbor_gather(h->st, d_srcs, n, d_offsets, d_out, (hipStream_t)stream));
    return LPMA_OK;
}

int lpma_check(lpma_t *h, uint64_t *violations, void *stream) {
    if (!h || !violations) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    hipStream_t st = (hipStream_t)stream;
    HIPCHECK(fill_u32(h->d_err, 1, 0, st));
    HIPCHECK(check_invariants(h->st, h->d_err, st));
    u32 e = 0;
    HIPCHECK(hipMemcpyAsync(&e, h->d_err, sizeof(u32), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    *violations = e;
    return LPMA_OK;
}

// Level-synchronous BFS over a DevState with scratch from the pool (so
// concurrent searches over one structure do not share buffers).
static int bfs_core(int device, const DevState &d, u32 src, u32 *d_levels, u32 *depth, u64 *edges_scanned,
                    u32 expand_grid, hipStream_t st) {
    const u64 V = d.V;
    DBuf<u32> q0, q1, len, cnt, scan, runs, visited;
    DBuf<u64> first;
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(st);
        release_on(device, q0); release_on(device, q1); release_on(device, len); release_on(device, cnt);
        release_on(device, scan); release_on(device, runs); release_on(device, first); release_on(device, visited);
    };
    int r;
    if ((r = ensure_on(device, st, q0, V + 1)) || (r = ensure_on(device, st, q1, V + 1)) ||
        (r = ensure_on(device, st, len, V + 2)) || (r = ensure_on(device, st, scan, scan_scratch_words(V + 1) + 8)) ||
        (r = ensure_on(device, st, cnt, 4)) || (r = ensure_on(device, st, first, V + 1)) ||
        (r = ensure_on(device, st, runs, bfs_run_words(nseg_of(d) << d.log2S))) ||
        (r = ensure_on(device, st, visited, bfs_visited_words(V)))) {
        cleanup();
        return r;
    }
    if (bfs_init(d_levels, d.V, src, q0.p, cnt.p, visited.p, st) != hipSuccess) {
        (void)hipGetLastError();
        cleanup();
        return LPMA_ERR_DEVICE;
    }
    u32 *q = q0.p, *qn_buf = q1.p;
    u32 qn = 1, level = 0;
    u64 scanned = 0;
    u32 c[2];
    int rc = LPMA_OK;
    while (qn) {
        if (bfs_level(d, q, qn, len.p, first.p, runs.p, scan.p, d_levels, visited.p, level, qn_buf, cnt.p, expand_grid,
                      st) != hipSuccess ||
            hipMemcpyAsync(c, cnt.p, sizeof(c), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess || fill_u32(cnt.p, 2, 0, st) != hipSuccess) {
            (void)hipGetLastError();
            rc = LPMA_ERR_DEVICE;
            break;
        }
        scanned += c[1];
        qn = c[0];
        std::swap(q, qn_buf);
        ++level;
    }
    cleanup();
    if (rc == LPMA_OK) {
        if (depth) *depth = level;
        if (edges_scanned) *edges_scanned = scanned;
    }
    return rc;
}

int lpma_bfs(lpma_t *h, uint32_t src, uint32_t *d_levels, uint32_t *depth, uint64_t *edges_scanned, void *stream) {
    if (!h || !d_levels) return LPMA_ERR_INVALID_ARGUMENT;
    if (src >= h->V) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    return bfs_core(h->device, h->st, src, d_levels, depth, edges_scanned, 0, (hipStream_t)stream);
}

// ---------------------------------------------------------------- snapshots

uint64_t lpma_snapshot_bytes(const lpma_t *h) {
    if (!h) return 0;
    u64 nseg = 1ull << h->st.H;
    return (nseg << h->log2S) * 8 + nseg * 12 + ((u64)h->V + 1) * 4;
}

int lpma_snapshot_create(lpma_snapshot_t **out, lpma_t *h, void *stream) {
    if (!out || !h) return LPMA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    DevGuard guard(h->device);
    hipStream_t st = (hipStream_t)stream;
    lpma_snapshot *s = new (std::nothrow) lpma_snapshot();
    if (!s) return LPMA_ERR_OUT_OF_MEMORY;
    s->device = h->device;
    const u64 nseg = 1ull << h->st.H, cap = nseg << h->log2S;
    int r;
    if ((r = ensure_on(h->device, st, s->keys, cap)) || (r = ensure_on(h->device, st, s->sep, nseg)) ||
        (r = ensure_on(h->device, st, s->cnt, nseg)) || (r = ensure_on(h->device, st, s->row_off, (u64)h->V + 1))) {
        lpma_snapshot_destroy(s);
        return r;
    }
    if (copy_device(s->keys.p, h->st.keys, cap * 8, st) != hipSuccess ||
        copy_device(s->sep.p, h->st.sep, nseg * 8, st) != hipSuccess ||
        copy_device(s->cnt.p, h->st.cnt, nseg * 4, st) != hipSuccess ||
        copy_device(s->row_off.p, h->st.row_off, ((u64)h->V + 1) * 4, st) != hipSuccess) {
        (void)hipGetLastError();
        lpma_snapshot_destroy(s);
        return LPMA_ERR_DEVICE;
    }
    s->st = h->st;
    s->st.keys = s->keys.p;
    s->st.sep = s->sep.p;
    s->st.cnt = s->cnt.p;
    s->st.row_off = s->row_off.p;
    s->st.pyr = nullptr;
    s->st.owner = nullptr;
    s->version = h->version;
    s->live = h->live;
    s->V = h->V;
    s->log2S = h->log2S;
    if (hipEventCreateWithFlags(&s->ready, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(s->ready, st) != hipSuccess) {
        (void)hipGetLastError();
        lpma_snapshot_destroy(s);
        return LPMA_ERR_DEVICE;
    }
    *out = s;
    return LPMA_OK;
}

// reads on any stream start after the snapshot's copies
static int snapshot_wait(lpma_snapshot *s, hipStream_t st) {
    HIPCHECK(hipStreamWaitEvent(st, s->ready, 0));
    return LPMA_OK;
}

int lpma_snapshot_destroy(lpma_snapshot_t *s) {
    if (!s) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(s->device);
    if (s->ready) {
        (void)hipEventSynchronize(s->ready);
        (void)hipEventDestroy(s->ready);
    }
    release_on(s->device, s->keys); release_on(s->device, s->sep); release_on(s->device, s->cnt);
    release_on(s->device, s->row_off);
    delete s;
    return LPMA_OK;
}

int lpma_snapshot_get_view(const lpma_snapshot_t *s, lpma_view_t *out) {
    if (!s || !out) return LPMA_ERR_INVALID_ARGUMENT;
    out->keys = s->st.keys;
    out->row_offsets = s->st.row_off;
    out->capacity = (u64)1 << (s->st.H + s->log2S);
    out->live_keys = s->live;
    out->edges = s->live - s->V;
    out->num_vertices = s->V;
    out->segment_slots = 1u << s->log2S;
    out->tree_height = s->st.H;
    out->version = s->version;
    return LPMA_OK;
}

int lpma_snapshot_neighbor_count(lpma_snapshot_t *s, const uint32_t *d_srcs, uint64_t n, uint32_t *d_len,
                                 void *stream) {
    if (!s || ((!d_srcs || !d_len) && n)) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(s->device);
    CHECKED(snapshot_wait(s, (hipStream_t)stream));
    HIPCHECK(neighbor_count(s->st, d_srcs, n, d_len, (hipStream_t)stream));
    return LPMA_OK;
}

int lpma_snapshot_neighbor_gather(lpma_snapshot_t *s, const uint32_t *d_srcs, uint64_t n, const uint32_t *d_offsets,
                                  uint32_t *d_out, void *stream) {
    if (!s || ((!d_srcs || !d_offsets || !d_out) && n)) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(s->device);
    CHECKED(snapshot_wait(s, (hipStream_t)stream));
    HIPCHECK(neighbor_gather(s->st, d_srcs, n, d_offsets, d_out, (hipStream_t)stream));
    return LPMA_OK;
}

int lpma_snapshot_edge_query(lpma_snapshot_t *s, const uint64_t *d_keys, uint64_t n, uint32_t *d_result,
                             void *stream) {
    if (!s || ((!d_keys || !d_result) && n)) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(s->device);
    CHECKED(snapshot_wait(s, (hipStream_t)stream));
    HIPCHECK(edge_query(s->st, d_keys, n, d_result, (hipStream_t)stream));
    return LPMA_OK;
}

int lpma_snapshot_bfs(lpma_snapshot_t *s, uint32_t src, uint32_t *d_levels, uint32_t *depth, uint64_t *edges_scanned,
                      void *stream) {
    if (!s || !d_levels || src >= s->V) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(s->device);
    CHECKED(snapshot_wait(s, (hipStream_t)stream));
    return bfs_core(s->device, s->st, src, d_levels, depth, edges_scanned, s->search_grid, (hipStream_t)stream);
}

int lpma_snapshot_set_search_width(lpma_snapshot_t *s, uint32_t workgroups) {
    if (!s || workgroups > 65536) return LPMA_ERR_INVALID_ARGUMENT;
    s->search_grid = workgroups;
    return LPMA_OK;
}

int lpma_pagerank(lpma_t *h, float damping, float tolerance, uint32_t max_iter, float *d_rank, uint32_t *iterations,
                  float *last_delta, void *stream) {
    if (!h || !d_rank || !(damping >= 0.f && damping <= 1.f) || max_iter == 0) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    hipStream_t st = (hipStream_t)stream;
    h->cur_stream = st;
    const u64 V = h->V;
    u64 nseg = 1ull << h->st.H;
    h->pl.cur_stream = st;
    CHECKED(ensure(h->pl, h->pl.gpre, nseg + 1));
    CHECKED(ensure(h->pl, h->pl.scan_scr, scan_scratch_words(nseg) + 8));
    CHECKED(ensure(h, h->aq0, V));   // out-degrees
    CHECKED(ensure(h, h->af0, V));   // contributions
    CHECKED(ensure(h, h->af2, V));   // second rank vector
    CHECKED(ensure(h, h->ared, 2 * pr_red_floats()));
    HIPCHECK(pr_outdeg(h->st, h->pl.gpre.p, h->pl.scan_scr.p, h->aq0.p, st));
    // Binned push where the destination bins fit the bin pass's LDS, else
    // the atomic push into per-XCD accumulators.
    const bool binned = pr_bins_fit(h->st.V);
    PrBins bins{};
    if (binned) {
        bins = pr_bins_plan(h->st);
        const u64 nc = pr_bins_count_words(bins);
        CHECKED(ensure(h, h->apoff, nc));
        CHECKED(ensure(h->pl, h->pl.scan_scr, scan_scratch_words(nc) + 8));
        HIPCHECK(pr_bins_count(h->st, bins, h->apoff.p, h->pl.scan_scr.p, st));
        u32 total = 0;
        HIPCHECK(hipMemcpyAsync(&total, h->apoff.p + (nc - 1), sizeof(u32), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        pr_bins_chunking(bins, total);
        CHECKED(ensure(h, h->apcb, (u64)bins.bins + 1));
        CHECKED(ensure(h, h->apairs, std::max<u64>(1, total)));
        CHECKED(ensure(h, h->apart, pr_bins_partial_floats(bins)));
        HIPCHECK(pr_bins_chunks(bins, h->apoff.p, h->apcb.p, h->pl.scan_scr.p, st));
    } else {
        CHECKED(ensure(h, h->af1, pr_acc_floats(h->V)));  // compensated accumulators
    }
    float *red[2] = {h->ared.p, h->ared.p + pr_red_floats()};
    float *vec[2] = {d_rank, h->af2.p};
    HIPCHECK(pr_init(vec[0], h->V, red[0], st));
    u32 it = 0;
    float delta = 0.f;
    int cur = 0;
    while (it < max_iter) {
        if (binned)
            HIPCHECK(pr_iterate_binned(h->st, bins, vec[cur], h->aq0.p, h->af0.p, h->apoff.p, h->apcb.p, h->apairs.p,
                                       h->apart.p, vec[cur ^ 1], damping, red[cur], red[cur ^ 1], st));
        else
            HIPCHECK(pr_iterate(h->st, vec[cur], h->aq0.p, h->af0.p, h->af1.p, vec[cur ^ 1], damping, red[cur],
                                red[cur ^ 1], st));
        cur ^= 1;
        ++it;
        float r[8];
        HIPCHECK(hipMemcpyAsync(r, red[cur], sizeof(r), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        delta = r[4] + r[5];
        if (delta < tolerance) break;
    }
    HIPCHECK(pr_scale(vec[cur], h->V, red[cur], st));
    if (vec[cur] != d_rank) HIPCHECK(copy_device(d_rank, vec[cur], V * sizeof(float), st));
    HIPCHECK(hipStreamSynchronize(st));
    if (iterations) *iterations = it;
    if (last_delta) *last_delta = delta;
    return LPMA_OK;
}

int lpma_get_stats(lpma_t *h, lpma_stats_t *out) {
    if (!h || !out) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    memset(out, 0, sizeof(*out));
    CHECKED(read_stats(h->pl, h->log2S, out));
    out->updates_calls = h->calls;
    out->effective_updates = h->effective;
    out->rebuilds = h->rebuilds;
    out->rebuild_slots = h->rebuild_slots;
    return LPMA_OK;
}

int lpma_reset_stats(lpma_t *h) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemset(h->pl.d_stats, 0, ST_ROWS * ST_NUM * sizeof(u64)));
    h->calls = h->effective = h->rebuilds = h->rebuild_slots = 0;
    return LPMA_OK;
}

int lpma_set_profiling(lpma_t *h, int on) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    h->pl.profiling = on != 0;
    return LPMA_OK;
}

int lpma_set_graph(lpma_t *h, int on) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    h->pl.graph_on = on != 0;
    return LPMA_OK;
}

static int slots_log2_of(uint32_t slots_per_lane, int *out) {
    if (slots_per_lane == 0) {
        *out = -1;
        return LPMA_OK;
    }
    if (slots_per_lane > 1024 || (slots_per_lane & (slots_per_lane - 1))) return LPMA_ERR_INVALID_ARGUMENT;
    *out = __builtin_ctz(slots_per_lane);
    return LPMA_OK;
}

int lpma_set_launch(lpma_t *h, uint32_t seg_slots_per_lane) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    return slots_log2_of(seg_slots_per_lane, &h->pl.seg_slots_log2);
}

int lpma_reserve(lpma_t *h, uint64_t max_batch, void *stream) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    h->cur_stream = (hipStream_t)stream;
    lpma *one[1] = {h};
    return reserve_pipeline(h->pl, one, 1, max_batch, (hipStream_t)stream);
}

// Rebuilds up to m live keys without device allocation.  A rebuild writes
// new arrays and keeps the old ones as spares, so the handle holds a spare
// of the current height (for root rebuilds) and two of every height above
// it up to the one m needs (one to grow into, one for root rebuilds there);
// the merged scratch of large windows is sized for m keys.
int reserve_growth(lpma *h, u64 m, hipStream_t st) {
    const u32 Ht = fit_height(h, m);
    if (!height_ok(h, Ht)) return LPMA_ERR_TOO_LARGE;
    std::vector<u32> added;
    int r = LPMA_OK;
    for (u32 H = h->st.H; H <= std::max(Ht, h->st.H) && r == LPMA_OK; ++H) {
        const size_t want = H == h->st.H ? 1 : 2;
        while (r == LPMA_OK && h->spares[H].size() < want) {
            const u64 nseg = 1ull << H;
            Geometry g;
            if ((r = ensure(h, g.keys, nseg << h->log2S)) || (r = ensure(h, g.cnt, nseg)) ||
                (r = ensure(h, g.pyr, nseg)) || (r = ensure(h, g.sep, nseg)) || (r = ensure(h, g.owner, 2 * nseg))) {
                release_geometry(h, g);
                break;
            }
            h->spares[H].push_back(g);
            added.push_back(H);
        }
    }
    h->pl.cur_stream = st;
    if (!r) r = ensure(h->pl, h->pl.merged, m + 1);
    if (!r) {
        DevState top = h->st;
        top.H = std::max(Ht, h->st.H);
        r = ensure(h->pl, h->pl.gpre, rebuild_index_words(top));  // a rebuild's chunk index
    }
    if (r) {
        for (u32 H : added) {
            release_geometry(h, h->spares[H].back());
            h->spares[H].pop_back();
        }
        return r;
    }
    h->keep_spares = true;
    return LPMA_OK;
}

int lpma_reserve_capacity(lpma_t *h, uint64_t max_edges, void *stream) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    if (max_edges > (1ull << 40)) return LPMA_ERR_TOO_LARGE;
    DevGuard guard(h->device);
    h->cur_stream = (hipStream_t)stream;
    return reserve_growth(h, max_edges + h->V, (hipStream_t)stream);
}

int lpma_get_profile(lpma_t *h, lpma_profile_t *out) {
    if (!h || !out) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    memset(out, 0, sizeof(*out));
    add_profile(h->pl, out);
    return LPMA_OK;
}

// ------------------------------------------------------------------ groups

int lpma_group_create(lpma_group_t **out, lpma_t *const *members, uint32_t n) {
    if (!out) return LPMA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    if (!members || n == 0) return LPMA_ERR_INVALID_ARGUMENT;
    for (u32 i = 0; i < n; ++i) {
        if (!members[i] || members[i]->device != members[0]->device) return LPMA_ERR_INVALID_ARGUMENT;
        for (u32 j = 0; j < i; ++j)
            if (members[j] == members[i]) return LPMA_ERR_INVALID_ARGUMENT;
    }
    lpma_group *g = new (std::nothrow) lpma_group();
    if (!g) return LPMA_ERR_OUT_OF_MEMORY;
    g->device = members[0]->device;
    g->members.assign(members, members + n);
    // partitions: same segment size, and the ids must fit above the widest
    // member's sort key (2 * vbits + 1 bits)
    std::map<u32, std::vector<u32>> by_s;
    for (u32 i = 0; i < n; ++i) by_s[members[i]->log2S].push_back(i);
    for (auto &kv : by_s) {
        std::vector<u32> cur;
        u32 vmax = 1;
        for (u32 i : kv.second) {
            u32 v2 = std::max(vmax, members[i]->vbits);
            if (!cur.empty() && 2 * v2 + 1 + bits_for(cur.size() + 1) > 64) {
                g->parts.push_back(cur);
                cur.clear();
                v2 = members[i]->vbits;
            }
            cur.push_back(i);
            vmax = v2;
        }
        if (!cur.empty()) g->parts.push_back(cur);
    }
    DevGuard guard(g->device);
    for (size_t k = 0; k < g->parts.size(); ++k) {
        Pipeline *pl = new (std::nothrow) Pipeline();
        int r = pl ? pipeline_init(*pl, g->device, nullptr) : LPMA_ERR_OUT_OF_MEMORY;
        if (pl) g->pls.push_back(pl);
        if (r != LPMA_OK) {
            lpma_group_destroy(g);
            return r;
        }
    }
    if (hipStreamSynchronize(nullptr) != hipSuccess) {
        (void)hipGetLastError();
        lpma_group_destroy(g);
        return LPMA_ERR_DEVICE;
    }
    *out = g;
    return LPMA_OK;
}

int lpma_group_destroy(lpma_group_t *g) {
    if (!g) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(g->device);
    (void)hipDeviceSynchronize();
    for (Pipeline *pl : g->pls) {
        pipeline_free(*pl);
        delete pl;
    }
    delete g;
    return LPMA_OK;
}

int lpma_group_update(lpma_group_t *g, const uint64_t *const *d_ins, const uint64_t *n_ins,
                      const uint64_t *const *d_del, const uint64_t *n_del, void *stream) {
    if (!g) return LPMA_ERR_INVALID_ARGUMENT;
    const u32 n = (u32)g->members.size();
    for (u32 i = 0; i < n; ++i) {
        u64 ni = n_ins ? n_ins[i] : 0, nd = n_del ? n_del[i] : 0;
        if ((ni && (!d_ins || !d_ins[i])) || (nd && (!d_del || !d_del[i]))) return LPMA_ERR_INVALID_ARGUMENT;
    }
    DevGuard guard(g->device);
    g->calls++;
    std::vector<Member> mem;
    std::vector<Req> reqs(n);
    for (u32 i = 0; i < n; ++i) {
        Req &q = reqs[i];
        q.nins = n_ins ? n_ins[i] : 0;
        q.ndel = n_del ? n_del[i] : 0;
        q.ins = q.nins ? d_ins[i] : nullptr;
        q.del = q.ndel ? d_del[i] : nullptr;
    }
    for (size_t k = 0; k < g->parts.size(); ++k) {
        mem.clear();
        for (u32 i : g->parts[k]) mem.push_back(Member{g->members[i], &reqs[i], 1});
        CHECKED(run_pipeline(*g->pls[k], mem.data(), (u32)mem.size(), (hipStream_t)stream));
    }
    return LPMA_OK;
}

uint32_t lpma_group_partitions(const lpma_group_t *g) { return g ? (uint32_t)g->parts.size() : 0; }

int lpma_group_get_stats(lpma_group_t *g, lpma_stats_t *out) {
    if (!g || !out) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(g->device);
    memset(out, 0, sizeof(*out));
    for (size_t k = 0; k < g->pls.size(); ++k) {
        Pipeline &pl = *g->pls[k];
        CHECKED(read_stats(pl, g->members[g->parts[k][0]]->log2S, out));
        out->effective_updates += pl.effective;
        out->rebuilds += pl.rebuilds;
        out->rebuild_slots += pl.rebuild_slots;
    }
    out->updates_calls = g->calls;
    return LPMA_OK;
}

int lpma_group_set_profiling(lpma_group_t *g, int on) {
    if (!g) return LPMA_ERR_INVALID_ARGUMENT;
    for (Pipeline *pl : g->pls) pl->profiling = on != 0;
    return LPMA_OK;
}

int lpma_group_set_graph(lpma_group_t *g, int on) {
    if (!g) return LPMA_ERR_INVALID_ARGUMENT;
    for (Pipeline *pl : g->pls) pl->graph_on = on != 0;
    return LPMA_OK;
}

int lpma_group_set_launch(lpma_group_t *g, uint32_t seg_slots_per_lane) {
    if (!g) return LPMA_ERR_INVALID_ARGUMENT;
    int v;
    CHECKED(slots_log2_of(seg_slots_per_lane, &v));
    for (Pipeline *pl : g->pls) pl->seg_slots_log2 = v;
    return LPMA_OK;
}

int lpma_group_reserve(lpma_group_t *g, uint64_t max_batch, void *stream) {
    if (!g) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(g->device);
    std::vector<lpma *> ms;
    for (size_t k = 0; k < g->parts.size(); ++k) {
        ms.clear();
        for (u32 i : g->parts[k]) ms.push_back(g->members[i]);
        CHECKED(reserve_pipeline(*g->pls[k], ms.data(), (u32)ms.size(), max_batch, (hipStream_t)stream));
    }
    return LPMA_OK;
}

int lpma_group_get_profile(lpma_group_t *g, lpma_profile_t *out) {
    if (!g || !out) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(g->device);
    memset(out, 0, sizeof(*out));
    for (Pipeline *pl : g->pls) add_profile(*pl, out);
    return LPMA_OK;
}

int lpma_gen_rmat_device(uint32_t scale, uint64_t n, double a, double b, double c, uint64_t seed, uint64_t *d_out,
                         void *stream) {
    if ((!d_out && n) || scale < 1 || scale > 31) return LPMA_ERR_INVALID_ARGUMENT;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return LPMA_ERR_INVALID_ARGUMENT;
    HIPCHECK(gen_rmat_device(scale, n, a, b, c, seed, d_out, (hipStream_t)stream));
    return LPMA_OK;
}

int lpma_set_pool_budget(uint64_t bytes) {
    std::lock_guard<std::mutex> g(g_pools_mu);
    g_budget = bytes;
    return LPMA_OK;
}

int lpma_inject_alloc_failure(uint64_t nth, uint64_t *remaining) {
    long long prev =
        g_fail_countdown.exchange(nth > (uint64_t)LLONG_MAX ? LLONG_MAX : (long long)nth, std::memory_order_relaxed);
    if (remaining) *remaining = prev > 0 ? (uint64_t)prev : 0;  // -1: it fired
    return LPMA_OK;
}

int lpma_pool_trim(void) {
    std::vector<Pool *> ps;
    {
        std::lock_guard<std::mutex> g(g_pools_mu);
        for (auto &kv : g_pools) ps.push_back(kv.second);
    }
    for (Pool *p : ps) {
        std::lock_guard<std::mutex> g(p->mu);
        p->trim_locked();
    }
    return LPMA_OK;
}

uint64_t lpma_pool_reserved_bytes(void) {
    u64 t = 0;
    std::lock_guard<std::mutex> g(g_pools_mu);
    for (auto &kv : g_pools) {
        std::lock_guard<std::mutex> g2(kv.second->mu);
        t += kv.second->reserved;
    }
    return t;
}

}  // extern "C"
