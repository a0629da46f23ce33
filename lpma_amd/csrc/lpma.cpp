// lpma.cpp -- host side of the C ABI: handle, device pool allocator, and
// the update orchestrator that replaces update_rpma (rpma.cuh:1549-1718).
//
// One update call = one stream-ordered kernel sequence with exactly one
// host synchronisation (after the density check, to learn whether the batch
// is a plain rebalance, a root rebalance or a resize), instead of the
// reference's per-level host loop with a device sync after every step.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
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

Pool &pool_for(int dev) {
    std::lock_guard<std::mutex> g(g_pools_mu);
    auto it = g_pools.find(dev);
    if (it != g_pools.end()) return *it->second;
    Pool *p = new Pool();
    g_pools[dev] = p;
    return *p;
}

int pool_alloc(int dev, void **out, size_t bytes) {
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

// ------------------------------------------------------------------ handle

struct lpma {
    int device = 0;
    hipStream_t cur_stream = nullptr;  // stream of the call in progress
    lpma_config_t cfg;
    u32 V = 0, vbits = 1, log2S = 5;
    u64 live = 0;
    u64 version = 0;
    DevState st;
    Density dn;

    DBuf<u64> keys;
    DBuf<u32> cnt, pyr, row_off;
    DBuf<u64> owner;
    DBuf<u64> sep;

    // batch scratch
    DBuf<u64> pk0, pk1, ukey, ckey, merged;
    DBuf<u32> sort_scr, useg, cseg, rlevel, gpre, scan_scr, lchunk, lhdr;
    DBuf<u64> lmbase;
    // analytics scratch
    DBuf<u32> aq0, aq1, alen, acnt;
    DBuf<float> af0, af1, af2, ared;
    DBuf<i32> pd, part;
    DBuf<Work> work[NUM_CLASSES];
    Status *d_status = nullptr;
    Status *h_status = nullptr;
    u64 *d_stats = nullptr;
    u32 *d_err = nullptr;

    // host counters
    u64 calls = 0, effective = 0, rebuilds = 0, rebuild_slots = 0;

    // profiling
    bool profiling = false;
    struct Ev {
        hipEvent_t a, b;
        int phase;
    };
    std::vector<Ev> events;
    std::vector<hipEvent_t> ev_pool;
    double phase_ms[9] = {0};
    u64 launches_wave = 0, launches_block = 0, launches_seg = 0;
};

namespace {

enum Phase { PH_SORT, PH_LOCATE, PH_ELECT, PH_WAVE, PH_BLOCK, PH_LARGE, PH_REBUILD, PH_PYR, PH_SEG };

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
int ensure(lpma *h, DBuf<T> &b, u64 n) {
    if (b.n >= n && b.p) return LPMA_OK;
    u64 want = std::max<u64>(n, b.n + b.n / 2);
    if (want < 64) want = 64;
    void *p = nullptr;
    int r = pool_alloc(h->device, &p, want * sizeof(T));
    if (r != LPMA_OK) {
        // retry exactly
        r = pool_alloc(h->device, &p, n * sizeof(T));
        if (r != LPMA_OK) return r;
        want = n;
    }
    if (b.p) {
        // queued kernels may still read the old block; it must not reach
        // another stream through the shared pool before they finish
        (void)hipStreamSynchronize(h->cur_stream);
        pool_free(h->device, b.p);
    }
    b.p = (T *)p;
    b.n = want;
    return LPMA_OK;
}

template <typename T>
void release(lpma *h, DBuf<T> &b) {
    pool_free(h->device, b.p);
    b.p = nullptr;
    b.n = 0;
}

hipEvent_t take_event(lpma *h) {
    if (!h->ev_pool.empty()) {
        hipEvent_t e = h->ev_pool.back();
        h->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
}

struct PhaseTimer {
    lpma *h;
    hipStream_t st;
    int phase;
    hipEvent_t a = nullptr;
    PhaseTimer(lpma *h_, hipStream_t s, int p) : h(h_), st(s), phase(p) {
        if (h->profiling) {
            a = take_event(h);
            (void)hipEventRecord(a, st);
        }
    }
    ~PhaseTimer() {
        if (a) {
            hipEvent_t b = take_event(h);
            (void)hipEventRecord(b, st);
            h->events.push_back({a, b, phase});
        }
    }
};

void collect_events(lpma *h) {
    for (auto &e : h->events) {
        (void)hipEventSynchronize(e.b);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e.a, e.b);
        h->phase_ms[e.phase] += ms;
        h->ev_pool.push_back(e.a);
        h->ev_pool.push_back(e.b);
    }
    h->events.clear();
}

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

struct Geometry {
    DBuf<u64> keys, sep, owner;
    DBuf<u32> cnt, pyr;
};

int alloc_geometry(lpma *h, u32 H, Geometry &g, hipStream_t st) {
    u64 nseg = 1ull << H;
    u64 cap = nseg << h->log2S;
    int r;
    if ((r = ensure(h, g.keys, cap)) || (r = ensure(h, g.cnt, nseg)) || (r = ensure(h, g.pyr, nseg)) ||
        (r = ensure(h, g.sep, nseg)) || (r = ensure(h, g.owner, 2 * nseg))) {
        release(h, g.keys); release(h, g.cnt); release(h, g.pyr); release(h, g.sep); release(h, g.owner);
        return r;
    }
    if (fill_u64(g.keys.p, cap, KEY_NONE, st) != hipSuccess || fill_u32(g.cnt.p, nseg, 0, st) != hipSuccess ||
        fill_u32(g.pyr.p, nseg, 0, st) != hipSuccess || fill_u64(g.sep.p, nseg, 0, st) != hipSuccess ||
        fill_u64(g.owner.p, 2 * nseg, 0, st) != hipSuccess) {
        (void)hipGetLastError();
        release(h, g.keys); release(h, g.cnt); release(h, g.pyr); release(h, g.sep); release(h, g.owner);
        return LPMA_ERR_DEVICE;
    }
    return LPMA_OK;
}

void install_geometry(lpma *h, u32 H, Geometry &g) {
    release(h, h->keys); release(h, h->cnt); release(h, h->pyr); release(h, h->sep); release(h, h->owner);
    h->keys = g.keys; h->cnt = g.cnt; h->pyr = g.pyr; h->sep = g.sep; h->owner = g.owner;
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

BatchBufs batch_bufs(lpma *h) {
    BatchBufs b;
    b.ukey = h->ukey.p;
    b.useg = h->useg.p;
    b.ckey = h->ckey.p;
    b.cseg = h->cseg.p;
    b.pd = h->pd.p;
    b.rlevel = h->rlevel.p;
    b.part = h->part.p;
    for (int c = 0; c < NUM_CLASSES; ++c) b.work[c] = h->work[c].p;
    b.lchunk = h->lchunk.p;
    b.lmbase = h->lmbase.p;
    b.lhdr = h->lhdr.p;
    b.status = h->d_status;
    b.stats = h->d_stats;
    return b;
}

int ensure_batch(lpma *h, u64 n) {
    int r;
    u64 nb = locate_blocks(n);
    if ((r = ensure(h, h->pk0, n)) || (r = ensure(h, h->pk1, n)) || (r = ensure(h, h->sort_scr, sort_scratch_words(n))) ||
        (r = ensure(h, h->ukey, n)) || (r = ensure(h, h->useg, n)) || (r = ensure(h, h->ckey, n)) ||
        (r = ensure(h, h->cseg, n)) || (r = ensure(h, h->pd, n + 1)) || (r = ensure(h, h->rlevel, n)) ||
        (r = ensure(h, h->part, 2 * nb + 2)))
        return r;
    u64 runs = std::min<u64>(n, 1ull << h->st.H);
    for (int c = 0; c < NUM_CLASSES; ++c)
        if ((r = ensure(h, h->work[c], runs))) return r;
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

// Spread m sorted keys from `merged` over a fresh geometry of height H and
// install it (create and resize).
int respread_into_new(lpma *h, u32 H, u64 m, hipStream_t st) {
    Geometry g;
    CHECKED(alloc_geometry(h, H, g, st));
    DevState d2 = state_with(h, H, g);
    HIPCHECK(rebuild_spread(d2, h->merged.p, m, st));
    HIPCHECK(rebuild_pyramid(d2, true, st));
    // old arrays may still be read by queued kernels: wait before they
    // return to the pool
    HIPCHECK(hipStreamSynchronize(st));
    install_geometry(h, H, g);
    return LPMA_OK;
}

int run_update(lpma *h, const u64 *a, u64 na, u32 opa, const u64 *b, u64 nb, u32 opb, hipStream_t st) {
    const u64 n = na + nb;
    h->cur_stream = st;
    h->calls++;
    if (n == 0) {
        h->version++;
        return LPMA_OK;
    }
    if (n > 0x7FFFFFF0ull) return LPMA_ERR_TOO_LARGE;
    CHECKED(ensure_batch(h, n));
    if (++h->st.epoch == 0) {
        // claims are tagged with the epoch: after a wrap, clear the old tags
        h->st.epoch = 1;
        HIPCHECK(fill_u64(h->owner.p, 2 * (1ull << h->st.H), 0, st));
    }
    BatchBufs bb = batch_bufs(h);
    // bit 0 of a packed key is the operation: sorted only when the batch mixes them
    const u32 lo_bit = (na && nb && opa != opb) ? 0 : 1;
    const u32 hi_bit = 2 * h->vbits + 1;
    u64 *pk;
    {
        PhaseTimer t(h, st, PH_SORT);
        HIPCHECK(batch_begin(h->d_status, h->st, st));
        pk = pack_and_sort(a, na, opa, b, nb, opb, h->V, h->vbits, lo_bit, hi_bit, h->pk0.p, h->pk1.p,
                           h->sort_scr.p, &h->d_status->flags, st);
        HIPCHECK(hipGetLastError());
    }
    {
        PhaseTimer t(h, st, PH_LOCATE);
        HIPCHECK(locate_and_compact(h->st, pk, n, bb, st));
        HIPCHECK(density_check(h->st, h->dn, n, bb, st));
    }
    {
        // window election before the single host synchronisation, so the
        // host learns the exact work of every rebalance class and sizes each
        // grid to it (idle workgroups still cost dispatch time)
        PhaseTimer t(h, st, PH_ELECT);
        HIPCHECK(elect_windows(h->st, n, bb, st));
    }
    HIPCHECK(hipMemcpyAsync(h->h_status, h->d_status, sizeof(Status), hipMemcpyDeviceToHost, st));
    HIPCHECK(hipStreamSynchronize(st));
    const Status s = *h->h_status;
    if (s.flags & FLAG_INVALID_KEY) return LPMA_ERR_INVALID_KEY;
    const u64 n_eff = s.n_eff;
    if (n_eff == 0) {
        h->version++;
        return LPMA_OK;
    }
    const u64 M = (u64)((i64)h->live + s.delta);

    if (s.flags & (FLAG_ROOT | FLAG_RESIZE)) {
        // root rebalance or resize: every live key and every update is
        // merged into scratch and spread again (rebuild_rpma, rpma.cuh:1417-1472)
        u32 newH = (s.flags & FLAG_RESIZE) ? fit_height(h, M) : h->st.H;
        if (!height_ok(h, newH)) return LPMA_ERR_TOO_LARGE;
        u64 nseg = 1ull << h->st.H;
        CHECKED(ensure(h, h->merged, M + 1));
        CHECKED(ensure(h, h->gpre, nseg + 1));
        CHECKED(ensure(h, h->scan_scr, scan_scratch_words(nseg) + 8));
        PhaseTimer t(h, st, PH_REBUILD);
        HIPCHECK(rebuild_push(h->st, n_eff, bb, h->merged.p, M, h->gpre.p, h->scan_scr.p, st));
        if (newH != h->st.H) {
            CHECKED(respread_into_new(h, newH, M, st));
        } else {
            HIPCHECK(rebuild_spread(h->st, h->merged.p, M, st));
            HIPCHECK(rebuild_pyramid(h->st, true, st));
        }
        h->rebuilds++;
        h->rebuild_slots += (u64)1 << (h->st.H + h->log2S);
    } else {
        if (s.nwork[CLASS_SEG]) {
            PhaseTimer t(h, st, PH_SEG);
            HIPCHECK(rebalance_seg(h->st, bb, s.nwork[CLASS_SEG], st));
            h->launches_seg++;
        }
        if (s.nwork[CLASS_WAVE]) {
            PhaseTimer t(h, st, PH_WAVE);
            HIPCHECK(rebalance_lds(h->st, n_eff, bb, CLASS_WAVE, std::min<u32>(s.nwork[CLASS_WAVE], 65535), st));
            h->launches_wave++;
        }
        if (s.nwork[CLASS_BLOCK]) {
            PhaseTimer t(h, st, PH_BLOCK);
            HIPCHECK(rebalance_lds(h->st, n_eff, bb, CLASS_BLOCK, std::min<u32>(s.nwork[CLASS_BLOCK], 65535), st));
            h->launches_block++;
        }
        if (s.nwork[CLASS_LARGE]) {
            CHECKED(ensure(h, h->merged, M + 1));
            CHECKED(ensure(h, h->lchunk, (u64)s.nwork[CLASS_LARGE] + 1));
            CHECKED(ensure(h, h->lmbase, s.nwork[CLASS_LARGE]));
            CHECKED(ensure(h, h->lhdr, 4));
            bb = batch_bufs(h);
            PhaseTimer t(h, st, PH_LARGE);
            HIPCHECK(rebalance_large(h->st, bb, h->merged.p, 1024, st));
        }
        {
            PhaseTimer t(h, st, PH_PYR);
            HIPCHECK(rebuild_pyramid(h->st, false, st));
        }
    }
    h->live = M;
    h->effective += n_eff;
    h->version++;
    return LPMA_OK;
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
    if (hipMalloc(&h->d_status, sizeof(Status)) != hipSuccess ||
        hipHostMalloc((void **)&h->h_status, sizeof(Status), hipHostMallocDefault) != hipSuccess ||
        hipMalloc(&h->d_stats, ST_ROWS * ST_NUM * sizeof(u64)) != hipSuccess || hipMalloc(&h->d_err, 16) != hipSuccess) {
        (void)hipGetLastError();
        lpma_destroy(h);
        return LPMA_ERR_OUT_OF_MEMORY;
    }
    if (hipMemsetAsync(h->d_stats, 0, ST_ROWS * ST_NUM * sizeof(u64), st) != hipSuccess ||
        hipMemsetAsync(h->d_status, 0, sizeof(Status), st) != hipSuccess) {
        (void)hipGetLastError();
        lpma_destroy(h);
        return LPMA_ERR_DEVICE;
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
    if ((r = ensure(h, h->merged, (u64)num_vertices + 1)) != LPMA_OK) {
        lpma_destroy(h);
        return r;
    }
    if (make_markers(h->merged.p, num_vertices, st) != hipSuccess) {
        (void)hipGetLastError();
        lpma_destroy(h);
        return LPMA_ERR_DEVICE;
    }
    h->st.H = H;
    compute_density(h);
    if ((r = respread_into_new(h, H, num_vertices, st)) != LPMA_OK) {
        lpma_destroy(h);
        return r;
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
    collect_events(h);
    for (auto e : h->ev_pool) (void)hipEventDestroy(e);
    release(h, h->keys); release(h, h->cnt); release(h, h->pyr); release(h, h->owner);
    release(h, h->row_off); release(h, h->sep);
    release(h, h->pk0); release(h, h->pk1); release(h, h->ukey); release(h, h->ckey); release(h, h->merged);
    release(h, h->sort_scr); release(h, h->useg); release(h, h->cseg); release(h, h->rlevel);
    release(h, h->gpre); release(h, h->scan_scr); release(h, h->pd); release(h, h->part);
    release(h, h->lchunk); release(h, h->lhdr); release(h, h->lmbase);
    release(h, h->aq0); release(h, h->aq1); release(h, h->alen); release(h, h->acnt);
    release(h, h->af0); release(h, h->af1); release(h, h->af2); release(h, h->ared);
    for (int c = 0; c < NUM_CLASSES; ++c) release(h, h->work[c]);
    if (h->d_status) (void)hipFree(h->d_status);
    if (h->h_status) (void)hipHostFree(h->h_status);
    if (h->d_stats) (void)hipFree(h->d_stats);
    if (h->d_err) (void)hipFree(h->d_err);
    delete h;
    return LPMA_OK;
}

int lpma_update(lpma_t *h, const uint64_t *d_keys, uint64_t n, int op, void *stream) {
    if (!h || (!d_keys && n) || (op != LPMA_OP_INSERT && op != LPMA_OP_DELETE)) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    return run_update(h, d_keys, n, (u32)op, nullptr, 0, 0, (hipStream_t)stream);
}

int lpma_update_mixed(lpma_t *h, const uint64_t *d_ins, uint64_t n_ins, const uint64_t *d_del, uint64_t n_del,
                      void *stream) {
    if (!h || (!d_ins && n_ins) || (!d_del && n_del)) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    return run_update(h, d_del, n_del, 0, d_ins, n_ins, 1, (hipStream_t)stream);
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
    if (d_keys)
        HIPCHECK(hipMemcpyAsync(d_keys, h->st.keys, ((u64)1 << (h->st.H + h->log2S)) * sizeof(u64),
                                hipMemcpyDeviceToDevice, st));
    if (d_row_offsets)
        HIPCHECK(hipMemcpyAsync(d_row_offsets, h->st.row_off, ((u64)h->V + 1) * sizeof(u32), hipMemcpyDeviceToDevice,
                                st));
    return LPMA_OK;
}

static int export_common(lpma_t *h, u64 *edges, u32 *row_std, u64 *all, u32 *csr_row, u32 *csr_col,
                         void *stream) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    hipStream_t st = (hipStream_t)stream;
    h->cur_stream = st;
    u64 nseg = 1ull << h->st.H;
    CHECKED(ensure(h, h->gpre, nseg + 1));
    CHECKED(ensure(h, h->scan_scr, scan_scratch_words(nseg) + 8));
    HIPCHECK(export_compacted(h->st, h->gpre.p, h->scan_scr.p, edges, row_std, all, csr_row, csr_col, st));
    return LPMA_OK;
}

int lpma_export_compacted(lpma_t *h, uint64_t *d_edge_keys, uint32_t *d_row_offsets, void *stream) {
    return export_common(h, d_edge_keys, d_row_offsets, nullptr, nullptr, nullptr, stream);
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

int lpma_bfs(lpma_t *h, uint32_t src, uint32_t *d_levels, uint32_t *depth, uint64_t *edges_scanned, void *stream) {
    if (!h || !d_levels) return LPMA_ERR_INVALID_ARGUMENT;
    if (src >= h->V) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    hipStream_t st = (hipStream_t)stream;
    h->cur_stream = st;
    const u64 V = h->V;
    CHECKED(ensure(h, h->aq0, V + 1));
    CHECKED(ensure(h, h->aq1, V + 1));
    CHECKED(ensure(h, h->alen, V + 2));
    CHECKED(ensure(h, h->scan_scr, scan_scratch_words(V + 1) + 8));
    CHECKED(ensure(h, h->acnt, 4));
    HIPCHECK(bfs_init(d_levels, h->V, src, h->aq0.p, h->acnt.p, st));
    u32 *q = h->aq0.p, *qn_buf = h->aq1.p;
    u32 qn = 1, level = 0;
    u64 scanned = 0;
    u32 cnt[2];
    while (qn) {
        HIPCHECK(bfs_level(h->st, q, qn, h->alen.p, h->scan_scr.p, d_levels, level, qn_buf, h->acnt.p, st));
        HIPCHECK(hipMemcpyAsync(cnt, h->acnt.p, sizeof(cnt), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        HIPCHECK(fill_u32(h->acnt.p, 2, 0, st));
        scanned += cnt[1];
        qn = cnt[0];
        std::swap(q, qn_buf);
        ++level;
    }
    if (depth) *depth = level;
    if (edges_scanned) *edges_scanned = scanned;
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
    CHECKED(ensure(h, h->gpre, nseg + 1));
    CHECKED(ensure(h, h->scan_scr, scan_scratch_words(nseg) + 8));
    CHECKED(ensure(h, h->aq0, V));   // out-degrees
    CHECKED(ensure(h, h->af0, V));   // contributions
    CHECKED(ensure(h, h->af1, V));   // accumulator
    CHECKED(ensure(h, h->af2, V));   // second rank vector
    CHECKED(ensure(h, h->ared, 8));
    HIPCHECK(pr_outdeg(h->st, h->gpre.p, h->scan_scr.p, h->aq0.p, st));
    float *red[2] = {h->ared.p, h->ared.p + 4};
    float *vec[2] = {d_rank, h->af2.p};
    HIPCHECK(pr_init(vec[0], h->V, red[0], st));
    u32 it = 0;
    float delta = 0.f;
    int cur = 0;
    while (it < max_iter) {
        HIPCHECK(pr_iterate(h->st, vec[cur], h->aq0.p, h->af0.p, h->af1.p, vec[cur ^ 1], damping, red[cur],
                            red[cur ^ 1], st));
        cur ^= 1;
        ++it;
        float r[4];
        HIPCHECK(hipMemcpyAsync(r, red[cur], sizeof(r), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        delta = r[2];
        if (delta < tolerance) break;
    }
    HIPCHECK(pr_scale(vec[cur], h->V, red[cur], st));
    if (vec[cur] != d_rank) HIPCHECK(hipMemcpyAsync(d_rank, vec[cur], V * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIPCHECK(hipStreamSynchronize(st));
    if (iterations) *iterations = it;
    if (last_delta) *last_delta = delta;
    return LPMA_OK;
}

int lpma_get_stats(lpma_t *h, lpma_stats_t *out) {
    if (!h || !out) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    std::vector<u64> rows((size_t)ST_ROWS * ST_NUM);
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemcpy(rows.data(), h->d_stats, rows.size() * sizeof(u64), hipMemcpyDeviceToHost));
    u64 s[ST_NUM] = {0};
    for (u32 r = 0; r < ST_ROWS; ++r)
        for (u32 i = 0; i < ST_NUM; ++i) s[i] += rows[(size_t)r * ST_NUM + i];
    memset(out, 0, sizeof(*out));
    out->updates_calls = h->calls;
    out->effective_updates = h->effective;
    for (int l = 0; l < 32; ++l) {
        out->windows_per_level[l] = s[ST_WINDOWS + l];
        out->slots_per_level[l] = s[ST_WINDOWS + l] << (l + h->log2S);
    }
    out->rebuilds = h->rebuilds;
    out->rebuild_slots = h->rebuild_slots;
    out->marker_moves = s[ST_MARKERS];
    out->alg_bytes_wave = s[ST_ALG_BYTES_WAVE];
    out->alg_bytes_block = s[ST_ALG_BYTES_BLOCK];
    out->alg_bytes_large = s[ST_ALG_BYTES_LARGE];
    out->alg_bytes_seg = s[ST_ALG_BYTES_SEG];
    return LPMA_OK;
}

int lpma_reset_stats(lpma_t *h) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    HIPCHECK(hipDeviceSynchronize());
    HIPCHECK(hipMemset(h->d_stats, 0, ST_ROWS * ST_NUM * sizeof(u64)));
    h->calls = h->effective = h->rebuilds = h->rebuild_slots = 0;
    return LPMA_OK;
}

int lpma_set_profiling(lpma_t *h, int on) {
    if (!h) return LPMA_ERR_INVALID_ARGUMENT;
    h->profiling = on != 0;
    return LPMA_OK;
}

int lpma_get_profile(lpma_t *h, lpma_profile_t *out) {
    if (!h || !out) return LPMA_ERR_INVALID_ARGUMENT;
    DevGuard guard(h->device);
    collect_events(h);
    out->sort_ms = h->phase_ms[PH_SORT];
    out->locate_ms = h->phase_ms[PH_LOCATE];
    out->elect_ms = h->phase_ms[PH_ELECT];
    out->rebalance_wave_ms = h->phase_ms[PH_WAVE];
    out->rebalance_block_ms = h->phase_ms[PH_BLOCK];
    out->large_ms = h->phase_ms[PH_LARGE];
    out->rebuild_ms = h->phase_ms[PH_REBUILD];
    out->pyramid_ms = h->phase_ms[PH_PYR];
    out->rebalance_seg_ms = h->phase_ms[PH_SEG];
    out->launches_wave = h->launches_wave;
    out->launches_block = h->launches_block;
    out->launches_seg = h->launches_seg;
    for (double &v : h->phase_ms) v = 0;
    h->launches_wave = h->launches_block = h->launches_seg = 0;
    return LPMA_OK;
}

int lpma_set_pool_budget(uint64_t bytes) {
    std::lock_guard<std::mutex> g(g_pools_mu);
    g_budget = bytes;
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
