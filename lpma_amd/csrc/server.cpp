// server.cpp -- the served form: a request loop over one resident PMA.
//
// The reference has no serving layer: its drivers call update_rpma in a
// loop (test_update_rpma.cu:31-165) and query the structure in between
// (test.cu:122-182).  Here requests -- update batches, neighbor reads and
// breadth-first searches -- are submitted from any host thread and served
// with continuous batching, on top of the public C ABI only:
//
//   * one engine thread takes every update request queued when it comes
//     round and applies them as ONE pipeline run (lpma_update_sequence keeps
//     their order: for each key the last request decides), then every
//     queued neighbor read as ONE count / scan / gather pass over the live
//     structure, so reads see the version the updates just made;
//   * a BFS runs on a snapshot -- the newest one, copied at most a set age
//     ago -- on a worker thread with its own stream, beside the engine's
//     updates and reads (searches share snapshots);
//   * admission: the device memory a request needs (update and read
//     staging, BFS snapshots and scratch) is charged against a byte budget;
//     a request waits in its queue while the budget has no room for it, and
//     one larger than the whole budget fails with LPMA_ERR_OUT_OF_MEMORY.
//
// Every result records how many update requests had been applied when it
// was served, so a reader's answer can be checked against the sequence of
// updates up to that point.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/lpma.h"

namespace {

typedef uint64_t u64;
typedef uint32_t u32;
using clk = std::chrono::steady_clock;

struct Request {
    u64 id = 0;
    int kind = 0;
    clk::time_point t_submit, t_start, t_done;
    std::vector<u64> ins, del;
    std::vector<u32> verts;
    u32 root = 0;
    bool want_levels = false;
    // result
    int status = LPMA_OK;
    bool done = false;
    u64 applied = 0, version = 0;
    std::vector<u32> offsets, values;
    u32 depth = 0;
    u64 scanned = 0;
    struct Snap *snap = nullptr;
    u64 charged = 0;
};

struct Snap {
    lpma_snapshot_t *s = nullptr;
    u64 applied = 0;
    int refs = 0;
    u64 bytes = 0;
    clk::time_point taken;
};

// device buffer owned by the engine, grown on demand
template <typename T>
struct Dev {
    T *p = nullptr;
    u64 cap = 0;
};

// pinned host buffer
template <typename T>
struct Pin {
    T *p = nullptr;
    u64 cap = 0;
};

double us_between(clk::time_point a, clk::time_point b) {
    return std::chrono::duration<double, std::micro>(b - a).count();
}

}  // namespace

struct lpma_server {
    lpma_t *h = nullptr;
    int device = 0;
    u32 V = 0;
    lpma_server_config_t cfg;
    u32 max_reqs = 1;  // request order bits that fit above the sort keys
    hipStream_t wst = nullptr;

    std::mutex mu;
    std::condition_variable cv_engine, cv_done, cv_work;
    std::deque<Request *> q_upd, q_read, q_bfs, q_work;
    std::unordered_map<u64, Request *> reqs;
    u64 next_id = 1;
    bool stop = false, stop_workers = false;
    u64 applied = 0;  // update requests applied so far
    u64 in_use = 0, peak = 0;
    int busy = 0;
    bool bfs_blocked = false;  // the next search waits for memory held by running ones
    Snap *cur = nullptr;
    std::vector<Snap *> snaps;
    std::thread engine;
    std::vector<std::thread> workers;

    // engine staging
    Dev<u64> d_keys;
    Pin<u64> h_keys;
    Dev<u32> d_verts, d_len, d_off, d_out;
    Pin<u32> h_verts, h_len, h_off, h_out;

    lpma_server_stats_t st;
};

namespace {

template <typename T>
int grow_dev(lpma_server *s, Dev<T> &b, u64 n) {
    if (b.p && b.cap >= n) return LPMA_OK;
    u64 want = std::max<u64>(n, b.cap + b.cap / 2);
    if (want < 1024) want = 1024;
    u64 add = (want - b.cap) * sizeof(T);
    {
        std::lock_guard<std::mutex> g(s->mu);
        if (s->in_use + add > s->cfg.hbm_budget_bytes) return LPMA_ERR_OUT_OF_MEMORY;
        s->in_use += add;
        s->peak = std::max(s->peak, s->in_use);
    }
    if (b.p) {
        (void)hipStreamSynchronize(s->wst);
        (void)hipFree(b.p);
    }
    b.p = nullptr;
    if (hipMalloc((void **)&b.p, want * sizeof(T)) != hipSuccess) {
        (void)hipGetLastError();
        std::lock_guard<std::mutex> g(s->mu);
        s->in_use -= b.cap * sizeof(T) + add;
        b.p = nullptr;
        b.cap = 0;
        return LPMA_ERR_OUT_OF_MEMORY;
    }
    b.cap = want;
    return LPMA_OK;
}

template <typename T>
int grow_pin(Pin<T> &b, u64 n, hipStream_t st) {
    if (b.p && b.cap >= n) return LPMA_OK;
    u64 want = std::max<u64>(n, 2 * b.cap);
    if (want < 1024) want = 1024;
    if (b.p) {
        (void)hipStreamSynchronize(st);
        (void)hipHostFree(b.p);
        b.p = nullptr;
        b.cap = 0;
    }
    if (hipHostMalloc((void **)&b.p, want * sizeof(T), hipHostMallocDefault) != hipSuccess) {
        (void)hipGetLastError();
        b.p = nullptr;
        return LPMA_ERR_OUT_OF_MEMORY;
    }
    b.cap = want;
    return LPMA_OK;
}

// device scratch of one search: per vertex two queues, the counts, the first
// slots and the levels (32 B) and a visited bit; one run entry per 1024 slots
// (a snapshot holds at least 8 B per slot, so snapshot bytes / 2048 covers
// it); the scan
u64 bfs_bytes(u32 V, u64 snap_bytes) {
    return 32ull * ((u64)V + 2) + (u64)V / 8 + 64 + snap_bytes / 2048 + (1ull << 20);
}

void finish(lpma_server *s, Request *r, int status) {
    // caller holds s->mu
    r->status = status;
    r->t_done = clk::now();
    r->done = true;
    s->st.completed[r->kind]++;
}

// One pipeline run for the queued update requests (in order).  If the run
// fails on an invalid key, the requests are applied one by one instead so
// that only the offending ones fail.
void serve_updates(lpma_server *s, std::vector<Request *> &batch) {
    auto t0 = clk::now();
    u64 K = 0;
    for (Request *r : batch) K += r->ins.size() + r->del.size();
    int rc = grow_dev(s, s->d_keys, K);
    if (rc == LPMA_OK) rc = grow_pin(s->h_keys, K, s->wst);
    std::vector<const uint64_t *> pi(batch.size()), pd(batch.size());
    std::vector<uint64_t> ni(batch.size()), nd(batch.size());
    if (rc == LPMA_OK) {
        u64 off = 0;
        for (size_t i = 0; i < batch.size(); ++i) {
            Request *r = batch[i];
            if (!r->del.empty()) memcpy(s->h_keys.p + off, r->del.data(), r->del.size() * 8);
            pd[i] = s->d_keys.p + off;
            nd[i] = r->del.size();
            off += r->del.size();
            if (!r->ins.empty()) memcpy(s->h_keys.p + off, r->ins.data(), r->ins.size() * 8);
            pi[i] = s->d_keys.p + off;
            ni[i] = r->ins.size();
            off += r->ins.size();
        }
        if (K && hipMemcpyAsync(s->d_keys.p, s->h_keys.p, K * 8, hipMemcpyHostToDevice, s->wst) != hipSuccess) {
            (void)hipGetLastError();
            rc = LPMA_ERR_DEVICE;
        }
    }
    std::vector<int> status(batch.size(), rc);
    auto t1 = clk::now(), t2 = t1;
    if (rc == LPMA_OK) {
        rc = lpma_update_sequence(s->h, pi.data(), ni.data(), pd.data(), nd.data(), (u32)batch.size(), s->wst);
        t2 = clk::now();
        if (rc == LPMA_ERR_INVALID_KEY && batch.size() > 1) {
            for (size_t i = 0; i < batch.size(); ++i)
                status[i] = lpma_update_mixed(s->h, pi[i], ni[i], pd[i], nd[i], s->wst);
        } else {
            std::fill(status.begin(), status.end(), rc);
        }
        if (hipStreamSynchronize(s->wst) != hipSuccess) {
            (void)hipGetLastError();
            std::fill(status.begin(), status.end(), (int)LPMA_ERR_DEVICE);
        }
    }
    auto t3 = clk::now();
    lpma_view_t v;
    lpma_get_view(s->h, &v);
    std::lock_guard<std::mutex> g(s->mu);
    s->st.stage_us += us_between(t0, t1);
    s->st.pipeline_us += us_between(t1, t2);
    s->st.finish_us += us_between(t2, t3);
    s->st.update_runs++;
    for (size_t i = 0; i < batch.size(); ++i) {
        Request *r = batch[i];
        if (status[i] == LPMA_OK) s->applied++;
        r->applied = s->applied;
        r->version = v.version;
        finish(s, r, status[i]);
    }
    s->cv_done.notify_all();
}

// One count / scan / gather pass for the queued neighbor reads.
void serve_reads(lpma_server *s, std::vector<Request *> &batch) {
    u64 N = 0;
    for (Request *r : batch) N += r->verts.size();
    int rc = LPMA_OK;
    if ((rc = grow_dev(s, s->d_verts, N + 1)) || (rc = grow_dev(s, s->d_len, N + 1)) ||
        (rc = grow_dev(s, s->d_off, N + 1)) || (rc = grow_pin(s->h_verts, N + 1, s->wst)) ||
        (rc = grow_pin(s->h_len, N + 1, s->wst)) || (rc = grow_pin(s->h_off, N + 1, s->wst))) {
    }
    u64 total = 0;
    if (rc == LPMA_OK) {
        u64 off = 0;
        for (Request *r : batch) {
            if (!r->verts.empty()) memcpy(s->h_verts.p + off, r->verts.data(), r->verts.size() * 4);
            off += r->verts.size();
        }
        if (hipMemcpyAsync(s->d_verts.p, s->h_verts.p, N * 4, hipMemcpyHostToDevice, s->wst) != hipSuccess) rc =
            LPMA_ERR_DEVICE;
        if (rc == LPMA_OK) rc = lpma_neighbor_count(s->h, s->d_verts.p, N, s->d_len.p, s->wst);
        if (rc == LPMA_OK &&
            (hipMemcpyAsync(s->h_len.p, s->d_len.p, N * 4, hipMemcpyDeviceToHost, s->wst) != hipSuccess ||
             hipStreamSynchronize(s->wst) != hipSuccess))
            rc = LPMA_ERR_DEVICE;
        if (rc == LPMA_OK) {
            for (u64 i = 0; i < N; ++i) {
                s->h_off.p[i] = (u32)total;
                total += s->h_len.p[i];
            }
            s->h_off.p[N] = (u32)total;
            if (total >= (1ull << 32)) rc = LPMA_ERR_TOO_LARGE;
        }
        if (rc == LPMA_OK && ((rc = grow_dev(s, s->d_out, total + 1)) == LPMA_OK) &&
            ((rc = grow_pin(s->h_out, total + 1, s->wst)) == LPMA_OK)) {
            if (hipMemcpyAsync(s->d_off.p, s->h_off.p, (N + 1) * 4, hipMemcpyHostToDevice, s->wst) != hipSuccess)
                rc = LPMA_ERR_DEVICE;
            if (rc == LPMA_OK) rc = lpma_neighbor_gather(s->h, s->d_verts.p, N, s->d_off.p, s->d_out.p, s->wst);
            if (rc == LPMA_OK &&
                (hipMemcpyAsync(s->h_out.p, s->d_out.p, total * 4, hipMemcpyDeviceToHost, s->wst) != hipSuccess ||
                 hipStreamSynchronize(s->wst) != hipSuccess))
                rc = LPMA_ERR_DEVICE;
        }
    }
    if (rc == LPMA_ERR_DEVICE) (void)hipGetLastError();
    lpma_view_t v;
    lpma_get_view(s->h, &v);
    // results: per request its own offsets (from 0) and neighbor ids
    u64 off = 0;
    for (Request *r : batch) {
        const u64 k = r->verts.size();
        if (rc == LPMA_OK) {
            const u32 base = s->h_off.p[off];
            r->offsets.resize(k + 1);
            for (u64 i = 0; i <= k; ++i) r->offsets[i] = s->h_off.p[off + i] - base;
            r->values.assign(s->h_out.p + base, s->h_out.p + base + r->offsets[k]);
        }
        off += k;
    }
    std::lock_guard<std::mutex> g(s->mu);
    s->st.read_runs++;
    for (Request *r : batch) {
        r->applied = s->applied;
        r->version = v.version;
        finish(s, r, rc);
    }
    s->cv_done.notify_all();
}

void destroy_unused_snaps(lpma_server *s) {
    // caller holds s->mu; a snapshot goes once no search uses it and a newer
    // version has one (or the server stops)
    for (size_t i = 0; i < s->snaps.size();) {
        Snap *p = s->snaps[i];
        if (p->refs == 0 && (p != s->cur || s->stop)) {
            lpma_snapshot_destroy(p->s);
            s->in_use -= p->bytes;
            if (p == s->cur) s->cur = nullptr;
            delete p;
            s->snaps.erase(s->snaps.begin() + i);
        } else {
            ++i;
        }
    }
}

// Admits queued searches while workers and budget allow; caller holds mu.
void dispatch_bfs(lpma_server *s, std::unique_lock<std::mutex> &lk) {
    const u64 bb = bfs_bytes(s->V, lpma_snapshot_bytes(s->h));
    while (!s->q_bfs.empty() && s->busy < (int)s->cfg.bfs_workers) {
        // a search reads the newest snapshot while it is younger than the
        // configured age (its result reports the version it saw); a newer
        // version is copied only after that, so a stream of searches does not
        // turn into a stream of full copies beside the writer
        const bool stale = s->cur && s->cur->applied != s->applied &&
                           us_between(s->cur->taken, clk::now()) >= (double)s->cfg.snapshot_max_age_us;
        const bool need_snap = !s->cur || stale;
        if (need_snap && s->cur) {
            // the newest snapshot is out of date: it goes once unused
            s->cur = nullptr;
            destroy_unused_snaps(s);
        }
        const u64 snap_bytes = need_snap ? lpma_snapshot_bytes(s->h) : 0;
        Request *r = s->q_bfs.front();
        if (snap_bytes + bb > s->cfg.hbm_budget_bytes ||
            (s->in_use + snap_bytes + bb > s->cfg.hbm_budget_bytes && s->busy == 0)) {
            // larger than the budget, or nothing running will return memory
            s->q_bfs.pop_front();
            finish(s, r, LPMA_ERR_OUT_OF_MEMORY);
            s->cv_done.notify_all();
            continue;
        }
        if (s->in_use + snap_bytes + bb > s->cfg.hbm_budget_bytes) {
            s->st.admission_waits++;
            s->bfs_blocked = true;  // a finishing search clears it
            break;
        }
        if (need_snap) {
            // the engine is the only writer and holds the loop: the handle
            // does not change while the copies are queued
            lpma_snapshot_t *snap = nullptr;
            lk.unlock();
            int rc = lpma_snapshot_create(&snap, s->h, s->wst);
            lk.lock();
            if (rc != LPMA_OK) {
                s->q_bfs.pop_front();
                finish(s, r, rc);
                s->cv_done.notify_all();
                continue;
            }
            lpma_snapshot_set_search_width(snap, s->cfg.search_workgroups);
            Snap *p = new Snap();
            p->s = snap;
            p->applied = s->applied;
            p->bytes = snap_bytes;
            p->taken = clk::now();
            s->snaps.push_back(p);
            s->cur = p;
            s->in_use += snap_bytes;
            s->st.snapshots++;
        }
        s->q_bfs.pop_front();
        r->snap = s->cur;
        r->snap->refs++;
        r->charged = bb;
        r->applied = s->cur->applied;
        s->in_use += bb;
        s->peak = std::max(s->peak, s->in_use);
        s->busy++;
        s->st.max_concurrent_bfs = std::max<u64>(s->st.max_concurrent_bfs, (u64)s->busy);
        s->q_work.push_back(r);
        s->cv_work.notify_one();
    }
}

void engine_loop(lpma_server *s) {
    (void)hipSetDevice(s->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_engine.wait(lk, [&] {
            return !s->q_upd.empty() || !s->q_read.empty() ||
                   (!s->q_bfs.empty() && s->busy < (int)s->cfg.bfs_workers && !s->bfs_blocked) ||
                   (s->stop && s->q_bfs.empty() && s->busy == 0);
        });
        if (s->stop && s->q_upd.empty() && s->q_read.empty() && s->q_bfs.empty() && s->busy == 0) break;
        std::vector<Request *> ub, rb;
        u64 keys = 0, verts = 0;
        while (!s->q_upd.empty() && ub.size() < s->max_reqs) {
            Request *r = s->q_upd.front();
            u64 k = r->ins.size() + r->del.size();
            if (!ub.empty() && keys + k > s->cfg.max_merge_keys) break;
            ub.push_back(r);
            keys += k;
            s->q_upd.pop_front();
        }
        while (!s->q_read.empty()) {
            Request *r = s->q_read.front();
            if (!rb.empty() && verts + r->verts.size() > s->cfg.max_merge_vertices) break;
            rb.push_back(r);
            verts += r->verts.size();
            s->q_read.pop_front();
        }
        auto now = clk::now();
        for (Request *r : ub) r->t_start = now;
        for (Request *r : rb) r->t_start = now;
        lk.unlock();
        auto t1 = clk::now();
        if (!ub.empty()) serve_updates(s, ub);
        auto t2 = clk::now();
        if (!rb.empty()) serve_reads(s, rb);
        auto t3 = clk::now();
        lk.lock();
        s->st.update_us += us_between(t1, t2);
        s->st.read_us += us_between(t2, t3);
        auto t4 = clk::now();
        destroy_unused_snaps(s);
        dispatch_bfs(s, lk);
        s->st.dispatch_us += us_between(t4, clk::now());
    }
    s->stop_workers = true;
    s->cv_work.notify_all();
}

hipStream_t worker_stream() {
    hipStream_t ws = nullptr;
    if (hipStreamCreateWithFlags(&ws, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        ws = nullptr;
    }
    return ws;
}

void worker_loop(lpma_server *s) {
    (void)hipSetDevice(s->device);
    hipStream_t ws = worker_stream();
    u32 *d_levels = nullptr;
    if (hipMalloc((void **)&d_levels, ((u64)s->V + 1) * 4) != hipSuccess) {
        (void)hipGetLastError();
        d_levels = nullptr;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        s->cv_work.wait(lk, [&] { return s->stop_workers || !s->q_work.empty(); });
        if (s->q_work.empty() && s->stop_workers) break;
        Request *r = s->q_work.front();
        s->q_work.pop_front();
        r->t_start = clk::now();
        lk.unlock();
        int rc = d_levels ? LPMA_OK : LPMA_ERR_OUT_OF_MEMORY;
        u32 depth = 0;
        u64 scanned = 0;
        lpma_view_t v;
        lpma_snapshot_get_view(r->snap->s, &v);
        if (rc == LPMA_OK) rc = lpma_snapshot_bfs(r->snap->s, r->root, d_levels, &depth, &scanned, ws);
        if (rc == LPMA_OK && r->want_levels) {
            r->values.resize(s->V);
            if (hipMemcpyAsync(r->values.data(), d_levels, (u64)s->V * 4, hipMemcpyDeviceToHost, ws) != hipSuccess ||
                hipStreamSynchronize(ws) != hipSuccess) {
                (void)hipGetLastError();
                rc = LPMA_ERR_DEVICE;
            }
        }
        lk.lock();
        r->depth = depth;
        r->scanned = scanned;
        r->version = v.version;
        r->snap->refs--;
        r->snap = nullptr;
        s->in_use -= r->charged;
        s->busy--;
        s->bfs_blocked = false;
        finish(s, r, rc);
        s->cv_done.notify_all();
        s->cv_engine.notify_one();
    }
    lk.unlock();
    if (d_levels) (void)hipFree(d_levels);
    if (ws) (void)hipStreamDestroy(ws);
}

int submit(lpma_server *s, Request *r, std::deque<Request *> &q, uint64_t *ticket) {
    std::lock_guard<std::mutex> g(s->mu);
    if (s->stop) {
        delete r;
        return LPMA_ERR_INVALID_ARGUMENT;
    }
    r->id = s->next_id++;
    r->t_submit = clk::now();
    s->reqs[r->id] = r;
    q.push_back(r);
    s->st.submitted[r->kind]++;
    *ticket = r->id;
    s->cv_engine.notify_one();
    return LPMA_OK;
}

void fill_result(const Request *r, lpma_server_result_t *out) {
    memset(out, 0, sizeof(*out));
    out->status = r->status;
    out->kind = r->kind;
    out->applied_updates = r->applied;
    out->version = r->version;
    out->queue_us = us_between(r->t_submit, r->t_start);
    out->latency_us = us_between(r->t_submit, r->t_done);
    if (r->kind == LPMA_REQ_READ) {
        out->count = r->verts.size();
        out->offsets = r->offsets.empty() ? nullptr : r->offsets.data();
        out->values = r->values.empty() ? nullptr : r->values.data();
    } else if (r->kind == LPMA_REQ_BFS) {
        out->count = r->values.size();
        out->values = r->values.empty() ? nullptr : r->values.data();
        out->depth = r->depth;
        out->edges_scanned = r->scanned;
    }
}

}  // namespace

extern "C" {

void lpma_server_config_default(lpma_server_config_t *cfg) {
    if (!cfg) return;
    cfg->bfs_workers = 2;
    cfg->hbm_budget_bytes = 0;  // 0: 90% of the device memory free at creation
    cfg->max_merge_keys = 1ull << 24;
    cfg->max_merge_requests = 4096;
    cfg->max_merge_vertices = 1ull << 20;
    cfg->snapshot_max_age_us = 50000;
    cfg->search_workgroups = 256;
}

int lpma_server_create(lpma_server_t **out, lpma_t *h, const lpma_server_config_t *cfg_in) {
    if (!out || !h) return LPMA_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    lpma_server_config_t cfg;
    lpma_server_config_default(&cfg);
    if (cfg_in) cfg = *cfg_in;
    if (cfg.bfs_workers == 0 || cfg.bfs_workers > 64 || cfg.max_merge_requests == 0 || cfg.max_merge_keys == 0 ||
        cfg.max_merge_vertices == 0)
        return LPMA_ERR_INVALID_ARGUMENT;
    lpma_view_t v;
    if (lpma_get_view(h, &v) != LPMA_OK) return LPMA_ERR_INVALID_ARGUMENT;
    lpma_server *s = new (std::nothrow) lpma_server();
    if (!s) return LPMA_ERR_OUT_OF_MEMORY;
    memset(&s->st, 0, sizeof(s->st));
    s->h = h;
    s->V = v.num_vertices;
    (void)hipGetDevice(&s->device);
    if (cfg.hbm_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
            (void)hipGetLastError();
            fr = 0;
        }
        cfg.hbm_budget_bytes = (u64)(0.9 * (double)fr);
    }
    s->cfg = cfg;
    // the merged requests' order bits sit above the widest key (2 vbits + 1)
    u32 vb = 1;
    while ((1ull << vb) < s->V) ++vb;
    u32 free_bits = 64 - (2 * vb + 1);
    u64 fit = free_bits >= 31 ? (1ull << 31) : (1ull << free_bits);
    s->max_reqs = (u32)std::min<u64>(cfg.max_merge_requests, fit);
    if (s->max_reqs == 0) s->max_reqs = 1;
    if (hipStreamCreateWithFlags(&s->wst, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete s;
        return LPMA_ERR_DEVICE;
    }
    s->st.budget_bytes = cfg.hbm_budget_bytes;
    s->engine = std::thread(engine_loop, s);
    for (u32 i = 0; i < cfg.bfs_workers; ++i) s->workers.emplace_back(worker_loop, s);
    *out = s;
    return LPMA_OK;
}

int lpma_server_destroy(lpma_server_t *s) {
    if (!s) return LPMA_ERR_INVALID_ARGUMENT;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv_engine.notify_all();
    }
    s->engine.join();
    for (auto &t : s->workers) t.join();
    {
        std::lock_guard<std::mutex> g(s->mu);
        destroy_unused_snaps(s);
    }
    (void)hipStreamSynchronize(s->wst);
    if (s->d_keys.p) (void)hipFree(s->d_keys.p);
    if (s->d_verts.p) (void)hipFree(s->d_verts.p);
    if (s->d_len.p) (void)hipFree(s->d_len.p);
    if (s->d_off.p) (void)hipFree(s->d_off.p);
    if (s->d_out.p) (void)hipFree(s->d_out.p);
    if (s->h_keys.p) (void)hipHostFree(s->h_keys.p);
    if (s->h_verts.p) (void)hipHostFree(s->h_verts.p);
    if (s->h_len.p) (void)hipHostFree(s->h_len.p);
    if (s->h_off.p) (void)hipHostFree(s->h_off.p);
    if (s->h_out.p) (void)hipHostFree(s->h_out.p);
    (void)hipStreamDestroy(s->wst);
    for (auto &kv : s->reqs) delete kv.second;
    delete s;
    return LPMA_OK;
}

int lpma_server_submit_update(lpma_server_t *s, const uint64_t *h_ins, uint64_t n_ins, const uint64_t *h_del,
                              uint64_t n_del, uint64_t *ticket) {
    if (!s || !ticket || (!h_ins && n_ins) || (!h_del && n_del)) return LPMA_ERR_INVALID_ARGUMENT;
    if (n_ins + n_del > s->cfg.hbm_budget_bytes / 8) return LPMA_ERR_OUT_OF_MEMORY;
    Request *r = new (std::nothrow) Request();
    if (!r) return LPMA_ERR_OUT_OF_MEMORY;
    r->kind = LPMA_REQ_UPDATE;
    r->ins.assign(h_ins, h_ins + n_ins);
    r->del.assign(h_del, h_del + n_del);
    return submit(s, r, s->q_upd, ticket);
}

int lpma_server_submit_read(lpma_server_t *s, const uint32_t *h_vertices, uint64_t n, uint64_t *ticket) {
    if (!s || !ticket || (!h_vertices && n)) return LPMA_ERR_INVALID_ARGUMENT;
    Request *r = new (std::nothrow) Request();
    if (!r) return LPMA_ERR_OUT_OF_MEMORY;
    r->kind = LPMA_REQ_READ;
    r->verts.assign(h_vertices, h_vertices + n);
    return submit(s, r, s->q_read, ticket);
}

int lpma_server_submit_bfs(lpma_server_t *s, uint32_t root, int want_levels, uint64_t *ticket) {
    if (!s || !ticket || root >= s->V) return LPMA_ERR_INVALID_ARGUMENT;
    Request *r = new (std::nothrow) Request();
    if (!r) return LPMA_ERR_OUT_OF_MEMORY;
    r->kind = LPMA_REQ_BFS;
    r->root = root;
    r->want_levels = want_levels != 0;
    return submit(s, r, s->q_bfs, ticket);
}

int lpma_server_wait(lpma_server_t *s, uint64_t ticket, lpma_server_result_t *out) {
    if (!s || !out) return LPMA_ERR_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return LPMA_ERR_INVALID_ARGUMENT;
    Request *r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    fill_result(r, out);
    return LPMA_OK;
}

int lpma_server_poll(lpma_server_t *s, uint64_t ticket, int *done, lpma_server_result_t *out) {
    if (!s || !done) return LPMA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) return LPMA_ERR_INVALID_ARGUMENT;
    *done = it->second->done ? 1 : 0;
    if (*done && out) fill_result(it->second, out);
    return LPMA_OK;
}

int lpma_server_release(lpma_server_t *s, uint64_t ticket) {
    if (!s) return LPMA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end() || !it->second->done) return LPMA_ERR_INVALID_ARGUMENT;
    delete it->second;
    s->reqs.erase(it);
    return LPMA_OK;
}

int lpma_server_get_stats(lpma_server_t *s, lpma_server_stats_t *out) {
    if (!s || !out) return LPMA_ERR_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(s->mu);
    *out = s->st;
    out->in_use_bytes = s->in_use;
    out->peak_bytes = s->peak;
    out->applied_updates = s->applied;
    return LPMA_OK;
}

}  // extern "C"
