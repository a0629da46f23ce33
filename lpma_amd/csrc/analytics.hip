// analytics.hip -- BFS and PageRank straight over the gapped layout.
//
// The reference runs nvGraph on a CSR copy (test_algos.cu:44-90 PageRank,
// :136-186 BFS; test.cu:170-182 converts with to_csr_rpma first).  Here both
// read the PMA in place: row r's out-edges are the live slots between the
// markers of rows r-1 and r, i.e. slots (row_off[r], row_off[r+1]).
#include <atomic>

#include "common.hpp"
#include "kernels.hpp"

namespace pma {

static constexpr u32 INF_LEVEL = 0xFFFFFFFFu;

__device__ __forceinline__ void slot_range(const DevState &d, u32 r, u64 &lo, u64 &hi) {
    lo = r == 0 ? 0 : (u64)d.row_off[r] + 1;
    hi = d.row_off[r + 1];
}

// ------------------------------------------------------------------ BFS
//
// Level-synchronous, top-down.  A level expands the frontier's slot ranges
// with one thread per slot, so hub vertices are split over many threads:
//   k_bfs_len   each frontier row's slot count and first slot;
//   scan        the counts' exclusive prefix: slot t of the level belongs to
//               the last entry whose offset is <= t;
//   k_bfs_runs  the first entry of every run of EXP_RUN slots (one binary
//               search per run, all in parallel);
//   k_bfs_expand  a workgroup per run stages the run's entries (offsets and
//               first slots) in LDS; each thread takes EXP_ITEMS slots and
//               issues their loads phase by phase (segment counts, keys,
//               visited words), so dead lines of the gapped rows are not read
//               and a thread has EXP_ITEMS loads in flight per phase.
// A visited bitmap (V bits: 2 MB at V = 2^24, held by every XCD's L2) filters
// the edges; the levels array (32 bits per vertex) is only written.  A vertex
// is claimed by the atomic OR that sets its bit, so it enters the next
// frontier once; a plain read of a word another XCD has just changed can only
// miss bits (they go from 0 to 1), which costs an extra atomic, never a wrong
// level.

__global__ void k_bfs_init(u32 *levels, u32 V, u32 src, u32 *queue, u32 *counters, u32 *visited) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (u64)gridDim.x * blockDim.x)
        levels[i] = i == src ? 0u : INF_LEVEL;
    for (u64 w = (u64)blockIdx.x * blockDim.x + threadIdx.x; w < bfs_visited_words(V); w += (u64)gridDim.x * blockDim.x)
        visited[w] = w == (src >> 5) ? 1u << (src & 31) : 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        queue[0] = src;
        counters[0] = 0;  // next frontier size
        counters[1] = 0;  // live edges scanned
    }
}

__global__ void k_bfs_len(DevState d, const u32 *__restrict__ q, u32 qn, u32 *__restrict__ len,
                          u64 *__restrict__ first) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < qn; i += (u64)gridDim.x * blockDim.x) {
        u64 lo, hi;
        slot_range(d, q[i], lo, hi);
        len[i] = (u32)(hi - lo);
        first[i] = lo;
    }
}

static constexpr u32 EXP_TPB = 512;
static constexpr u32 EXP_ITEMS = 4;
static constexpr u32 EXP_RUN = EXP_TPB * EXP_ITEMS;  // slots per run
static constexpr u32 EXP_STAGE = EXP_RUN + 1;        // staged entries: every entry of a run of non-empty
                                                     // rows; empty rows past it use the global search

// last entry e in [lo, hi) with off[e] <= t (off[lo] <= t)
__device__ __forceinline__ u32 entry_of(const u32 *__restrict__ off, u32 lo, u32 hi, u64 t) {
    while (hi - lo > 1) {
        const u32 m = (lo + hi) >> 1;
        if ((u64)off[m] <= t) lo = m; else hi = m;
    }
    return lo;
}

// runs[r] = entry holding slot r * EXP_RUN
__global__ void k_bfs_runs(const u32 *__restrict__ off, u32 qn, u32 *__restrict__ runs) {
    const u64 total = off[qn];
    const u64 nruns = (total + EXP_RUN - 1) / EXP_RUN;
    for (u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x; r < nruns; r += (u64)gridDim.x * blockDim.x)
        runs[r] = entry_of(off, 0, qn, r * EXP_RUN);
}

// Moves the workgroup's discovered vertices from LDS to the next frontier:
// one atomic on the frontier size per flush.
__device__ __forceinline__ void bfs_flush(u32 *s_q, u32 &s_n, u32 &s_base, u32 *__restrict__ qnext,
                                          u32 *__restrict__ counters) {
    const u32 n = s_n;  // read by all after the barrier that precedes the call
    if (threadIdx.x == 0) s_base = atomicAdd(&counters[0], n);
    __syncthreads();
    const u32 base = s_base;
    for (u32 k = threadIdx.x; k < n; k += EXP_TPB) qnext[base + k] = s_q[k];
    __syncthreads();
    if (threadIdx.x == 0) s_n = 0;
    __syncthreads();
}

// entries of a run's stage each thread loads (and later writes to LDS)
static constexpr u32 EXP_SK = (EXP_STAGE + EXP_TPB - 1) / EXP_TPB;

// the stage of the run starting at entry e0, into registers
__device__ __forceinline__ void bfs_stage_load(const u32 *__restrict__ off, const u64 *__restrict__ first, u32 qn,
                                               u32 e0, u32 (&p_off)[EXP_SK], u64 (&p_first)[EXP_SK]) {
#pragma unroll
    for (u32 k = 0; k < EXP_SK; ++k) {
        const u32 kk = threadIdx.x + k * EXP_TPB;
        const u64 idx = (u64)e0 + kk;
        p_off[k] = (kk < EXP_STAGE && idx <= qn) ? off[idx] : 0xFFFFFFFFu;
        p_first[k] = (kk < EXP_STAGE && idx < qn) ? first[idx] : 0ull;
    }
}

__global__ __launch_bounds__(EXP_TPB) void k_bfs_expand(DevState d, u32 qn, const u32 *__restrict__ off,
                                                         const u64 *__restrict__ first,
                                                         const u32 *__restrict__ runs, u32 *__restrict__ levels,
                                                         u32 *__restrict__ visited, u32 level,
                                                         u32 *__restrict__ qnext, u32 *__restrict__ counters) {
    __shared__ u32 s_off[EXP_STAGE];
    __shared__ u64 s_first[EXP_STAGE];
    __shared__ u32 s_q[2 * EXP_RUN];  // discovered vertices; flushed once it holds EXP_RUN
    __shared__ u32 s_n, s_base;
    if (threadIdx.x == 0) s_n = 0;
    const u64 total = off[qn];
    const u64 nruns = (total + EXP_RUN - 1) / EXP_RUN;
    const u32 lane = lane_id();
    const u32 Smask = (1u << d.log2S) - 1;
    const gu64 *keys = (const gu64 *)d.keys;
    u32 scanned = 0;
    // a run's entry and stage are loaded one run ahead, into registers, and
    // written to LDS once the previous run is done with the stage
    u32 p_off[EXP_SK];
    u64 p_first[EXP_SK];
    u32 e0n = 0;
    if (blockIdx.x < nruns) {
        e0n = runs[blockIdx.x];
        bfs_stage_load(off, first, qn, e0n, p_off, p_first);
    }
    for (u64 run = blockIdx.x; run < nruns; run += gridDim.x) {
        const u64 t0 = run * EXP_RUN;
        const u32 e0 = e0n;
#pragma unroll
        for (u32 k = 0; k < EXP_SK; ++k) {
            const u32 kk = threadIdx.x + k * EXP_TPB;
            if (kk < EXP_STAGE) {
                s_off[kk] = p_off[k];
                s_first[kk] = p_first[k];
            }
        }
        __syncthreads();
        const bool more = run + gridDim.x < nruns;  // uniform
        if (more) e0n = runs[run + gridDim.x];
        // slot of each item: the staged entry holding it, or the global search
        // past the stage (only when empty rows fill it)
        u64 slot[EXP_ITEMS];
        bool ok[EXP_ITEMS];
        const bool staged_all = (u64)s_off[EXP_STAGE - 1] > t0 + EXP_RUN - 1;
#pragma unroll
        for (u32 i = 0; i < EXP_ITEMS; ++i) {
            const u64 t = t0 + (u64)i * EXP_TPB + threadIdx.x;
            ok[i] = t < total;
            slot[i] = 0;
            if (!ok[i]) continue;
            if (staged_all || (u64)s_off[EXP_STAGE - 1] > t) {
                u32 lo = 0, hi = EXP_STAGE - 1;  // s_off[0] <= t0 <= t
                while (hi - lo > 1) {
                    const u32 m = (lo + hi) >> 1;
                    if ((u64)s_off[m] <= t) lo = m; else hi = m;
                }
                slot[i] = s_first[lo] + (t - s_off[lo]);
            } else {
                const u32 e = entry_of(off, e0 + EXP_STAGE - 1, qn, t);
                slot[i] = first[e] + (t - off[e]);
            }
        }
        // the next run's stage: its entry has arrived during the searches
        if (more) bfs_stage_load(off, first, qn, e0n, p_off, p_first);
        u32 live[EXP_ITEMS];
#pragma unroll
        for (u32 i = 0; i < EXP_ITEMS; ++i) live[i] = d.cnt[slot[i] >> d.log2S];  // slot 0 when !ok: in bounds
        u64 key[EXP_ITEMS];
#pragma unroll
        for (u32 i = 0; i < EXP_ITEMS; ++i) {
            u64 k = KEY_NONE;
            if (ok[i] && (u32)(slot[i] & Smask) < live[i]) k = keys[slot[i]];
            key[i] = k;
        }
        u32 vw[EXP_ITEMS];
#pragma unroll
        for (u32 i = 0; i < EXP_ITEMS; ++i) vw[i] = key[i] != KEY_NONE ? visited[(u32)key[i] >> 5] : ~0u;
#pragma unroll
        for (u32 i = 0; i < EXP_ITEMS; ++i) {
            bool found = false;
            if (key[i] != KEY_NONE) {
                ++scanned;
                const u32 v = (u32)key[i], bit = 1u << (v & 31);
                if (!(vw[i] & bit) && !(atomicOr(&visited[v >> 5], bit) & bit)) {
                    found = true;
                    levels[v] = level + 1;
                }
            }
            const u64 m = __ballot(found);
            if (m) {
                const u32 lead = (u32)__builtin_ctzll(m);
                u32 base = 0;
                if (lane == lead) base = atomicAdd(&s_n, (u32)__popcll(m));  // LDS
                base = __shfl(base, lead, 64);
                if (found) s_q[base + __popcll(m & ((1ull << lane) - 1ull))] = (u32)key[i];
            }
        }
        __syncthreads();  // the stage is rewritten by the next run; s_n is final for this run
        if (s_n >= EXP_RUN) bfs_flush(s_q, s_n, s_base, qnext, counters);  // a run adds at most EXP_RUN
    }
    __syncthreads();
    if (s_n) bfs_flush(s_q, s_n, s_base, qnext, counters);
    scanned = wave_sum(scanned);
    if (lane == 0 && scanned) atomicAdd(&counters[1], scanned);
}

hipError_t bfs_init(u32 *levels, u32 V, u32 src, u32 *queue, u32 *counters, u32 *visited, hipStream_t st) {
    k_bfs_init<<<grid_for(V, 256, 16384), 256, 0, st>>>(levels, V, src, queue, counters, visited);
    return hipGetLastError();
}

u64 bfs_run_words(u64 slots) { return slots / EXP_RUN + 2; }

// Default expansion grid: every workgroup resident at once (compute units x
// workgroups per CU that fit), so the runs are shared out in one round; a
// grid one resident round larger leaves its last workgroups a tail of their
// own (2048 workgroups of 512 threads: 9.1 ms per search, 1024: 11.6, one
// round of 768: 8.1 on MI355X).
static u32 bfs_expand_grid() {
    static std::atomic<u32> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 2048;
    u32 g = cached[dev].load(std::memory_order_relaxed);
    if (g) return g;
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_bfs_expand, EXP_TPB, 0) != hipSuccess || cus <= 0 ||
        per_cu <= 0)
        return 2048;
    g = (u32)(cus * per_cu);
    cached[dev].store(g, std::memory_order_relaxed);
    return g;
}

hipError_t bfs_level(const DevState &d, const u32 *q, u32 qn, u32 *len, u64 *first, u32 *runs, u32 *scan_scratch,
                     u32 *levels, u32 *visited, u32 level, u32 *qnext, u32 *counters, u32 expand_grid,
                     hipStream_t st) {
    k_bfs_len<<<grid_for(qn, 256, 16384), 256, 0, st>>>(d, q, qn, len, first);
    (void)device_excl_scan(len, qn, len, scan_scratch, st);
    k_bfs_runs<<<grid_for(bfs_run_words(nseg_of(d) << d.log2S), 256, 16384), 256, 0, st>>>(len, qn, runs);
    k_bfs_expand<<<expand_grid ? expand_grid : bfs_expand_grid(), EXP_TPB, 0, st>>>(d, qn, len, first, runs, levels, visited,
                                                                      level, qnext, counters);
    return hipGetLastError();
}

// ------------------------------------------------------------- PageRank
//
// nvGraph's formulation (include/pagerank.hxx:66-88): with a = alpha on
// dangling vertices and 1-alpha elsewhere,
//     pr' = alpha * A^T (pr / outdeg) + (sum_u a_u pr_u) / n,
// iterated from the uniform vector; here until the L1 change is below the
// tolerance or max_iter, then normalised by the L1 norm.
//
// Values and accumulation are FP32.  Every accumulation is compensated: an
// atomic add returns the old value, the thread recomputes the sum and its
// exact rounding error (TwoSum, Knuth TAOCP vol. 2, 4.2.2) and adds that
// error to a second FP32 accumulator, so a hub that receives millions of
// contributions keeps the precision of its partial sums.  The push goes to
// one accumulator copy per XCD (the XCD read from the hardware register), so
// the atomics of a hub are spread over eight addresses; the copies are
// summed (Kahan) when the iteration finishes.

static constexpr u32 PR_COPIES = 8;  // accumulator copies, one per XCD

__device__ __forceinline__ u32 xcc_id() {
    u32 x;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID, 0, 4)" : "=s"(x));
    return x & (PR_COPIES - 1);
}

// hi += x with the rounding error of that add going to lo
__device__ __forceinline__ void atomic_add2(float *hi, float *lo, float x) {
    const float old = atomicAdd(hi, x);
    const float s = old + x;
    const float bp = s - old;
    const float e = (old - (s - bp)) + (x - bp);
    if (e != 0.f) atomicAdd(lo, e);
}

// sum over a workgroup of PR_TPB threads (every thread gets it); tmp holds
// PR_TPB / 64 floats of LDS
static constexpr u32 PR_TPB = 256;
static constexpr u32 PR_GRID = 2048;  // vector passes: few enough blocks for one atomic each
__device__ __forceinline__ float block_sum(float v, float *tmp) {
    v = wave_sum(v);
    __syncthreads();  // tmp may still be read from a previous call
    if (lane_id() == 0) tmp[threadIdx.x >> 6] = v;
    __syncthreads();
    float t = 0.f;
#pragma unroll
    for (u32 w = 0; w < PR_TPB / 64; ++w) t += tmp[w];
    return t;
}

// red layout (floats): [0,1] dangling mass hi/lo, [2,3] sum hi/lo,
// [4,5] L1 change hi/lo
static constexpr u32 PR_RED = 8;

// out-degree of every row from marker ranks (rank = live slots before)
__global__ void k_outdeg(DevState d, const u32 *__restrict__ gpre, u32 *__restrict__ outdeg) {
    const u32 Smask = (1u << d.log2S) - 1;
    for (u64 r = (u64)blockIdx.x * blockDim.x + threadIdx.x; r < d.V; r += (u64)gridDim.x * blockDim.x) {
        u64 hs = d.row_off[r + 1];
        u64 rank_hi = (u64)gpre[hs >> d.log2S] + (hs & Smask);
        u64 rank_lo = 0;
        if (r > 0) {
            u64 ls = d.row_off[r];
            rank_lo = (u64)gpre[ls >> d.log2S] + (ls & Smask) + 1;
        }
        outdeg[r] = (u32)(rank_hi - rank_lo);
    }
}

// contributions pr/outdeg, the dangling mass, and cleared accumulators
// (PR_COPIES hi copies, then PR_COPIES lo copies, V floats each)
__global__ __launch_bounds__(PR_TPB) void k_pr_prep(const float *__restrict__ pr, const u32 *__restrict__ outdeg, u32 V, float *__restrict__ c,
                          float *__restrict__ acc, u64 nclear, float *__restrict__ red) {
    float dang = 0.f;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x; u < V; u += stride) {
        u32 od = outdeg[u];
        float p = pr[u];
        c[u] = od ? p / (float)od : 0.f;
        if (!od) dang += p;
    }
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nclear; i += stride) acc[i] = 0.f;
    __shared__ float tmp[PR_TPB / 64];
    dang = block_sum(dang, tmp);
    if (threadIdx.x == 0 && dang != 0.f) atomic_add2(&red[0], &red[1], dang);
}

__global__ void k_pr_push(DevState d, const float *__restrict__ c, float *__restrict__ acc) {
    const u64 W = nseg_of(d) << d.log2S;
    const u64 V = d.V;
    float *hi = acc + (u64)xcc_id() * V;
    float *lo = hi + PR_COPIES * V;
    for (u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x; t < W; t += (u64)gridDim.x * blockDim.x) {
        u64 key = d.keys[t];
        if (key == KEY_NONE || is_marker(key)) continue;
        const u32 v = (u32)key;
        atomic_add2(&hi[v], &lo[v], c[(u32)(key >> 32)]);
    }
}

// pr' = alpha*acc + (alpha*dangling + (1-alpha)*sum(pr)) / n; L1 change and sum
__global__ __launch_bounds__(PR_TPB) void k_pr_finish(const float *__restrict__ acc, const float *__restrict__ pr, float *__restrict__ out,
                            u32 V, float alpha, const float *__restrict__ red_in, float *__restrict__ red_out) {
    const float dang = red_in[0] + red_in[1];
    const float sum_in = red_in[2] + red_in[3];
    const float gamma = (alpha * dang + (1.f - alpha) * sum_in) / (float)V;
    float dl = 0.f, sum = 0.f;
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (u64)gridDim.x * blockDim.x) {
        // Kahan sum of the copies, lo parts first (the small terms)
        float a = 0.f, comp = 0.f;
#pragma unroll
        for (u32 k = 2 * PR_COPIES; k-- > 0;) {
            float y = acc[(u64)k * V + v] - comp;
            float t = a + y;
            comp = (t - a) - y;
            a = t;
        }
        float p = alpha * a + gamma;
        out[v] = p;
        dl += fabsf(p - pr[v]);
        sum += p;
    }
    __shared__ float tmp[PR_TPB / 64];
    dl = block_sum(dl, tmp);
    sum = block_sum(sum, tmp);
    if (threadIdx.x == 0) {
        atomic_add2(&red_out[4], &red_out[5], dl);
        atomic_add2(&red_out[2], &red_out[3], sum);
    }
}

__global__ void k_pr_init(float *pr, u32 V, float *red) {
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (u64)gridDim.x * blockDim.x)
        pr[v] = 1.f / (float)V;
    if (blockIdx.x == 0 && threadIdx.x < PR_RED) red[threadIdx.x] = threadIdx.x == 2 ? 1.f : 0.f;  // sum = 1
}

__global__ void k_pr_reset(float *red_next) {
    if (threadIdx.x < PR_RED) red_next[threadIdx.x] = 0.f;
}

__global__ void k_pr_scale(float *pr, u32 V, const float *red) {
    const float inv = 1.f / (red[2] + red[3]);
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (u64)gridDim.x * blockDim.x) pr[v] *= inv;
}

hipError_t pr_outdeg(const DevState &d, u32 *gpre, u32 *scan_scratch, u32 *outdeg, hipStream_t st) {
    (void)device_excl_scan(d.cnt, nseg_of(d), gpre, scan_scratch, st);
    k_outdeg<<<grid_for(d.V, 256, 16384), 256, 0, st>>>(d, gpre, outdeg);
    return hipGetLastError();
}

u64 pr_acc_floats(u32 V) { return 2ull * PR_COPIES * V; }
u32 pr_red_floats() { return PR_RED; }

hipError_t pr_init(float *pr, u32 V, float *red, hipStream_t st) {
    k_pr_init<<<grid_for(V, 256, 16384), 256, 0, st>>>(pr, V, red);
    return hipGetLastError();
}

// one iteration: pr -> out; red holds (dangling, sum) of pr on entry and
// (dangling of out, sum of out, L1 change) on exit in red_next
hipError_t pr_iterate(const DevState &d, const float *pr, const u32 *outdeg, float *c, float *acc, float *out,
                      float alpha, float *red, float *red_next, hipStream_t st) {
    const u32 V = d.V;
    k_pr_reset<<<1, 64, 0, st>>>(red_next);
    k_pr_prep<<<grid_for(pr_acc_floats(V), PR_TPB, PR_GRID), PR_TPB, 0, st>>>(pr, outdeg, V, c, acc, pr_acc_floats(V),
                                                                         red);
    k_pr_push<<<grid_for(nseg_of(d) << d.log2S, 256, 65536), 256, 0, st>>>(d, c, acc);
    k_pr_finish<<<grid_for(V, PR_TPB, PR_GRID), PR_TPB, 0, st>>>(acc, pr, out, V, alpha, red, red_next);
    return hipGetLastError();
}

hipError_t pr_scale(float *pr, u32 V, const float *red, hipStream_t st) {
    k_pr_scale<<<grid_for(V, 256, 16384), 256, 0, st>>>(pr, V, red);
    return hipGetLastError();
}

// ------------------------------------------------- binned push (PageRank)
//
// The push above sends every edge's contribution to a random destination in
// HBM: one returning atomic per edge on a 64 MB vector, which no L2 holds.
// Propagation blocking (Beamer, Asanovic, Patterson, "Reducing PageRank
// communication via propagation blocking", IPDPS 2017) splits it in two
// streaming passes over the same gapped layout:
//   bin pass   the slot array is cut into parts of 2^part_log slots; each
//              part's workgroup writes (dst mod PB_BIN, contribution) pairs
//              into per-destination-bin runs, at offsets fixed once per call
//              by a count pass (bin-major exclusive scan of count[bin][part]);
//   acc pass   each chunk of a bin's pairs is summed in LDS (PB_BIN vertices,
//              compensated like the atomics above) and written as one
//              partial vector; the finish pass adds a bin's partials (Kahan).
// Every access to HBM is then a stream; the random adds stay in LDS.  The
// graph does not change during a call, so the count pass runs once.

static constexpr u32 PB_BIN_LOG = 13;  // destination vertices per bin
static constexpr u32 PB_BIN = 1u << PB_BIN_LOG;
static constexpr u32 PB_MAX_BINS = 16384;  // LDS cursors of the bin pass: 64 KB
static constexpr u32 PB_TPB = 1024;
static constexpr u32 PB_ITEMS = 8;
static constexpr u32 PB_MAX_PARTS = 1024;
static constexpr u32 PB_CHUNK = 4;                // pairs per staged chunk (32 B, one HBM sector)
static constexpr u32 PB_STAGE_MAX_BINS = 2048;    // staged bin pass LDS: 40 B per bin
static constexpr u32 PBS_TPB = 512;               // staged bin pass: two workgroups per CU (80 KB LDS each)
static constexpr u32 PBS_ITEMS = 8;
static constexpr u32 PBA_TPB = 512;               // acc pass: fewer waves per LDS accumulator contend less

// Lanes holding a live edge of the same bin form runs (adjacent lanes; keys
// are sorted, so a row's edges to one bin are adjacent).  Returns the lane
// that leads this lane's run, and for a leader the run's length in `len`.
__device__ __forceinline__ u32 bin_run(bool valid, u32 bin, u32 &len) {
    const u32 lane = lane_id();
    const u32 bb = valid ? bin : 0xFFFFFFFFu;
    const u32 prev = __shfl_up(bb, 1, 64);
    const u32 next = __shfl_down(bb, 1, 64);
    const u64 starts = __ballot(valid && (lane == 0 || prev != bb));
    const u64 ends = __ballot(valid && (lane == 63 || next != bb));
    // the sentinel bits keep both masks non-zero; a valid lane's own run
    // bits are the highest start at or below it and the lowest end at or above
    const u32 lead = 63 - (u32)__clzll((starts & ((2ull << lane) - 1ull)) | 1ull);
    const u32 end = (u32)__builtin_ctzll((ends & ~((1ull << lane) - 1ull)) | (1ull << 63));
    len = end - lane + 1;
    return lead;
}

// PB_ITEMS slots t0 + i * PB_TPB + threadIdx.x of a part ending at p1: the
// live ones' keys, KEY_NONE elsewhere.  All segment counts are loaded first
// (clamped, unconditional), then the keys with exec-masked global loads, so a
// thread has PB_ITEMS loads in flight per phase and dead lines are not read.
template <u32 TPB = PB_TPB, u32 ITEMS = PB_ITEMS>
__device__ __forceinline__ void pb_load(const DevState &d, u64 t0, u64 p1, u64 (&key)[ITEMS]) {
    const u32 Smask = (1u << d.log2S) - 1;
    const gu64 *keys = (const gu64 *)d.keys;
    u32 live[ITEMS];
#pragma unroll
    for (u32 i = 0; i < ITEMS; ++i) {
        const u64 t = t0 + (u64)i * TPB + threadIdx.x;
        live[i] = d.cnt[(t < p1 ? t : p1 - 1) >> d.log2S];
    }
#pragma unroll
    for (u32 i = 0; i < ITEMS; ++i) {
        const u64 t = t0 + (u64)i * TPB + threadIdx.x;
        u64 k = KEY_NONE;
        if (t < p1 && (u32)(t & Smask) < live[i]) k = keys[t];
        key[i] = k;
    }
}

__device__ __forceinline__ bool pb_edge(u64 key) { return key != KEY_NONE && !is_marker(key); }

// pb_load in separate phases, for a loop that prefetches its next step:
// the segment counts, then the keys (after the counts), then the
// contributions of the keys' sources (after the keys)
template <u32 TPB, u32 ITEMS>
__device__ __forceinline__ void pb_counts(const DevState &d, u64 t0, u64 p1, u32 (&live)[ITEMS]) {
#pragma unroll
    for (u32 i = 0; i < ITEMS; ++i) {
        const u64 t = t0 + (u64)i * TPB + threadIdx.x;
        live[i] = d.cnt[(t < p1 ? t : p1 - 1) >> d.log2S];
    }
}

template <u32 TPB, u32 ITEMS>
__device__ __forceinline__ void pb_keys(const DevState &d, u64 t0, u64 p1, const u32 (&live)[ITEMS],
                                        u64 (&key)[ITEMS]) {
    const u32 Smask = (1u << d.log2S) - 1;
    const gu64 *keys = (const gu64 *)d.keys;
#pragma unroll
    for (u32 i = 0; i < ITEMS; ++i) {
        const u64 t = t0 + (u64)i * TPB + threadIdx.x;
        u64 k = KEY_NONE;
        if (t < p1 && (u32)(t & Smask) < live[i]) k = keys[t];
        key[i] = k;
    }
}

template <u32 ITEMS>
__device__ __forceinline__ void pb_contrib(const float *__restrict__ c, const u64 (&key)[ITEMS], float (&x)[ITEMS]) {
#pragma unroll
    for (u32 i = 0; i < ITEMS; ++i) x[i] = c[pb_edge(key[i]) ? (u32)(key[i] >> 32) : 0u];  // always in bounds
}

// count[bin * P + part] = live edges of `part` into `bin`, rounded up to a
// multiple of pad + 1 (pad = PB_CHUNK - 1 for the staged bin pass, else 0)
__global__ __launch_bounds__(PB_TPB) void k_pb_count(DevState d, u32 part_log, u32 P, u32 NB, u32 pad,
                                                     u32 *__restrict__ count) {
    extern __shared__ u32 pb_lds[];
    for (u32 b = threadIdx.x; b < NB; b += PB_TPB) pb_lds[b] = 0;
    __syncthreads();
    const u64 W = nseg_of(d) << d.log2S;
    const u64 p0 = (u64)blockIdx.x << part_log;
    const u64 p1 = min(W, p0 + (1ull << part_log));
    for (u64 t0 = p0; t0 < p1; t0 += (u64)PB_TPB * PB_ITEMS) {
        u64 key[PB_ITEMS];
        pb_load(d, t0, p1, key);
#pragma unroll
        for (u32 i = 0; i < PB_ITEMS; ++i) {
            const bool valid = pb_edge(key[i]);
            const u32 bin = (u32)key[i] >> PB_BIN_LOG;
            u32 len;
            const u32 lead = bin_run(valid, bin, len);
            if (valid && lead == lane_id()) atomicAdd(&pb_lds[bin], len);
        }
    }
    __syncthreads();
    for (u32 b = threadIdx.x; b < NB; b += PB_TPB) count[(u64)b * P + blockIdx.x] = (pb_lds[b] + pad) & ~pad;
}

// pairs[off[bin * P + part] + rank] = (contribution bits << 32) | (dst mod PB_BIN)
// Phases per step: all key loads, all contribution loads, all cursor adds,
// then the stores, so each thread keeps PB_ITEMS loads in flight at a time.
__global__ __launch_bounds__(PB_TPB) void k_pb_bin(DevState d, u32 part_log, u32 P, u32 NB, const u32 *__restrict__ off,
                                                   const float *__restrict__ c, u64 *__restrict__ pairs, u64 npairs) {
    extern __shared__ u32 pb_lds[];
    for (u32 b = threadIdx.x; b < NB; b += PB_TPB) pb_lds[b] = off[(u64)b * P + blockIdx.x];
    __syncthreads();
    const u64 W = nseg_of(d) << d.log2S;
    const u64 p0 = (u64)blockIdx.x << part_log;
    const u64 p1 = min(W, p0 + (1ull << part_log));
    const u32 lane = lane_id();
    for (u64 t0 = p0; t0 < p1; t0 += (u64)PB_TPB * PB_ITEMS) {
        u64 key[PB_ITEMS];
        pb_load(d, t0, p1, key);
        float x[PB_ITEMS];
#pragma unroll
        for (u32 i = 0; i < PB_ITEMS; ++i) x[i] = c[pb_edge(key[i]) ? (u32)(key[i] >> 32) : 0u];  // always in bounds
        u32 lead[PB_ITEMS], base[PB_ITEMS];
#pragma unroll
        for (u32 i = 0; i < PB_ITEMS; ++i) {
            const bool valid = pb_edge(key[i]);
            u32 len;
            lead[i] = bin_run(valid, (u32)key[i] >> PB_BIN_LOG, len);
            base[i] = 0;
            if (valid && lead[i] == lane) base[i] = atomicAdd(&pb_lds[(u32)key[i] >> PB_BIN_LOG], len);
            if (!valid) lead[i] = lane;
        }
#pragma unroll
        for (u32 i = 0; i < PB_ITEMS; ++i) {
            const u32 b0 = __shfl(base[i], lead[i], 64);
            const u64 pos = (u64)b0 + (lane - lead[i]);
            if (pb_edge(key[i]) && pos < npairs)
                pairs[pos] = ((u64)__float_as_uint(x[i]) << 32) | ((u32)key[i] & (PB_BIN - 1));
        }
    }
}

// Staged form of the bin pass (bins <= PB_STAGE_MAX_BINS): a bin's pairs
// reach HBM only as whole aligned chunks of PB_CHUNK pairs (32 B), so no
// partly written line is evicted from L2 and written twice.  Every run of
// the count is padded to a multiple of PB_CHUNK; padding pairs are zero (no
// contribution).  Per step, phase by phase:
//   1. ranks: each edge takes rank r on its bin's fill counter (run-aggregated);
//   2. r < PB_CHUNK goes to the bin's LDS chunk, r inside a further complete
//      chunk straight to HBM (a run's lanes are adjacent: coalesced), the
//      rest waits;
//   3. full LDS chunks are written out and the bin's position advances;
//   4. the waiting edges start the next LDS chunk.
// At the end of the part each bin's last chunk is written with zero padding.
__global__ __launch_bounds__(PBS_TPB) void k_pb_bin_staged(DevState d, u32 part_log, u32 P, u32 NB,
                                                          const u32 *__restrict__ off, const float *__restrict__ c,
                                                          u64 *__restrict__ pairs, u64 npairs) {
    extern __shared__ u64 pb_stage[];
    u64 *buf = pb_stage;                                  // NB * PB_CHUNK staged pairs
    u32 *fill = (u32 *)(buf + (u64)NB * PB_CHUNK);        // staged pairs per bin (< PB_CHUNK between steps)
    u32 *gpos = fill + NB;                                // pair index of the bin's LDS chunk in HBM
    for (u32 b = threadIdx.x; b < NB; b += PBS_TPB) {
        fill[b] = 0;
        gpos[b] = off[(u64)b * P + blockIdx.x];
    }
    __syncthreads();
    const u64 W = nseg_of(d) << d.log2S;
    const u64 p0 = (u64)blockIdx.x << part_log;
    const u64 p1 = min(W, p0 + (1ull << part_log));
    const u32 lane = lane_id();
    // the next step's loads go out one phase ahead of their use: counts
    // during the ranking, keys during the placement, contributions during
    // the flush (the barriers wait for LDS only)
    constexpr u64 STEP = (u64)PBS_TPB * PBS_ITEMS;
    u32 live_n[PBS_ITEMS];
    u64 key_n[PBS_ITEMS];
    float x_n[PBS_ITEMS];
    pb_counts<PBS_TPB, PBS_ITEMS>(d, p0, p1, live_n);
    pb_keys<PBS_TPB, PBS_ITEMS>(d, p0, p1, live_n, key_n);
    pb_contrib<PBS_ITEMS>(c, key_n, x_n);
    for (u64 t0 = p0; t0 < p1; t0 += STEP) {
        u64 key[PBS_ITEMS];
        float x[PBS_ITEMS];
#pragma unroll
        for (u32 i = 0; i < PBS_ITEMS; ++i) {
            key[i] = key_n[i];
            x[i] = x_n[i];
        }
        const bool more = t0 + STEP < p1;  // uniform
        if (more) pb_counts<PBS_TPB, PBS_ITEMS>(d, t0 + STEP, p1, live_n);
        u32 rank[PBS_ITEMS];
#pragma unroll
        for (u32 i = 0; i < PBS_ITEMS; ++i) {
            const bool valid = pb_edge(key[i]);
            u32 len;
            u32 lead = bin_run(valid, (u32)key[i] >> PB_BIN_LOG, len);
            u32 base = 0;
            if (valid && lead == lane) base = atomicAdd(&fill[(u32)key[i] >> PB_BIN_LOG], len);
            if (!valid) lead = lane;
            rank[i] = __shfl(base, lead, 64) + (lane - lead);
        }
        __syncthreads();
        if (more) pb_keys<PBS_TPB, PBS_ITEMS>(d, t0 + STEP, p1, live_n, key_n);
        u32 full[PBS_ITEMS];
        u32 wait = 0;
#pragma unroll
        for (u32 i = 0; i < PBS_ITEMS; ++i) {
            full[i] = 0;
            if (!pb_edge(key[i])) continue;
            const u32 bin = (u32)key[i] >> PB_BIN_LOG;
            const u64 pair = ((u64)__float_as_uint(x[i]) << 32) | ((u32)key[i] & (PB_BIN - 1));
            full[i] = fill[bin] & ~(PB_CHUNK - 1);
            if (rank[i] < PB_CHUNK) {
                buf[bin * PB_CHUNK + rank[i]] = pair;
            } else if (rank[i] < full[i]) {
                const u64 pos = (u64)gpos[bin] + rank[i];
                if (pos < npairs) pairs[pos] = pair;
            } else {
                wait |= 1u << i;
            }
        }
        __syncthreads();
        if (more) pb_contrib<PBS_ITEMS>(c, key_n, x_n);
        for (u32 b = threadIdx.x; b < NB; b += PBS_TPB) {
            const u32 F = fill[b];
            if (F >= PB_CHUNK) {
                const u64 pos = gpos[b];
                if (pos + PB_CHUNK <= npairs) {
                    const uint4 *src = (const uint4 *)(buf + b * PB_CHUNK);
                    uint4 *dst = (uint4 *)(pairs + pos);
#pragma unroll
                    for (u32 k = 0; k < PB_CHUNK / 2; ++k) dst[k] = src[k];
                }
                gpos[b] = (u32)(pos + (F & ~(PB_CHUNK - 1)));
                fill[b] = F & (PB_CHUNK - 1);
            }
        }
        __syncthreads();
#pragma unroll
        for (u32 i = 0; i < PBS_ITEMS; ++i)
            if (wait & (1u << i)) {
                const u32 bin = (u32)key[i] >> PB_BIN_LOG;
                buf[bin * PB_CHUNK + (rank[i] - full[i])] =
                    ((u64)__float_as_uint(x[i]) << 32) | ((u32)key[i] & (PB_BIN - 1));
            }
        // the next step's ranks read fill (settled before the last barrier);
        // its LDS chunk writes come after its first barrier
    }
    __syncthreads();
    for (u32 b = threadIdx.x; b < NB; b += PBS_TPB) {
        const u32 F = fill[b];
        const u64 pos = gpos[b];
        if (F && pos + PB_CHUNK <= npairs)
            for (u32 k = 0; k < PB_CHUNK; ++k) pairs[pos + k] = k < F ? buf[b * PB_CHUNK + k] : 0ull;
    }
}

// chunks per bin: nch[b] = ceil(pairs of bin b / chunk)
__global__ void k_pb_nchunks(const u32 *__restrict__ off, u32 P, u32 NB, u32 chunk, u32 *__restrict__ nch) {
    for (u32 b = blockIdx.x * blockDim.x + threadIdx.x; b < NB; b += gridDim.x * blockDim.x) {
        const u32 n = off[(u64)(b + 1) * P] - off[(u64)b * P];
        nch[b] = (n + chunk - 1) / chunk;
    }
}

// acc[v] += x, compensated: acc[v] packs (hi, lo) as two FP32 halves of one
// 64-bit LDS word; the new hi is the FP32 sum hi + x and its exact rounding
// error (TwoSum) goes to lo, both installed by one compare-and-swap.  FP32
// add atomics on LDS run about 20x slower than integer ones here (0.33 vs
// 6.5 lanes per clock per CU measured on MI355X), the CAS form at 1.4
// (uniform addresses) to 0.25 (R-MAT-skewed) and keeps FP32 arithmetic.
__device__ __forceinline__ u64 add2(u64 cur, float x) {
    const float hi = __uint_as_float((u32)cur), lo = __uint_as_float((u32)(cur >> 32));
    const float s = hi + x;
    const float bp = s - hi;
    const float e = (hi - (s - bp)) + (x - bp);
    return ((u64)__float_as_uint(lo + e) << 32) | __float_as_uint(s);
}

// retry from a failed compare: cur is the word's value the CAS returned
__device__ __forceinline__ void lds_add2_retry(u64 *acc, u32 v, float x, u64 cur) {
    while (true) {
        const u64 prev = atomicCAS((unsigned long long *)&acc[v], (unsigned long long)cur,
                                   (unsigned long long)add2(cur, x));
        if (prev == cur) break;
        cur = prev;
    }
}

// chunk k of bin b (cb = exclusive scan of nch, cb[NB] = chunks) summed in
// LDS into partial[k * PB_BIN + v] = (hi, lo)
__global__ __launch_bounds__(PBA_TPB) void k_pb_acc(const u64 *__restrict__ pairs, const u32 *__restrict__ off, u32 P,
                                                   u32 NB, const u32 *__restrict__ cb, u32 chunk,
                                                   float2 *__restrict__ partial) {
    extern __shared__ u64 pb_acc[];  // PB_BIN words: (lo << 32) | hi
    const u32 k = blockIdx.x;
    if (k >= cb[NB]) return;  // the grid is an upper bound
    u32 a = 0, z = NB;        // last b with cb[b] <= k
    while (z - a > 1) {
        const u32 m = (a + z) >> 1;
        if (cb[m] <= k) a = m; else z = m;
    }
    const u32 b = a;
    const u64 e1 = off[(u64)(b + 1) * P];
    const u64 s0 = (u64)off[(u64)b * P] + (u64)(k - cb[b]) * chunk;
    const u64 s1 = min(e1, s0 + chunk);
    for (u32 v = threadIdx.x; v < PB_BIN; v += PBA_TPB) pb_acc[v] = 0ull;
    __syncthreads();
    for (u64 i0 = s0; i0 < s1; i0 += (u64)PBA_TPB * PB_ITEMS) {
        u64 pr[PB_ITEMS];
#pragma unroll
        for (u32 r = 0; r < PB_ITEMS; ++r) {
            const u64 i = i0 + (u64)r * PBA_TPB + threadIdx.x;
            pr[r] = pairs[i < s1 ? i : s1 - 1];  // s1 > s0: a chunk is never empty
            if (i >= s1) pr[r] = 0ull;           // x = +0 adds nothing
        }
        // all reads, then all first compare-and-swaps, then the retries of
        // the lanes that lost a race: two LDS round trips per step when
        // nothing collides
        u64 cur[PB_ITEMS];
#pragma unroll
        for (u32 r = 0; r < PB_ITEMS; ++r) cur[r] = pb_acc[(u32)pr[r] & (PB_BIN - 1)];
        u64 got[PB_ITEMS];
#pragma unroll
        for (u32 r = 0; r < PB_ITEMS; ++r) {
            const float x = __uint_as_float((u32)(pr[r] >> 32));
            got[r] = cur[r];
            if (x != 0.f)
                got[r] = atomicCAS((unsigned long long *)&pb_acc[(u32)pr[r] & (PB_BIN - 1)],
                                   (unsigned long long)cur[r], (unsigned long long)add2(cur[r], x));
        }
#pragma unroll
        for (u32 r = 0; r < PB_ITEMS; ++r)
            if (got[r] != cur[r])
                lds_add2_retry(pb_acc, (u32)pr[r] & (PB_BIN - 1), __uint_as_float((u32)(pr[r] >> 32)), got[r]);
    }
    __syncthreads();
    float2 *out = partial + (u64)k * PB_BIN;
    for (u32 v = threadIdx.x; v < PB_BIN; v += PBA_TPB) {
        const u64 w = pb_acc[v];
        out[v] = make_float2(__uint_as_float((u32)w), __uint_as_float((u32)(w >> 32)));
    }
}

// as k_pr_finish, with vertex v's sum taken from its bin's chunk partials
__global__ __launch_bounds__(PR_TPB) void k_pb_finish(const float2 *__restrict__ partial, const u32 *__restrict__ cb,
                            const float *__restrict__ pr, float *__restrict__ out, u32 V, float alpha,
                            const float *__restrict__ red_in, float *__restrict__ red_out) {
    const float dang = red_in[0] + red_in[1];
    const float sum_in = red_in[2] + red_in[3];
    const float gamma = (alpha * dang + (1.f - alpha) * sum_in) / (float)V;
    float dl = 0.f, sum = 0.f;
    for (u64 v = (u64)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (u64)gridDim.x * blockDim.x) {
        const u32 b = (u32)(v >> PB_BIN_LOG);
        const u32 k0 = cb[b], k1 = cb[b + 1];
        float a = 0.f, comp = 0.f;
        for (u32 k = k0; k < k1; ++k) {
            const float2 h = partial[(u64)k * PB_BIN + (v & (PB_BIN - 1))];
            float y = h.y - comp;  // lo first: the small term
            float t = a + y;
            comp = (t - a) - y;
            a = t;
            y = h.x - comp;
            t = a + y;
            comp = (t - a) - y;
            a = t;
        }
        float p = alpha * a + gamma;
        out[v] = p;
        dl += fabsf(p - pr[v]);
        sum += p;
    }
    __shared__ float tmp[PR_TPB / 64];
    dl = block_sum(dl, tmp);
    sum = block_sum(sum, tmp);
    if (threadIdx.x == 0) {
        atomic_add2(&red_out[4], &red_out[5], dl);
        atomic_add2(&red_out[2], &red_out[3], sum);
    }
}

bool pr_bins_fit(u32 V) { return ((u64)V + PB_BIN - 1) / PB_BIN <= PB_MAX_BINS; }

PrBins pr_bins_plan(const DevState &d) {
    PrBins b{};
    const u64 W = nseg_of(d) << d.log2S;
    u32 lw = 0;
    while ((1ull << lw) < W) ++lw;
    b.part_log = std::max<u32>(d.log2S, lw > 10 ? lw - 10 : 0);  // at most PB_MAX_PARTS parts
    b.parts = (u32)((W + (1ull << b.part_log) - 1) >> b.part_log);
    b.bins = (u32)(((u64)d.V + PB_BIN - 1) / PB_BIN);
    b.staged = b.bins <= PB_STAGE_MAX_BINS;
    return b;
}

u64 pr_bins_count_words(const PrBins &b) { return (u64)b.bins * b.parts + 1; }

hipError_t pr_bins_count(const DevState &d, const PrBins &b, u32 *off, u32 *scan_scratch, hipStream_t st) {
    if (b.parts > PB_MAX_PARTS || b.bins > PB_MAX_BINS || b.bins == 0) return hipErrorInvalidValue;
    k_pb_count<<<b.parts, PB_TPB, b.bins * sizeof(u32), st>>>(d, b.part_log, b.parts, b.bins,
                                                             b.staged ? PB_CHUNK - 1 : 0u, off);
    return device_excl_scan(off, (u64)b.bins * b.parts, off, scan_scratch, st);
}

void pr_bins_chunking(PrBins &b, u64 pairs) {
    b.pairs = pairs;
    b.chunk = (u32)std::max<u64>(4096, (pairs + 2047) / 2048);
    b.max_chunks = (u32)((pairs + b.chunk - 1) / b.chunk) + b.bins;
}

u64 pr_bins_partial_floats(const PrBins &b) { return 2ull * PB_BIN * b.max_chunks; }

hipError_t pr_bins_chunks(const PrBins &b, const u32 *off, u32 *cb, u32 *scan_scratch, hipStream_t st) {
    k_pb_nchunks<<<grid_for(b.bins, 256, 1024), 256, 0, st>>>(off, b.parts, b.bins, b.chunk, cb);
    return device_excl_scan(cb, b.bins, cb, scan_scratch, st);
}

hipError_t pr_iterate_binned(const DevState &d, const PrBins &b, const float *pr, const u32 *outdeg, float *c,
                             const u32 *off, const u32 *cb, u64 *pairs, float *partial, float *out, float alpha,
                             float *red, float *red_next, hipStream_t st) {
    const u32 V = d.V;
    k_pr_reset<<<1, 64, 0, st>>>(red_next);
    k_pr_prep<<<grid_for(V, PR_TPB, PR_GRID), PR_TPB, 0, st>>>(pr, outdeg, V, c, nullptr, 0, red);
    if (b.staged)
        k_pb_bin_staged<<<b.parts, PBS_TPB, b.bins * (PB_CHUNK * sizeof(u64) + 2 * sizeof(u32)), st>>>(
            d, b.part_log, b.parts, b.bins, off, c, pairs, b.pairs);
    else
        k_pb_bin<<<b.parts, PB_TPB, b.bins * sizeof(u32), st>>>(d, b.part_log, b.parts, b.bins, off, c, pairs,
                                                                b.pairs);
    if (b.max_chunks)
        k_pb_acc<<<b.max_chunks, PBA_TPB, PB_BIN * sizeof(u64), st>>>(pairs, off, b.parts, b.bins, cb, b.chunk,
                                                                          (float2 *)partial);
    k_pb_finish<<<grid_for(V, PR_TPB, PR_GRID), PR_TPB, 0, st>>>((const float2 *)partial, cb, pr, out, V, alpha, red,
                                                         red_next);
    return hipGetLastError();
}

}  // namespace pma
