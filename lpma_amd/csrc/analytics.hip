// analytics.hip -- BFS and PageRank straight over the gapped layout.
//
// The reference runs nvGraph on a CSR copy (test_algos.cu:44-90 PageRank,
// :136-186 BFS; test.cu:170-182 converts with to_csr_rpma first).  Here both
// read the PMA in place: row r's out-edges are the live slots between the
// markers of rows r-1 and r, i.e. slots (row_off[r], row_off[r+1]).
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
// with one thread per slot (the range lengths are scanned, each thread finds
// its vertex by binary search), so hub vertices are split over many threads.

__global__ void k_bfs_init(u32 *levels, u32 V, u32 src, u32 *queue, u32 *counters) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (u64)gridDim.x * blockDim.x)
        levels[i] = i == src ? 0u : INF_LEVEL;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        queue[0] = src;
        counters[0] = 0;  // next frontier size
        counters[1] = 0;  // live edges scanned
    }
}

__global__ void k_bfs_len(DevState d, const u32 *__restrict__ q, u32 qn, u32 *__restrict__ len) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < qn; i += (u64)gridDim.x * blockDim.x) {
        u64 lo, hi;
        slot_range(d, q[i], lo, hi);
        len[i] = (u32)(hi - lo);
    }
}

// Each workgroup expands a run of EXP_TPB consecutive slots of the scanned
// frontier.  The frontier entries covering the run are found with one
// binary search over the global offsets (by one lane) and staged in LDS;
// every lane then finds its entry by a short search there instead of a
// search over the whole frontier.
static constexpr u32 EXP_TPB = 256;
static constexpr u32 EXP_STAGE = 512;  // staged offsets per run (>= EXP_TPB + 1 covers every run of
                                       // non-empty entries; empty rows fall back to the global search)

__global__ __launch_bounds__(EXP_TPB) void k_bfs_expand(DevState d, const u32 *__restrict__ q, u32 qn,
                                                         const u32 *__restrict__ off, u32 *__restrict__ levels,
                                                         u32 level, u32 *__restrict__ qnext, u32 *__restrict__ counters) {
    __shared__ u32 s_off[EXP_STAGE];
    __shared__ u32 s_e0;
    const u64 total = off[qn];
    const u32 lane = lane_id();
    u32 scanned = 0;
    for (u64 t0 = (u64)blockIdx.x * EXP_TPB; t0 < total; t0 += (u64)gridDim.x * EXP_TPB) {
        if (threadIdx.x == 0) {
            // last entry with off <= t0
            u32 lo = 0, hi = qn;
            while (hi - lo > 1) {
                u32 m = (lo + hi) >> 1;
                if (off[m] <= t0) lo = m; else hi = m;
            }
            s_e0 = lo;
        }
        __syncthreads();
        const u32 e0 = s_e0;
        for (u32 k = threadIdx.x; k < EXP_STAGE; k += EXP_TPB) {
            u64 idx = (u64)e0 + k;
            s_off[k] = idx <= qn ? off[idx] : 0xFFFFFFFFu;
        }
        __syncthreads();
        const u64 t = t0 + threadIdx.x;
        u32 v = INF_LEVEL;
        bool found = false;
        if (t < total) {
            u32 ent;
            if ((u64)s_off[EXP_STAGE - 1] > t) {
                // last k with s_off[k] <= t (s_off[0] <= t0 <= t)
                u32 lo = 0, hi = EXP_STAGE - 1;
                while (hi - lo > 1) {
                    u32 m = (lo + hi) >> 1;
                    if ((u64)s_off[m] <= t) lo = m; else hi = m;
                }
                ent = e0 + lo;
            } else {
                u32 lo = e0 + EXP_STAGE - 1, hi = qn;
                while (hi - lo > 1) {
                    u32 m = (lo + hi) >> 1;
                    if (off[m] <= t) lo = m; else hi = m;
                }
                ent = lo;
            }
            u64 slo, shi;
            slot_range(d, q[ent], slo, shi);
            u64 key = d.keys[slo + (t - off[ent])];
            if (key != KEY_NONE) {
                ++scanned;
                v = (u32)key;
                if (levels[v] == INF_LEVEL && atomicCAS(&levels[v], INF_LEVEL, level + 1) == INF_LEVEL) found = true;
            }
        }
        u64 m = __ballot(found);
        if (m) {
            u32 first = (u32)__builtin_ctzll(m);
            u32 base = 0;
            if (lane == first) base = atomicAdd(&counters[0], (u32)__popcll(m));
            base = __shfl(base, first, 64);
            if (found) qnext[base + __popcll(m & ((1ull << lane) - 1ull))] = v;
        }
        __syncthreads();  // s_off / s_e0 are rewritten by the next run
    }
    scanned = wave_sum(scanned);
    if (lane == 0 && scanned) atomicAdd(&counters[1], scanned);
}

hipError_t bfs_init(u32 *levels, u32 V, u32 src, u32 *queue, u32 *counters, hipStream_t st) {
    k_bfs_init<<<grid_for(V, 256, 16384), 256, 0, st>>>(levels, V, src, queue, counters);
    return hipGetLastError();
}

hipError_t bfs_level(const DevState &d, const u32 *q, u32 qn, u32 *len, u32 *scan_scratch, u32 *levels, u32 level,
                     u32 *qnext, u32 *counters, u32 expand_grid, hipStream_t st) {
    k_bfs_len<<<grid_for(qn, 256, 16384), 256, 0, st>>>(d, q, qn, len);
    (void)device_excl_scan(len, qn, len, scan_scratch, st);
    k_bfs_expand<<<expand_grid ? expand_grid : 2048, EXP_TPB, 0, st>>>(d, q, qn, len, levels, level, qnext, counters);
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
__global__ void k_pr_prep(const float *__restrict__ pr, const u32 *__restrict__ outdeg, u32 V, float *__restrict__ c,
                          float *__restrict__ acc, float *__restrict__ red) {
    float dang = 0.f;
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 u = (u64)blockIdx.x * blockDim.x + threadIdx.x; u < V; u += stride) {
        u32 od = outdeg[u];
        float p = pr[u];
        c[u] = od ? p / (float)od : 0.f;
        if (!od) dang += p;
    }
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < 2ull * PR_COPIES * V; i += stride) acc[i] = 0.f;
    dang = wave_sum(dang);
    if (lane_id() == 0 && dang != 0.f) atomic_add2(&red[0], &red[1], dang);
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
__global__ void k_pr_finish(const float *__restrict__ acc, const float *__restrict__ pr, float *__restrict__ out,
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
    dl = wave_sum(dl);
    sum = wave_sum(sum);
    if (lane_id() == 0) {
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
    k_pr_prep<<<grid_for(pr_acc_floats(V), 256, 16384), 256, 0, st>>>(pr, outdeg, V, c, acc, red);
    k_pr_push<<<grid_for(nseg_of(d) << d.log2S, 256, 65536), 256, 0, st>>>(d, c, acc);
    k_pr_finish<<<grid_for(V, 256, 16384), 256, 0, st>>>(acc, pr, out, V, alpha, red, red_next);
    return hipGetLastError();
}

hipError_t pr_scale(float *pr, u32 V, const float *red, hipStream_t st) {
    k_pr_scale<<<grid_for(V, 256, 16384), 256, 0, st>>>(pr, V, red);
    return hipGetLastError();
}

}  // namespace pma
