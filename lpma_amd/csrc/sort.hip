// sort.hip -- batch packing, LDS-staged wave64 LSD radix sort, and the
// device-wide u32 exclusive scan used across the library.
//
// Replaces thrust::sort_by_key over the batch (rpma.cuh:1557,
// gpma.cuh:751).  The batch key src<<32|dst and its operation are packed
// into ((src << vbits | dst) << 1 | op), vbits = ceil(log2 V), so only
// 2*vbits+1 bits are sorted (41 bits, 6 passes, for V = 2^20), and equal
// keys end up with the delete before the insert; the dedup that follows
// keeps the last of each run, which makes a batch "deletes, then inserts".
#include "common.hpp"
#include "kernels.hpp"

namespace pma {

static constexpr int RS_BITS = 8;
static constexpr int RS_RADIX = 1 << RS_BITS;
static constexpr int RS_TPB = 256;
static constexpr int RS_ITEMS = 16;
static constexpr int RS_TILE = RS_TPB * RS_ITEMS;
static constexpr int RS_SMALL_ITEMS = 8;
static constexpr int RS_SMALL_CAP = RS_TPB * RS_SMALL_ITEMS;  // single-workgroup sort

__global__ void k_pack(const u64 *__restrict__ in, u64 n, u32 op, u32 V, u32 vbits,
                       u64 *__restrict__ out, u32 *__restrict__ flags) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    bool bad = false;
    for (; i < n; i += (u64)gridDim.x * blockDim.x) {
        u64 k = in[i];
        u32 src = (u32)(k >> 32), dst = (u32)k;
        if (src >= V || dst >= V) { bad = true; src = 0; dst = 0; }
        out[i] = (((((u64)src) << vbits) | dst) << 1) | op;
    }
    if (__any(bad) && lane_id() == 0) atomicOr(flags, FLAG_INVALID_KEY);
}

// ------------------------------------------------------------ stable ranking
//
// One round ranks RS_TPB consecutive keys of a tile by their digit: lanes
// with equal digits are found with 8 wave ballots (a wave64 "match any"),
// a lane's rank inside its wave is the popcount of its lower peers, and
// per-wave digit counts in LDS give the offset of the waves before it.

struct RankLds {
    u32 running[RS_RADIX];
    u32 wcnt[RS_TPB / 64][RS_RADIX];
};

__device__ __forceinline__ void rank_round(bool valid, u32 digit, RankLds &s, u32 &pos) {
    const u32 lane = lane_id();
    const u32 w = threadIdx.x >> 6;
    u64 peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RS_BITS; ++b) {
        u32 bit = (digit >> b) & 1u;
        u64 bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
    }
    u32 rank = __popcll(peers & ((1ull << lane) - 1ull));
    if (valid && rank == 0) s.wcnt[w][digit] = __popcll(peers);
    __syncthreads();
    if (valid) {
        u32 off = s.running[digit];
        for (u32 v = 0; v < w; ++v) off += s.wcnt[v][digit];
        pos = off + rank;
    }
    __syncthreads();
    {
        u32 d = threadIdx.x;  // RS_TPB == RS_RADIX
        u32 add = 0;
#pragma unroll
        for (int v = 0; v < RS_TPB / 64; ++v) { add += s.wcnt[v][d]; s.wcnt[v][d] = 0; }
        s.running[d] += add;
    }
    __syncthreads();
}

__global__ __launch_bounds__(RS_TPB) void k_rs_hist(const u64 *__restrict__ keys, u64 n, u32 shift,
                                                     u32 *__restrict__ hist, u32 nblocks) {
    __shared__ u32 h[RS_RADIX];
    h[threadIdx.x] = 0;
    __syncthreads();
    u64 base = (u64)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int r = 0; r < RS_ITEMS; ++r) {
        u64 i = base + (u64)r * RS_TPB + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (RS_RADIX - 1)], 1u);
    }
    __syncthreads();
    hist[(u64)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

__global__ __launch_bounds__(RS_TPB) void k_rs_scatter(const u64 *__restrict__ in, u64 *__restrict__ out, u64 n,
                                                        u32 shift, const u32 *__restrict__ hist_scan,
                                                        u32 nblocks) {
    __shared__ RankLds s;
    __shared__ u32 gofs[RS_RADIX];
    gofs[threadIdx.x] = hist_scan[(u64)threadIdx.x * nblocks + blockIdx.x];
    s.running[threadIdx.x] = 0;
#pragma unroll
    for (int v = 0; v < RS_TPB / 64; ++v) s.wcnt[v][threadIdx.x] = 0;
    __syncthreads();
    u64 base = (u64)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; ++r) {
        u64 i = base + (u64)r * RS_TPB + threadIdx.x;
        bool valid = i < n;
        u64 key = valid ? in[i] : 0;
        u32 digit = (u32)(key >> shift) & (RS_RADIX - 1);
        u32 pos = 0;
        rank_round(valid, digit, s, pos);
        if (valid) out[gofs[digit] + pos] = key;
    }
}

// Whole sort inside one workgroup's LDS for n <= RS_SMALL_CAP.
__global__ __launch_bounds__(RS_TPB) void k_rs_small(u64 *__restrict__ keys, u64 n, u32 nbits) {
    __shared__ u64 buf[2][RS_SMALL_CAP];
    __shared__ RankLds s;
    __shared__ u32 h[RS_RADIX];
    __shared__ u32 scan_tmp[RS_TPB / 64 + 1];
    for (u32 i = threadIdx.x; i < n; i += RS_TPB) buf[0][i] = keys[i];
    int cur = 0;
    for (u32 shift = 0; shift < nbits; shift += RS_BITS) {
        h[threadIdx.x] = 0;
        s.running[threadIdx.x] = 0;
#pragma unroll
        for (int v = 0; v < RS_TPB / 64; ++v) s.wcnt[v][threadIdx.x] = 0;
        __syncthreads();
        for (u32 i = threadIdx.x; i < n; i += RS_TPB)
            atomicAdd(&h[(buf[cur][i] >> shift) & (RS_RADIX - 1)], 1u);
        __syncthreads();
        u32 tot;
        u32 ex = block_excl_scan<RS_TPB, u32>(h[threadIdx.x], tot, scan_tmp);
        __syncthreads();
        h[threadIdx.x] = ex;
        __syncthreads();
        for (u32 r = 0; r * RS_TPB < n; ++r) {
            u32 i = r * RS_TPB + threadIdx.x;
            bool valid = i < n;
            u64 key = valid ? buf[cur][i] : 0;
            u32 digit = (u32)(key >> shift) & (RS_RADIX - 1);
            u32 pos = 0;
            rank_round(valid, digit, s, pos);
            if (valid) buf[cur ^ 1][h[digit] + pos] = key;
        }
        __syncthreads();
        cur ^= 1;
    }
    for (u32 i = threadIdx.x; i < n; i += RS_TPB) keys[i] = buf[cur][i];
}

// ------------------------------------------------------ device-wide scan

static constexpr int SC_TPB = 256;
static constexpr int SC_ITEMS = 8;
static constexpr int SC_TILE = SC_TPB * SC_ITEMS;

__global__ __launch_bounds__(SC_TPB) void k_scan_reduce(const u32 *__restrict__ in, u64 n, u32 *__restrict__ part) {
    u64 base = (u64)blockIdx.x * SC_TILE;
    u32 sum = 0;
#pragma unroll
    for (int r = 0; r < SC_ITEMS; ++r) {
        u64 i = base + (u64)r * SC_TPB + threadIdx.x;
        if (i < n) sum += in[i];
    }
    __shared__ u32 tmp[SC_TPB / 64 + 1];
    u32 tot;
    block_excl_scan<SC_TPB, u32>(sum, tot, tmp);
    if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

// exclusive scan of part[0..np) in place by one workgroup; part[np] = total
__global__ __launch_bounds__(1024) void k_scan_partials(u32 *__restrict__ part, u32 np) {
    __shared__ u32 tmp[1024 / 64 + 1];
    u32 carry = 0;
    for (u32 b = 0; b < np; b += 1024) {
        u32 i = b + threadIdx.x;
        u32 v = i < np ? part[i] : 0;
        u32 tot;
        u32 ex = block_excl_scan<1024, u32>(v, tot, tmp);
        if (i < np) part[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) part[np] = carry;
}

__global__ __launch_bounds__(SC_TPB) void k_scan_down(const u32 *__restrict__ in, u64 n, const u32 *__restrict__ part,
                                                      u32 *__restrict__ out) {
    __shared__ u32 tile[SC_TILE];
    __shared__ u32 tmp[SC_TPB / 64 + 1];
    u64 base = (u64)blockIdx.x * SC_TILE;
#pragma unroll
    for (int r = 0; r < SC_ITEMS; ++r) {
        u64 i = base + (u64)r * SC_TPB + threadIdx.x;
        tile[r * SC_TPB + threadIdx.x] = i < n ? in[i] : 0;
    }
    __syncthreads();
    // each thread owns SC_ITEMS consecutive entries
    u32 loc[SC_ITEMS];
    u32 sum = 0;
#pragma unroll
    for (int r = 0; r < SC_ITEMS; ++r) { loc[r] = tile[threadIdx.x * SC_ITEMS + r]; sum += loc[r]; }
    u32 tot;
    u32 ex = block_excl_scan<SC_TPB, u32>(sum, tot, tmp) + part[blockIdx.x];
#pragma unroll
    for (int r = 0; r < SC_ITEMS; ++r) { tile[threadIdx.x * SC_ITEMS + r] = ex; ex += loc[r]; }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SC_ITEMS; ++r) {
        u64 i = base + (u64)r * SC_TPB + threadIdx.x;
        if (i < n) out[i] = tile[r * SC_TPB + threadIdx.x];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = part[gridDim.x];
}

u64 scan_scratch_words(u64 n) { return (n + SC_TILE - 1) / SC_TILE + 2; }

// out[0..n) = exclusive prefix of in, out[n] = total.  out may alias in.
hipError_t device_excl_scan(const u32 *in, u64 n, u32 *out, u32 *scratch, hipStream_t st) {
    u32 nb = (u32)((n + SC_TILE - 1) / SC_TILE);
    if (nb == 0) nb = 1;
    k_scan_reduce<<<nb, SC_TPB, 0, st>>>(in, n, scratch);
    k_scan_partials<<<1, 1024, 0, st>>>(scratch, nb);
    k_scan_down<<<nb, SC_TPB, 0, st>>>(in, n, scratch, out);
    return hipGetLastError();
}

// ------------------------------------------------------------ radix driver

u64 sort_scratch_words(u64 n) {
    u64 nb = (n + RS_TILE - 1) / RS_TILE;
    return (u64)RS_RADIX * nb + 1 + scan_scratch_words((u64)RS_RADIX * nb);
}

hipError_t pack_batch(const u64 *in, u64 n, u32 op, u32 V, u32 vbits, u64 *out, u32 *flags, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_pack<<<grid_for(n, 256, 8192), 256, 0, st>>>(in, n, op, V, vbits, out, flags);
    return hipGetLastError();
}

// Sorts keys[0..n) on bits [0, nbits).  Returns the buffer that holds the
// result (keys or alt).
u64 *radix_sort(u64 *keys, u64 *alt, u64 n, u32 nbits, u32 *scratch, hipStream_t st) {
    if (n <= 1) return keys;
    if (n <= (u64)RS_SMALL_CAP) {
        k_rs_small<<<1, RS_TPB, 0, st>>>(keys, n, nbits);
        return keys;
    }
    u32 nb = (u32)((n + RS_TILE - 1) / RS_TILE);
    u32 *hist = scratch;
    u32 *sscr = scratch + (u64)RS_RADIX * nb + 1;
    u64 *src = keys, *dst = alt;
    for (u32 shift = 0; shift < nbits; shift += RS_BITS) {
        k_rs_hist<<<nb, RS_TPB, 0, st>>>(src, n, shift, hist, nb);
        device_excl_scan(hist, (u64)RS_RADIX * nb, hist, sscr, st);
        k_rs_scatter<<<nb, RS_TPB, 0, st>>>(src, dst, n, shift, hist, nb);
        u64 *t = src; src = dst; dst = t;
    }
    return src;
}

}  // namespace pma
