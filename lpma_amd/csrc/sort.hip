// sort.hip -- batch packing, LDS-staged wave64 LSD radix sort, and the
// device-wide u32 exclusive scan used across the library.
//
// Replaces thrust::sort_by_key over the batch (rpma.cuh:1557,
// gpma.cuh:751).  The batch key src<<32|dst and its operation are packed
// into ((src << vbits | dst) << 1 | op), vbits = ceil(log2 V), so only
// 2*vbits+1 bits are sorted (41 bits, 6 passes, for V = 2^20), and equal
// keys end up with the delete before the insert; the dedup that follows
// keeps the last of each run, which makes a batch "deletes, then inserts".
// A group of PMAs sorts as one batch: the PMA id sits above the packed key.
#include "common.hpp"
#include "kernels.hpp"

namespace pma {

static constexpr int RS_BITS = 8;
static constexpr int RS_RADIX = 1 << RS_BITS;
static constexpr int RS_TPB = 256;
static constexpr int RS_ITEMS_BIG = 16;   // keys per thread per tile, large inputs
static constexpr int RS_ITEMS_SMALL = 4;  // smaller tiles for more workgroups on mid-size batches
static constexpr int RS_SMALL_TPB = 1024;
static constexpr int RS_SMALL_CAP = 8192;  // single-workgroup sort (2 x 64 KiB of LDS)
static constexpr u32 RS_PREFIX_MAX_TILES = 512;  // scatter derives its offsets itself up to here

// Batch element i comes from the source whose [off, off + n) holds it.
// Sources are few and sorted by off; each thread finds its first source by
// binary search and walks forward from there.
__device__ __forceinline__ u32 source_of(const BatchSeg *__restrict__ segs, u32 nsegs, u64 i) {
    u32 lo = 0, hi = nsegs;
    while (hi - lo > 1) {
        u32 mid = (lo + hi) >> 1;
        if (segs[mid].off <= i) lo = mid; else hi = mid;
    }
    return lo;
}

// Packs a group batch into sort keys pid << pshift | (src << vbits | dst) << 1 | op
// (keys with src or dst >= V raise FLAG_INVALID_KEY) and, with HIST, counts
// the first radix digit of every tile.
template <bool HIST>
__global__ __launch_bounds__(RS_TPB) void k_pack_group(const BatchSeg *__restrict__ segs, u32 nsegs, u64 n,
                                                        const u64 *n_dev, u32 pshift, u32 sbits, u64 *__restrict__ out,
                                                        u32 *__restrict__ flags, u32 shift, u32 items,
                                                        u32 *__restrict__ hist, u32 nblocks, int block_major) {
    __shared__ u32 h[RS_RADIX];
    if (n_dev) n = *n_dev;
    if (HIST) {
        h[threadIdx.x] = 0;
        __syncthreads();
    }
    const u64 base = (u64)blockIdx.x * items * RS_TPB + threadIdx.x;
    bool bad = false;
    if (base < n) {
        u32 s = source_of(segs, nsegs, base);
        BatchSeg sg = segs[s];
        for (u32 r = 0; r < items; ++r) {
            u64 i = base + (u64)r * RS_TPB;
            if (i >= n) break;
            while (i >= sg.off + sg.n) sg = segs[++s];
            u64 k = sg.keys[i - sg.off];
            u32 src = (u32)(k >> 32), dst = (u32)k;
            if (src >= sg.V || dst >= sg.V) { bad = true; src = 0; dst = 0; }
            u64 p = ((u64)sg.pid << pshift) | ((((((u64)src) << sg.vbits) | dst) << sbits | sg.seq) << 1) | sg.op;
            out[i] = p;
            if (HIST) atomicAdd(&h[(p >> shift) & (RS_RADIX - 1)], 1u);
        }
    }
    if (__any(bad) && lane_id() == 0) atomicOr(flags, FLAG_INVALID_KEY);
    if (HIST) {
        __syncthreads();
        if (block_major) hist[(u64)blockIdx.x * RS_RADIX + threadIdx.x] = h[threadIdx.x];
        else hist[(u64)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
    }
}

// ------------------------------------------------------------ stable ranking
//
// One round ranks NT consecutive keys of a tile by their digit: lanes with
// equal digits are found with 8 wave ballots (a wave64 "match any"), a
// lane's rank inside its wave is the popcount of its lower peers, and
// per-wave digit counts in LDS give the offset of the waves before it.

template <int NT>
struct RankLds {
    u32 running[RS_RADIX];
    u32 wcnt[NT / 64][RS_RADIX];
};

template <int NT>
__device__ __forceinline__ void rank_init(RankLds<NT> &s) {
    for (u32 i = threadIdx.x; i < RS_RADIX; i += NT) s.running[i] = 0;
    for (u32 i = threadIdx.x; i < (NT / 64) * RS_RADIX; i += NT) (&s.wcnt[0][0])[i] = 0;
}

template <int NT>
__device__ __forceinline__ void rank_round(bool valid, u32 digit, RankLds<NT> &s, u32 &pos) {
    const u32 lane = lane_id();
    const u32 w = threadIdx.x >> 6;
    u64 peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RS_BITS; ++b) {
        u32 bit = (digit >> b) & 1u;
        u64 bb = __ballot(bit);
        peers &= bit ? bb : ~bb;
    }
    const u32 rank = __popcll(peers & ((1ull << lane) - 1ull));
    const bool leader = valid && rank == 0;
    const u32 cnt = __popcll(peers);
    if (leader) s.wcnt[w][digit] = cnt;
    __syncthreads();
    if (valid) {
        u32 off = s.running[digit];
        for (u32 v = 0; v < w; ++v) off += s.wcnt[v][digit];
        pos = off + rank;
    }
    __syncthreads();
    // only the entries written this round change: their leaders fold them in
    if (leader) {
        atomicAdd(&s.running[digit], cnt);
        s.wcnt[w][digit] = 0;
    }
    __syncthreads();
}

// per-tile digit histogram; layout block-major (hist[b*RADIX + d]) when
// `block_major`, digit-major (hist[d*nb + b]) otherwise
__global__ __launch_bounds__(RS_TPB) void k_rs_hist(const u64 *__restrict__ keys, u64 n, const u64 *n_dev, u32 shift,
                                                     u32 items, u32 *__restrict__ hist, u32 nblocks, int block_major) {
    __shared__ u32 h[RS_RADIX];
    if (n_dev) n = *n_dev;
    h[threadIdx.x] = 0;
    __syncthreads();
    u64 base = (u64)blockIdx.x * items * RS_TPB;
    for (u32 r = 0; r < items; ++r) {
        u64 i = base + (u64)r * RS_TPB + threadIdx.x;
        if (i < n) atomicAdd(&h[(keys[i] >> shift) & (RS_RADIX - 1)], 1u);
    }
    __syncthreads();
    if (block_major) hist[(u64)blockIdx.x * RS_RADIX + threadIdx.x] = h[threadIdx.x];
    else hist[(u64)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Scatter of one pass.  With block-major histograms (few tiles) every
// workgroup derives its digit offsets itself: digit totals and the prefix
// over earlier tiles from the histogram rows, then a scan over the digits;
// otherwise hist_scan holds the device-wide exclusive scan (digit-major).
__global__ __launch_bounds__(RS_TPB) void k_rs_scatter(const u64 *__restrict__ in, u64 *__restrict__ out, u64 n,
                                                        const u64 *n_dev, u32 shift, u32 items,
                                                        const u32 *__restrict__ hist, u32 nblocks, int block_major) {
    __shared__ RankLds<RS_TPB> s;
    __shared__ u32 gofs[RS_RADIX];
    __shared__ u32 stmp[RS_TPB / 64 + 1];
    if (n_dev) n = *n_dev;
    const u32 d = threadIdx.x;
    if (block_major) {
        u32 tot = 0, pre = 0;
#pragma unroll 16
        for (u32 b = 0; b < nblocks; ++b) {
            u32 v = hist[(u64)b * RS_RADIX + d];
            tot += v;
            pre += b < blockIdx.x ? v : 0u;
        }
        u32 all;
        u32 dbase = block_excl_scan<RS_TPB, u32>(tot, all, stmp);
        gofs[d] = dbase + pre;
    } else {
        gofs[d] = hist[(u64)d * nblocks + blockIdx.x];
    }
    rank_init<RS_TPB>(s);
    __syncthreads();
    u64 base = (u64)blockIdx.x * items * RS_TPB;
    for (u32 r = 0; r < items; ++r) {
        u64 i = base + (u64)r * RS_TPB + threadIdx.x;
        bool valid = i < n;
        u64 key = valid ? in[i] : 0;
        u32 digit = (u32)(key >> shift) & (RS_RADIX - 1);
        u32 pos = 0;
        rank_round<RS_TPB>(valid, digit, s, pos);
        if (valid) out[gofs[digit] + pos] = key;
    }
}

// Whole sort inside one workgroup's LDS for n <= RS_SMALL_CAP.
__global__ __launch_bounds__(RS_SMALL_TPB) void k_rs_small(u64 *__restrict__ keys, u64 n, const u64 *n_dev,
                                                             u32 lo_bit, u32 hi_bit) {
    __shared__ u64 buf[2][RS_SMALL_CAP];
    __shared__ RankLds<RS_SMALL_TPB> s;
    __shared__ u32 h[RS_RADIX];
    if (n_dev) n = *n_dev;
    for (u32 i = threadIdx.x; i < n; i += RS_SMALL_TPB) buf[0][i] = keys[i];
    int cur = 0;
    for (u32 shift = lo_bit; shift < hi_bit; shift += RS_BITS) {
        for (u32 i = threadIdx.x; i < RS_RADIX; i += RS_SMALL_TPB) h[i] = 0;
        rank_init<RS_SMALL_TPB>(s);
        __syncthreads();
        for (u32 i = threadIdx.x; i < n; i += RS_SMALL_TPB)
            atomicAdd(&h[(buf[cur][i] >> shift) & (RS_RADIX - 1)], 1u);
        __syncthreads();
        if (threadIdx.x < 64) {
            // one wave scans the 256 digit counts
            u32 v[RS_RADIX / 64];
            u32 sum = 0;
#pragma unroll
            for (int r = 0; r < RS_RADIX / 64; ++r) { v[r] = h[threadIdx.x * (RS_RADIX / 64) + r]; sum += v[r]; }
            u32 ex = wave_incl_scan(sum) - sum;
#pragma unroll
            for (int r = 0; r < RS_RADIX / 64; ++r) { h[threadIdx.x * (RS_RADIX / 64) + r] = ex; ex += v[r]; }
        }
        __syncthreads();
        for (u32 r = 0; r * RS_SMALL_TPB < n; ++r) {
            u32 i = r * RS_SMALL_TPB + threadIdx.x;
            bool valid = i < n;
            u64 key = valid ? buf[cur][i] : 0;
            u32 digit = (u32)(key >> shift) & (RS_RADIX - 1);
            u32 pos = 0;
            rank_round<RS_SMALL_TPB>(valid, digit, s, pos);
            if (valid) buf[cur ^ 1][h[digit] + pos] = key;
        }
        __syncthreads();
        cur ^= 1;
    }
    for (u32 i = threadIdx.x; i < n; i += RS_SMALL_TPB) keys[i] = buf[cur][i];
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

// small tiles keep >= ~2000 workgroups in flight up to 8M keys
static u32 rs_items(u64 n) { return n <= (1ull << 23) ? RS_ITEMS_SMALL : RS_ITEMS_BIG; }

u64 sort_scratch_words(u64 n) {
    u64 nb = (n + (u64)rs_items(n) * RS_TPB - 1) / ((u64)rs_items(n) * RS_TPB);
    return (u64)RS_RADIX * nb + 1 + scan_scratch_words((u64)RS_RADIX * nb);
}

// Packs a group batch (n elements from nsegs sources) and sorts the packed
// keys on bits [lo_bit, hi_bit): the first histogram is taken while packing.
// With n_dev the kernels read the element count from device memory and n
// is only a capacity that sizes the grids (graph replays).
u64 *pack_and_sort(const BatchSeg *segs, u32 nsegs, u64 n, const u64 *n_dev, u32 pshift, u32 sbits, u32 lo_bit,
                   u32 hi_bit, u64 *k0, u64 *k1, u32 *scratch, u32 *flags, hipStream_t st) {
    if (n == 0) return k0;
    if (n <= (u64)RS_SMALL_CAP) {
        const u32 items = 4;
        k_pack_group<false><<<(u32)((n + items * RS_TPB - 1) / (items * RS_TPB)), RS_TPB, 0, st>>>(
            segs, nsegs, n, n_dev, pshift, sbits, k0, flags, 0, items, nullptr, 0, 0);
        if (n > 1 || n_dev) k_rs_small<<<1, RS_SMALL_TPB, 0, st>>>(k0, n, n_dev, lo_bit, hi_bit);
        return k0;
    }
    const u32 items = rs_items(n);
    const u64 tile = (u64)items * RS_TPB;
    const u32 nblk = (u32)((n + tile - 1) / tile);
    const int block_major = nblk <= RS_PREFIX_MAX_TILES;
    u32 *hist = scratch;
    u32 *sscr = scratch + (u64)RS_RADIX * nblk + 1;
    u64 *src = k0, *dst = k1;
    for (u32 shift = lo_bit; shift < hi_bit; shift += RS_BITS) {
        if (shift == lo_bit)
            k_pack_group<true><<<nblk, RS_TPB, 0, st>>>(segs, nsegs, n, n_dev, pshift, sbits, k0, flags, shift, items,
                                                       hist, nblk, block_major);
        else
            k_rs_hist<<<nblk, RS_TPB, 0, st>>>(src, n, n_dev, shift, items, hist, nblk, block_major);
        if (!block_major) (void)device_excl_scan(hist, (u64)RS_RADIX * nblk, hist, sscr, st);
        k_rs_scatter<<<nblk, RS_TPB, 0, st>>>(src, dst, n, n_dev, shift, items, hist, nblk, block_major);
        u64 *t = src; src = dst; dst = t;
    }
    return src;
}

}  // namespace pma
