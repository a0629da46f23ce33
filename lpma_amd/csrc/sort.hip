// sort.hip -- batch packing and sorting on gfx950, and the device-wide u32
// exclusive scan used across the library.
//
// Replaces thrust::sort_by_key over the batch (rpma.cuh:1557,
// gpma.cuh:751).  The batch key src<<32|dst and its operation are packed
// into ((src << vbits | dst) << 1 | op), vbits = ceil(log2 V), with the PMA id
// above it for a group and the request order between key and operation for a
// merged sequence, so equal keys end up ordered delete before insert, earlier
// request before later; the dedup that follows keeps the last of each run.
// Three paths by batch size:
//   <= 2^17 keys   sample-split sort: two kernels, buckets chosen from a
//                  sample, each sorted in LDS on the bits that vary in it;
//   <= 2^22 keys   LSD radix, one kernel per 10/11-bit digit (decoupled
//                  look-back), the digit counts of every pass taken at pack;
//   larger         LSD radix, 8-bit digits, histogram + scan + scatter.
#include "common.hpp"
#include "kernels.hpp"
#include "begin.hpp"

namespace pma {

static constexpr int RS_BITS = 8;
static constexpr int RS_RADIX = 1 << RS_BITS;
static constexpr int RS_TPB = 256;
static constexpr int RS_ITEMS_BIG = 16;   // keys per thread per tile, large inputs
static constexpr int RS_ITEMS_SMALL = 4;  // smaller tiles for more workgroups on mid-size batches

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
// (keys with src or dst >= V raise FLAG_INVALID_KEY) and counts the first
// radix digit of every tile (digit-major: hist[d * nblocks + tile]).
__global__ __launch_bounds__(RS_TPB) void k_pack_group(const BatchSeg *__restrict__ segs, u32 nsegs, u64 n,
                                                        const u64 *n_dev, u32 pshift, u32 sbits, u64 *__restrict__ out,
                                                        u32 *__restrict__ flags, u32 shift, u32 items,
                                                        u32 *__restrict__ hist, u32 nblocks) {
    __shared__ u32 h[RS_RADIX];
    if (n_dev) n = *n_dev;
    h[threadIdx.x] = 0;
    __syncthreads();
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
            atomicAdd(&h[(p >> shift) & (RS_RADIX - 1)], 1u);
        }
    }
    if (__any(bad) && lane_id() == 0) atomicOr(flags, FLAG_INVALID_KEY);
    __syncthreads();
    hist[(u64)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
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

// per-tile digit histogram, digit-major (hist[d * nblocks + tile])
__global__ __launch_bounds__(RS_TPB) void k_rs_hist(const u64 *__restrict__ keys, u64 n, const u64 *n_dev, u32 shift,
                                                     u32 items, u32 *__restrict__ hist, u32 nblocks) {
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
    hist[(u64)threadIdx.x * nblocks + blockIdx.x] = h[threadIdx.x];
}

// Scatter of one pass; hist holds the device-wide exclusive scan of the
// digit-major histograms (each tile's offset of each digit).
__global__ __launch_bounds__(RS_TPB) void k_rs_scatter(const u64 *__restrict__ in, u64 *__restrict__ out, u64 n,
                                                        const u64 *n_dev, u32 shift, u32 items,
                                                        const u32 *__restrict__ hist, u32 nblocks) {
    __shared__ RankLds<RS_TPB> s;
    __shared__ u32 gofs[RS_RADIX];
    if (n_dev) n = *n_dev;
    gofs[threadIdx.x] = hist[(u64)threadIdx.x * nblocks + blockIdx.x];
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

// ------------------------------------------------- one kernel per pass
//
// Batches of 8K..4M keys sort with DB = 10 or 11-bit digits (four passes
// for the 40-41 bits of a lone PMA with 2^20 vertices), one kernel each,
// over tiles of 4096 keys (1024 threads, 4 keys each): few enough tiles that
// a tile finds the counts of all tiles before it in few look-back steps,
// small enough that a tile's 16 waves rank it in 4 rounds each:
// the pack kernel counts the digits of every pass at once (global digit
// totals), and a pass kernel ranks its tile in LDS, publishes the tile's
// digit counts, reads the counts of the tiles before it (decoupled look-back,
// Merrill & Garland, "Single-pass Parallel Prefix Scan with Decoupled
// Look-back", NVR-2016-002) and scatters.  A published word is the count in
// bits 0..29 with OS_AGG (this tile only) or OS_PRE (this tile and every
// tile before it) in bits 30..31; 0 means not yet published.
//
// Tile t waits only on tiles < t.  Workgroups are dispatched in index order,
// so every tile it waits on is running or done; the wait is still bounded
// (OS_SPIN_LIMIT rounds), after which the batch is flagged FLAG_SORT_STALL
// and fails without changing anything.

static constexpr int OS_TPB = 1024;  // large tiles: few tiles, short look-back walks
static constexpr u32 OS_AGG = 1u << 30;
static constexpr u32 OS_PRE = 2u << 30;
static constexpr u32 OS_CNT = (1u << 30) - 1;
static constexpr u32 OS_SPIN_LIMIT = 1u << 20;
static constexpr u64 OS_MAX_N = 1ull << 22;       // larger batches: the two-kernel passes
static constexpr u32 OS_GH_WORDS = SORT_MAX_PASSES * 2048;
static constexpr int OS_ITEMS = 4;   // keys per thread: tiles of 4096
static constexpr int OS_LOOK = 32;   // look-back loads in flight per thread

__device__ __forceinline__ u32 os_load(const u32 *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void os_store(u32 *p, u32 v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pack (as k_pack_group) and count the digits of all npass passes:
// gh[p * R + d] += keys whose pass-p digit is d (gh zeroed by the caller);
// also clears the look-back row of this tile for the first pass.
template <int DB>
__global__ __launch_bounds__(OS_TPB) void k_pack_count(const BatchSeg *__restrict__ segs, u32 nsegs, u64 n,
                                                        const u64 *n_dev, u32 pshift, u32 sbits, u64 *__restrict__ out,
                                                        u32 *__restrict__ flags, u32 lo_bit, u32 npass, u32 items,
                                                        u32 *__restrict__ gh, u32 *__restrict__ look0) {
    constexpr u32 R = 1u << DB;
    __shared__ u32 h[SORT_MAX_PASSES][R];
    if (n_dev) n = *n_dev;
    for (u32 i = threadIdx.x; i < npass * R; i += OS_TPB) (&h[0][0])[i] = 0;
    for (u32 i = threadIdx.x; i < R; i += OS_TPB) look0[(u64)blockIdx.x * R + i] = 0;
    __syncthreads();
    const u64 base = (u64)blockIdx.x * items * OS_TPB + threadIdx.x;
    bool bad = false;
    if (base < n) {
        u32 s = source_of(segs, nsegs, base);
        BatchSeg sg = segs[s];
        for (u32 r = 0; r < items; ++r) {
            u64 i = base + (u64)r * OS_TPB;
            if (i >= n) break;
            while (i >= sg.off + sg.n) sg = segs[++s];
            u64 k = sg.keys[i - sg.off];
            u32 src = (u32)(k >> 32), dst = (u32)k;
            if (src >= sg.V || dst >= sg.V) { bad = true; src = 0; dst = 0; }
            u64 p = ((u64)sg.pid << pshift) | ((((((u64)src) << sg.vbits) | dst) << sbits | sg.seq) << 1) | sg.op;
            out[i] = p;
            for (u32 q = 0; q < npass; ++q) atomicAdd(&h[q][(u32)(p >> (lo_bit + q * DB)) & (R - 1)], 1u);
        }
    }
    if (__any(bad) && lane_id() == 0) atomicOr(flags, FLAG_INVALID_KEY);
    __syncthreads();
    for (u32 i = threadIdx.x; i < npass * R; i += OS_TPB) {
        u32 v = (&h[0][0])[i];
        if (v) atomicAdd(&gh[i], v);
    }
}

// One pass over bits [shift, shift + DB).  Tile t = keys [t*T, t*T + T),
// T = OS_TPB * ITEMS; wave w holds the 64*ITEMS consecutive keys from
// t*T + w*64*ITEMS, one per lane per round, so (wave, round, lane) order is
// input order and the ranking is stable.  W predecessors are read per
// look-back step.
template <int DB, int ITEMS, int W>
__global__ __launch_bounds__(OS_TPB) void k_os_pass(const u64 *__restrict__ in, u64 *__restrict__ out, u64 n,
                                                     const u64 *n_dev, u32 shift, const u32 *__restrict__ gh,
                                                     u32 *__restrict__ look, u32 *__restrict__ look_next,
                                                     u32 *__restrict__ flags) {
    constexpr u32 R = 1u << DB;
    constexpr u32 NW = OS_TPB / 64;
    constexpr u32 DPT = R / OS_TPB;  // digits per thread
    extern __shared__ u32 os_lds[];
    u32(*whist)[R] = (u32(*)[R])os_lds;  // per-wave digit counts, then per-wave offsets
    u32 *gofs = os_lds + NW * R;          // keys of smaller digits in the whole batch
    u32 *tofs = gofs + R;                 // keys of this digit in earlier tiles
    __shared__ u32 stmp[NW + 1];
    if (n_dev) n = *n_dev;
    const u32 t = blockIdx.x;
    if (look_next)
        for (u32 i = threadIdx.x; i < R; i += OS_TPB) look_next[(u64)t * R + i] = 0;
    const u64 tbase = (u64)t * OS_TPB * ITEMS;
    if (tbase >= n) return;  // whole workgroup: grids sized for a capacity
    const u32 lane = lane_id();
    const u32 w = threadIdx.x >> 6;
    const u64 lt = (1ull << lane) - 1ull;

    // batch-wide digit totals of this pass, contiguous per thread
    u32 gv[DPT];
#pragma unroll
    for (u32 j = 0; j < DPT; ++j) gv[j] = gh[threadIdx.x * DPT + j];
    u64 key[ITEMS];
    const u64 wbase = tbase + (u64)w * 64 * ITEMS + lane;
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        u64 i = wbase + (u64)r * 64;
        key[r] = i < n ? in[i] : 0ull;
    }
    for (u32 i = threadIdx.x; i < NW * R; i += OS_TPB) os_lds[i] = 0;
    __syncthreads();

    // rank inside the wave: lanes of equal digit found with DB ballots; the
    // wave's running count of each digit lives in whist[w] (one wave reads
    // and then updates its own row, in order)
    u32 rk[ITEMS];
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        const bool valid = wbase + (u64)r * 64 < n;
        const u32 dg = (u32)(key[r] >> shift) & (R - 1);
        u64 peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < DB; ++b) {
            u64 bb = __ballot((dg >> b) & 1u);
            peers &= ((dg >> b) & 1u) ? bb : ~bb;
        }
        const u32 below = __popcll(peers & lt);
        const u32 before = whist[w][dg];
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) whist[w][dg] = before + __popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rk[r] = before + below;
    }
    __syncthreads();

    // per digit: offsets of the waves, the tile's count, published at once
    u32 cnt[DPT];
#pragma unroll
    for (u32 j = 0; j < DPT; ++j) {
        const u32 d = j * OS_TPB + threadIdx.x;
        u32 run = 0;
#pragma unroll
        for (u32 v = 0; v < NW; ++v) {
            u32 c = whist[v][d];
            whist[v][d] = run;
            run += c;
        }
        cnt[j] = run;
        os_store(&look[(u64)t * R + d], (t == 0 ? OS_PRE : OS_AGG) | run);
    }
    // exclusive scan of the digit totals
    {
        u32 sum = 0;
#pragma unroll
        for (u32 j = 0; j < DPT; ++j) sum += gv[j];
        u32 tot;
        u32 ex = block_excl_scan<OS_TPB, u32>(sum, tot, stmp);
#pragma unroll
        for (u32 j = 0; j < DPT; ++j) { gofs[threadIdx.x * DPT + j] = ex; ex += gv[j]; }
    }

    // look-back: keys of each digit in tiles < t
    u32 excl[DPT];
    int p[DPT];
#pragma unroll
    for (u32 j = 0; j < DPT; ++j) { excl[j] = 0; p[j] = (int)t - 1; }
    bool stalled = false;
    for (u32 spin = 0;; ++spin) {
        u32 s[DPT][W];
#pragma unroll
        for (u32 j = 0; j < DPT; ++j) {
            const u32 d = j * OS_TPB + threadIdx.x;
            // unconditional loads (row 0 stands in past the first tile) so
            // that all W * DPT of them are in flight together
#pragma unroll
            for (int k = 0; k < W; ++k) {
                const int q = p[j] - k;
                s[j][k] = os_load(&look[(u64)(q > 0 ? q : 0) * R + d]);
            }
        }
        // consume predecessors in order until one is unpublished (read it
        // again next round) or carries the prefix of all before it (done);
        // branch-free so that s stays in registers
        bool pending = false;
#pragma unroll
        for (u32 j = 0; j < DPT; ++j) {
            int q = p[j];
            bool go = q >= 0;
            u32 acc = 0;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                const u32 v = s[j][k];
                const u32 f = v & ~OS_CNT;
                const bool use = go && f != 0;
                acc += use ? (v & OS_CNT) : 0u;
                q = use ? (f == OS_PRE ? -1 : q - 1) : q;
                go = use && f != OS_PRE;
            }
            excl[j] += acc;
            p[j] = q;
            pending |= q >= 0;
        }
        if (!pending) break;
        if (spin >= OS_SPIN_LIMIT) {
            stalled = true;
            break;
        }
    }
    if (stalled) atomicOr(flags, FLAG_SORT_STALL);
#pragma unroll
    for (u32 j = 0; j < DPT; ++j) {
        const u32 d = j * OS_TPB + threadIdx.x;
        tofs[d] = excl[j];
        if (t != 0) os_store(&look[(u64)t * R + d], OS_PRE | ((excl[j] + cnt[j]) & OS_CNT));
    }
    __syncthreads();

#pragma unroll
    for (int r = 0; r < ITEMS; ++r) {
        if (wbase + (u64)r * 64 >= n) break;
        const u32 dg = (u32)(key[r] >> shift) & (R - 1);
        u64 pos = (u64)gofs[dg] + tofs[dg] + whist[w][dg] + rk[r];
        if (pos >= n) pos = n - 1;  // only after a stall (flagged): stay in bounds
        out[pos] = key[r];
    }
}

// ------------------------------------------------------- sample-split sort
//
// Batches of up to SS_MAX_N keys sort in two kernels, whatever their key
// width.  k_ss_split: every workgroup draws the same regular sample of the
// batch (SS_SAMPLE packed keys), sorts it in LDS by counting, and takes its
// quantiles as the splitters of nb buckets (rounded down to a key boundary,
// so that equal keys of any request order and operation share a bucket);
// then it packs its tile of SS_TILE keys, assigns each to its bucket, and
// writes the tile back grouped by bucket (tile-major: no global offsets are
// needed), with its per-bucket counts.  k_ss_bucket: one workgroup per
// bucket gathers the bucket's keys from every tile into registers, finds
// the bits on which they differ (AND / OR over the bucket), sorts them by
// LSD radix passes of 8 bits over those bits only, ranked in LDS (wave64
// "match any" ballots, per-wave digit counts joined by a DPP scan), and
// writes them at the bucket's offset.  A sample-chosen split keeps buckets
// even under R-MAT skew, where a split on the top key bits puts a hub's keys
// in one bucket.  A bucket beyond the workgroup's SB_CAP keys (many equal
// keys) is LSD-radix-sorted through global memory, chunk by chunk, on the
// bits on which its keys differ (AND / OR of the keys), with wave64 ranking
// (the k_os_pass scheme).

static constexpr int SS_TPB = 256;
static constexpr int SS_ITEMS = 4;
static constexpr u32 SS_TILE = SS_TPB * SS_ITEMS;  // keys per tile
static constexpr u32 SS_SAMPLE = SS_TPB;            // sample keys (one per thread)
static constexpr u32 SS_MAX_BUCKETS = 128;
static constexpr u64 SS_MAX_N = 1ull << 17;
static constexpr u32 SS_MAX_TILES = (u32)(SS_MAX_N / SS_TILE);
static constexpr int SB_TPB = 1024;                 // bucket sort workgroup
static constexpr int SB_ITEMS = 4;
static constexpr u32 SB_CAP = SB_TPB * SB_ITEMS;    // keys sorted in LDS
static constexpr u32 SB_NW = SB_TPB / 64;
// a tile's count entry: offset << 16 | count, bit 31 on bucket 0's entry
// when the tile met an invalid key
static constexpr u32 SS_BAD = 1u << 31;
static_assert(SS_TILE <= (1u << 15), "a tile's offsets and counts stay below the bad-key bit");

// buckets of a batch of n keys: about a thousand keys each, a power of two
static u32 ss_buckets(u64 n) {
    u32 nb = 1;
    while (nb < SS_MAX_BUCKETS && (u64)nb * SS_TILE < n) nb <<= 1;
    return nb;
}

__device__ __forceinline__ u64 pack_element(const BatchSeg &sg, u64 i, u32 pshift, u32 sbits, bool &bad) {
    u64 k = sg.keys[i - sg.off];
    u32 src = (u32)(k >> 32), dst = (u32)k;
    if (src >= sg.V || dst >= sg.V) { bad = true; src = 0; dst = 0; }
    return ((u64)sg.pid << pshift) | ((((((u64)src) << sg.vbits) | dst) << sbits | sg.seq) << 1) | sg.op;
}

// Splitters and bucket choice work on 32-bit projections of the packed
// keys, x >> psh with psh >= the bits below the key (request order,
// operation): equal keys project equally, so they share a bucket, and any
// non-decreasing splitters give a correct split.
// FIRST: a lone PMA's batch of inline sources, whose first work (status,
// update index) one more workgroup, past the tiles, does here instead of a
// kernel of its own, and the next pyr_chunks workgroups the count pyramid's
// chunk levels the last batch left owing (its top levels follow in the
// bucket kernel); the sources come from the kernel arguments.  A tile that
// met an invalid key marks its count of bucket 0 (bit 31): the bucket
// kernel raises the flag, after the status is cleared.
template <bool FIRST>
__global__ __launch_bounds__(SS_TPB) void k_ss_split(const BatchSeg *__restrict__ segs, u32 nsegs, u64 n,
                                                      const u64 *n_dev, u32 pshift, u32 sbits, u32 psh, u32 nb,
                                                      u64 *__restrict__ tmp, u32 *__restrict__ bc, InlineSegs in,
                                                      Group g, Status *status, u32 *wfirst) {
    __shared__ u64 smp[SS_SAMPLE];  // projection << 32 | sample index: distinct
    __shared__ u32 spl[SS_MAX_BUCKETS];
    __shared__ u32 cnt[SS_MAX_BUCKETS];
    __shared__ u32 base[SS_MAX_BUCKETS];
    __shared__ BatchSeg lsegs[INLINE_SEGS];
    if (n_dev) n = *n_dev;
    const u32 t = blockIdx.x;
    const u32 ntiles = (u32)((n + SS_TILE - 1) / SS_TILE);
    const u32 tid = threadIdx.x;
    if (FIRST) {
        if (t >= ntiles) {
            if (t == ntiles)
                batch_first_work(g, status, 0, nullptr, 0, wfirst, false);
            else
                pyr_chunk(pma_of(g, 0), t - ntiles - 1);
            return;
        }
        if (tid < in.n) lsegs[tid] = in.s[tid];
        __syncthreads();
        segs = lsegs;
    }
    if (t >= ntiles) return;  // whole workgroup: grids sized for a capacity

    // the tile's keys and the sample key are loaded first, all in flight
    // while the sample is ranked
    bool bad = false;
    const u64 tbase = (u64)t * SS_TILE;
    u64 key[SS_ITEMS];
    bool valid[SS_ITEMS];
    {
        const u64 i0 = tbase + tid;
        u32 s = source_of(segs, nsegs, i0 < n ? i0 : n - 1);
        BatchSeg sg = segs[s];
#pragma unroll
        for (int r = 0; r < SS_ITEMS; ++r) {
            const u64 i = i0 + (u64)r * SS_TPB;
            valid[r] = i < n;
            key[r] = 0;
            if (valid[r]) {
                while (i >= sg.off + sg.n) sg = segs[++s];
                key[r] = pack_element(sg, i, pshift, sbits, bad);
            }
        }
    }
    {
        const u64 p = ((2ull * tid + 1) * n) / (2ull * SS_SAMPLE);
        bool ignore = false;
        const u64 x = pack_element(segs[source_of(segs, nsegs, p)], p, pshift, sbits, ignore);
        smp[tid] = ((u64)(u32)(x >> psh) << 32) | tid;
    }
    if (tid < SS_MAX_BUCKETS) cnt[tid] = 0;
    __syncthreads();
    // rank of each (distinct) sample entry by counting, two entries per LDS read
    {
        const u64 x = smp[tid];
        const ulonglong2 *s2 = (const ulonglong2 *)smp;
        u32 r = 0;
#pragma unroll 16
        for (u32 q = 0; q < SS_SAMPLE / 2; ++q) {
            const ulonglong2 y = s2[q];
            r += (y.x < x ? 1u : 0u) + (y.y < x ? 1u : 0u);
        }
        __syncthreads();
        smp[r] = x;
    }
    __syncthreads();
    if (tid + 1 < nb) spl[tid] = (u32)(smp[(tid + 1) * (SS_SAMPLE / nb)] >> 32);
    __syncthreads();

    // bucket = number of splitters <= the key's projection
    u32 bk[SS_ITEMS], loc[SS_ITEMS];
#pragma unroll
    for (int r = 0; r < SS_ITEMS; ++r) {
        const u32 x = (u32)(key[r] >> psh);
        u32 lo = 0, hi = nb - 1;
        while (lo < hi) {
            const u32 mid = (lo + hi) >> 1;
            if (spl[mid] <= x) lo = mid + 1; else hi = mid;
        }
        bk[r] = lo;
        loc[r] = valid[r] ? atomicAdd(&cnt[lo], 1u) : 0u;
    }
    const bool tbad = __syncthreads_or(bad);
    if (tid < 64) {
        // exclusive scan of the bucket counts (nb <= 128: two per lane)
        const u32 c0 = 2 * tid < nb ? cnt[2 * tid] : 0u, c1 = 2 * tid + 1 < nb ? cnt[2 * tid + 1] : 0u;
        const u32 ex = wave_incl_scan(c0 + c1) - (c0 + c1);
        if (2 * tid < nb) base[2 * tid] = ex;
        if (2 * tid + 1 < nb) base[2 * tid + 1] = ex + c0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SS_ITEMS; ++r)
        if (valid[r]) tmp[tbase + base[bk[r]] + loc[r]] = key[r];
    if (tid < nb) bc[(u64)tid * ntiles + t] = (base[tid] << 16) | cnt[tid] | (tid == 0 && tbad ? SS_BAD : 0u);
}

// Keys of a workgroup are spread over every wave: with R = ceil(m / SB_TPB)
// rounds, wave w holds keys w*64R + r*64 + lane (r < R), so a bucket of up
// to SB_TPB keys is ranked in one round by all waves at once.
__device__ __forceinline__ u32 sb_rounds(u32 m) { return (m + SB_TPB - 1) / SB_TPB; }
__device__ __forceinline__ u32 sb_index(u32 w, u32 r, u32 lane, u32 R) { return w * (64 * R) + r * 64 + lane; }

// One LSD pass over the (up to SB_CAP) keys x[] of the workgroup, digit
// (x >> shift) & 255: returns each key's rank inside the chunk (stable, in
// the order of sb_index) in rk[] and leaves in dofs[d] the chunk's exclusive
// count of smaller digits, in whist[w][d] the keys of digit d in waves
// before w.  Every thread of the workgroup calls it.
__device__ __forceinline__ void sb_rank(const u64 (&x)[SB_ITEMS], u32 m, u32 R, u32 shift, u32 (*whist)[256],
                                        u32 *dofs, u32 *stmp, u32 (&rk)[SB_ITEMS]) {
    const u32 lane = lane_id();
    const u32 w = threadIdx.x >> 6;
    const u64 lt = (1ull << lane) - 1ull;
    for (u32 i = threadIdx.x; i < SB_NW * 256; i += SB_TPB) (&whist[0][0])[i] = 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SB_ITEMS; ++r) {
        if (r >= (int)R || sb_index(w, r, 0, R) >= m) break;  // the rest of this wave's rounds are empty
        const u32 j = sb_index(w, r, lane, R);
        const bool valid = j < m;
        const u32 dg = (u32)(x[r] >> shift) & 255u;
        u64 peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const u64 bb = __ballot((dg >> b) & 1u);
            peers &= ((dg >> b) & 1u) ? bb : ~bb;
        }
        const u32 below = __popcll(peers & lt);
        const u32 before = whist[w][dg];
        __builtin_amdgcn_wave_barrier();
        if (valid && below == 0) whist[w][dg] = before + __popcll(peers);
        __builtin_amdgcn_wave_barrier();
        rk[r] = before + below;
    }
    __syncthreads();
    // the first four waves: digit d = thread; its count per wave turns into
    // the prefix of the waves before, its total into a wave-wide DPP scan
    u32 total = 0, inc = 0;
    if (threadIdx.x < 256) {
        const u32 d = threadIdx.x;
        u32 c[SB_NW];
#pragma unroll
        for (u32 v = 0; v < SB_NW; ++v) c[v] = whist[v][d];
#pragma unroll
        for (u32 v = 0; v < SB_NW; ++v) {
            whist[v][d] = total;
            total += c[v];
        }
        inc = dpp_incl_scan32(total);
        if (lane == 63) stmp[w] = inc;
    }
    __syncthreads();
    if (threadIdx.x < 256) {
        u32 pre = 0;
        for (u32 v = 0; v < w; ++v) pre += stmp[v];
        dofs[threadIdx.x] = pre + inc - total;
    }
    __syncthreads();
}

// top: one more workgroup, past the buckets, sums the lone PMA's pyramid
// top levels (after the split kernel's chunk levels)
__global__ __launch_bounds__(SB_TPB) void k_ss_bucket(const u64 *__restrict__ tmp, u64 *__restrict__ out,
                                                       u64 *__restrict__ aux, u64 n, const u64 *n_dev, u32 nb,
                                                       const u32 *__restrict__ bc, u32 *__restrict__ flags,
                                                       bool top, DevState topd) {
    __shared__ u64 kb[1][SB_CAP];
    __shared__ u32 whist[SB_NW][256];
    __shared__ u32 dofs[256];
    __shared__ u32 run[256];
    __shared__ u32 toff[SS_MAX_TILES + 1], tb[SS_MAX_TILES];
    __shared__ u32 stmp[SB_NW + 1];
    __shared__ u64 red[2][SB_NW];
    (void)kb;
    if (n_dev) n = *n_dev;
    const u32 b = blockIdx.x;
    const u32 tid = threadIdx.x;
    const u32 ntiles = (u32)((n + SS_TILE - 1) / SS_TILE);
    if (top && b == nb) {
        pyr_top_levels(topd);
        return;
    }
    if (b >= nb || ntiles == 0) return;
    // the bucket's offset: the keys of the buckets before it, whose
    // per-tile counts are the first b rows of bc
    u64 obase;
    {
        // at most SS_MAX_BUCKETS * SS_MAX_TILES / SB_TPB = 16 words a thread,
        // all loads in flight at once
        constexpr u32 PER = SS_MAX_BUCKETS * SS_MAX_TILES / SB_TPB;
        const u32 lim = b * ntiles;
        u32 part[PER];
#pragma unroll
        for (u32 q = 0; q < PER; ++q) {
            const u32 i = q * SB_TPB + tid;
            part[q] = i < lim ? bc[i] & 0xFFFFu : 0u;
        }
        u32 sum = 0;
#pragma unroll
        for (u32 q = 0; q < PER; ++q) sum += part[q];
        u32 tot;
        (void)block_excl_scan<SB_TPB, u32>(sum, tot, stmp);
        obase = tot;
    }
    // where the bucket lies in each tile: offsets inside the bucket
    u32 m;
    {
        const u32 e = tid < ntiles ? bc[(u64)b * ntiles + tid] : 0u;
        if (b == 0 && __syncthreads_or((e & SS_BAD) != 0u) && tid == 0) atomicOr(flags, FLAG_INVALID_KEY);
        u32 tot;
        const u32 ex = block_excl_scan<SB_TPB, u32>(e & 0xFFFFu, tot, stmp);
        if (tid < ntiles) {
            toff[tid] = ex;
            tb[tid] = (e & ~SS_BAD) >> 16;
        }
        if (tid == 0) toff[ntiles] = tot;
        m = tot;
    }
    if (m == 0) return;
    __syncthreads();
    const u32 lane = lane_id();
    const u32 w = tid >> 6;
    auto addr_of = [&](u32 j) -> u64 {
        u32 lo = 0, hi = ntiles;  // last tile with toff <= j
        while (hi - lo > 1) {
            const u32 mid = (lo + hi) >> 1;
            if (toff[mid] <= j) lo = mid; else hi = mid;
        }
        return (u64)lo * SS_TILE + tb[lo] + (j - toff[lo]);
    };

    if (m <= SB_CAP) {
        // gather: element j = sb_index(w, r, lane, R) into x[r] (addresses
        // first, then the loads), and the bits on which the keys differ
        const u32 R = sb_rounds(m);
        u64 x[SB_ITEMS];
        u32 rk[SB_ITEMS];
        {
            u64 ad[SB_ITEMS];
#pragma unroll
            for (int r = 0; r < SB_ITEMS; ++r) {
                const u32 j = r < (int)R ? sb_index(w, r, lane, R) : m;
                ad[r] = j < m ? addr_of(j) : ~0ull;
            }
#pragma unroll
            for (int r = 0; r < SB_ITEMS; ++r) x[r] = ad[r] != ~0ull ? tmp[ad[r]] : 0ull;
        }
        u64 andv = ~0ull, orv = 0;
#pragma unroll
        for (int r = 0; r < SB_ITEMS; ++r) {
            const u32 j = r < (int)R ? sb_index(w, r, lane, R) : m;
            if (j < m) {
                andv &= x[r];
                orv |= x[r];
            }
        }
        andv = wave_and64(andv);
        orv = wave_or64(orv);
        if (lane == 0) { red[0][w] = andv; red[1][w] = orv; }
        __syncthreads();
        andv = ~0ull;
        orv = 0;
#pragma unroll
        for (u32 q = 0; q < SB_NW; ++q) { andv &= red[0][q]; orv |= red[1][q]; }
        const u64 diff = andv ^ orv;
        const u32 lo_bit = diff ? (u32)__builtin_ctzll(diff) : 64u;
        const u32 hi_bit = diff ? 64u - (u32)__builtin_clzll(diff) : 0u;
        u64 *kbuf = kb[0];
        for (u32 shift = lo_bit; shift < hi_bit; shift += 8) {
            sb_rank(x, m, R, shift, whist, dofs, stmp, rk);
#pragma unroll
            for (int r = 0; r < SB_ITEMS; ++r) {
                const u32 j = r < (int)R ? sb_index(w, r, lane, R) : m;
                if (j < m) {
                    const u32 dg = (u32)(x[r] >> shift) & 255u;
                    kbuf[dofs[dg] + whist[w][dg] + rk[r]] = x[r];
                }
            }
            __syncthreads();
#pragma unroll
            for (int r = 0; r < SB_ITEMS; ++r) {
                const u32 j = r < (int)R ? sb_index(w, r, lane, R) : m;
                if (j < m) x[r] = kbuf[j];
            }
            // (the next pass's first barrier keeps kbuf until every thread has read it)
        }
#pragma unroll
        for (int r = 0; r < SB_ITEMS; ++r) {
            const u32 j = r < (int)R ? sb_index(w, r, lane, R) : m;
            if (j < m) out[obase + j] = x[r];
        }
        return;
    }

    // beyond the LDS (many equal keys): gathered into the output range, then
    // LSD passes over the bits that vary, chunk by chunk through global
    // memory (from the output range to aux and back); each digit's running
    // offset carries across the chunks, which are taken in order (stable)
    u64 andv = ~0ull, orv = 0;
    for (u32 j = tid; j < m; j += SB_TPB) {
        const u64 x = tmp[addr_of(j)];
        andv &= x;
        orv |= x;
        out[obase + j] = x;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        andv &= __shfl_xor(andv, o, 64);
        orv |= __shfl_xor(orv, o, 64);
    }
    if (lane == 0) { red[0][w] = andv; red[1][w] = orv; }
    __syncthreads();  // also makes the gather's stores visible to the workgroup
    andv = ~0ull;
    orv = 0;
#pragma unroll
    for (u32 q = 0; q < SB_NW; ++q) { andv &= red[0][q]; orv |= red[1][q]; }
    const u64 diff = andv ^ orv;
    const u32 lo_bit = diff ? (u32)__builtin_ctzll(diff) : 64u;
    const u32 hi_bit = diff ? 64u - (u32)__builtin_clzll(diff) : 0u;
    u64 *src = out + obase, *dst = aux + obase;
    for (u32 shift = lo_bit; shift < hi_bit; shift += 8) {
        // digit totals of the bucket, then their exclusive scan as the start
        for (u32 i = tid; i < SB_NW * 256; i += SB_TPB) (&whist[0][0])[i] = 0;
        __syncthreads();
        for (u32 j = tid; j < m; j += SB_TPB) atomicAdd(&whist[0][(u32)(src[j] >> shift) & 255u], 1u);
        __syncthreads();
        {
            const u32 c = tid < 256 ? whist[0][tid] : 0u;
            u32 tot;
            const u32 ex = block_excl_scan<SB_TPB, u32>(c, tot, stmp);
            if (tid < 256) run[tid] = ex;
        }
        __syncthreads();
        for (u32 c0 = 0; c0 < m; c0 += SB_CAP) {
            const u32 mc = m - c0 < SB_CAP ? m - c0 : SB_CAP;
            const u32 R = sb_rounds(mc);
            u64 x[SB_ITEMS];
            u32 rk[SB_ITEMS];
#pragma unroll
            for (int r = 0; r < SB_ITEMS; ++r) {
                const u32 j = r < (int)R ? sb_index(w, r, lane, R) : mc;
                x[r] = j < mc ? src[c0 + j] : 0ull;
            }
            sb_rank(x, mc, R, shift, whist, dofs, stmp, rk);
#pragma unroll
            for (int r = 0; r < SB_ITEMS; ++r) {
                const u32 j = r < (int)R ? sb_index(w, r, lane, R) : mc;
                if (j < mc) {
                    const u32 dg = (u32)(x[r] >> shift) & 255u;
                    dst[run[dg] + whist[w][dg] + rk[r]] = x[r];
                }
            }
            __syncthreads();
            // the chunk's count of digit d: the next digit's chunk offset
            // minus this one's
            if (tid < 256) run[tid] += (tid + 1 < 256 ? dofs[tid + 1] : mc) - dofs[tid];
            __syncthreads();
        }
        u64 *t2 = src; src = dst; dst = t2;
        __syncthreads();
    }
    if (src != out + obase)
        for (u32 j = tid; j < m; j += SB_TPB) out[obase + j] = src[j];
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

static u32 os_items(u64) { return OS_ITEMS; }

// digit width for a sort over `width` bits: 10 unless 11 saves a pass
static u32 os_digit_bits(u32 width) { return (width + 9) / 10 == (width + 10) / 11 ? 10 : 11; }

u64 sort_scratch_words(u64 n) {
    u64 nb = (n + (u64)rs_items(n) * RS_TPB - 1) / ((u64)rs_items(n) * RS_TPB);
    u64 two = (u64)RS_RADIX * nb + 1 + scan_scratch_words((u64)RS_RADIX * nb);
    u64 tile = (u64)os_items(n) * OS_TPB;
    u64 one = 2 * ((n + tile - 1) / tile) * 2048;
    return OS_GH_WORDS + (n <= OS_MAX_N ? std::max(two, one) : two);
}

u32 sort_zero_words() { return OS_GH_WORDS; }

template <int DB>
static void os_sort(const BatchSeg *segs, u32 nsegs, u64 n, const u64 *n_dev, u32 pshift, u32 sbits, u32 lo_bit,
                    u32 npass, u64 *k0, u64 *k1, u32 *scratch, u32 *flags, u64 **result, hipStream_t st) {
    constexpr u32 R = 1u << DB;
    const u32 items = os_items(n);
    const u64 tile = (u64)items * OS_TPB;
    const u32 ntiles = (u32)((n + tile - 1) / tile);
    u32 *gh = scratch;
    u32 *look[2] = {scratch + OS_GH_WORDS, scratch + OS_GH_WORDS + (u64)ntiles * R};
    k_pack_count<DB><<<ntiles, OS_TPB, 0, st>>>(segs, nsegs, n, n_dev, pshift, sbits, k0, flags, lo_bit, npass, items,
                                                gh, look[0]);
    u64 *src = k0, *dst = k1;
    for (u32 q = 0; q < npass; ++q) {
        const u32 shift = lo_bit + q * DB;
        u32 *nxt = q + 1 < npass ? look[(q + 1) & 1] : nullptr;
        const size_t lds = ((size_t)(OS_TPB / 64) * R + 2 * R) * sizeof(u32);
        k_os_pass<DB, OS_ITEMS, OS_LOOK / (R / OS_TPB)><<<ntiles, OS_TPB, lds, st>>>(src, dst, n, n_dev, shift,
                                                                                   gh + q * R, look[q & 1], nxt, flags);
        u64 *t = src; src = dst; dst = t;
    }
    *result = src;
}

// Packs a group batch (n elements from nsegs sources) and sorts the packed
// keys on bits [lo_bit, hi_bit): the first histogram is taken while packing.
// With n_dev the kernels read the element count from device memory and n
// is only a capacity that sizes the grids (graph replays).  The first
// sort_zero_words() words of `scratch` must be zero on entry (the batch's
// first kernel clears them).
u64 *pack_and_sort(const BatchSeg *segs, u32 nsegs, u64 n, const u64 *n_dev, u32 pshift, u32 sbits, u32 lo_bit,
                   u32 hi_bit, u64 *k0, u64 *k1, u64 *k2, u32 *scratch, u32 *flags, hipStream_t st) {
    if (n == 0) return k0;
    if (n <= SS_MAX_N) {
        // sample-split sort: k1 holds the tiles grouped by bucket, k0 the
        // result, k2 the second buffer of oversized buckets; the tiles'
        // bucket counts go to the scratch after its cleared words
        const u32 nb = ss_buckets(n);
        const u32 ntiles = (u32)((n + SS_TILE - 1) / SS_TILE);
        u32 *bc = scratch + OS_GH_WORDS;
        const u32 ks = 1 + sbits;
        const u32 psh = hi_bit > 32 + ks ? hi_bit - 32 : ks;
        k_ss_split<false><<<ntiles, SS_TPB, 0, st>>>(segs, nsegs, n, n_dev, pshift, sbits, psh, nb, k1, bc,
                                                      InlineSegs{}, Group{}, nullptr, nullptr);
        k_ss_bucket<<<nb, SB_TPB, 0, st>>>(k1, k0, k2, n, n_dev, nb, bc, flags, false, DevState{});
        return k0;
    }
    if (n <= OS_MAX_N) {
        const u32 width = hi_bit - lo_bit;
        const u32 db = os_digit_bits(width);
        const u32 npass = (width + db - 1) / db;
        u64 *res = k0;
        if (db == 10)
            os_sort<10>(segs, nsegs, n, n_dev, pshift, sbits, lo_bit, npass, k0, k1, scratch, flags, &res, st);
        else
            os_sort<11>(segs, nsegs, n, n_dev, pshift, sbits, lo_bit, npass, k0, k1, scratch, flags, &res, st);
        return res;
    }
    scratch += OS_GH_WORDS;
    const u32 items = rs_items(n);
    const u64 tile = (u64)items * RS_TPB;
    const u32 nblk = (u32)((n + tile - 1) / tile);
    u32 *hist = scratch;
    u32 *sscr = scratch + (u64)RS_RADIX * nblk + 1;
    u64 *src = k0, *dst = k1;
    for (u32 shift = lo_bit; shift < hi_bit; shift += RS_BITS) {
        if (shift == lo_bit)
            k_pack_group<<<nblk, RS_TPB, 0, st>>>(segs, nsegs, n, n_dev, pshift, sbits, k0, flags, shift, items,
                                                  hist, nblk);
        else
            k_rs_hist<<<nblk, RS_TPB, 0, st>>>(src, n, n_dev, shift, items, hist, nblk);
        (void)device_excl_scan(hist, (u64)RS_RADIX * nblk, hist, sscr, st);
        k_rs_scatter<<<nblk, RS_TPB, 0, st>>>(src, dst, n, n_dev, shift, items, hist, nblk);
        u64 *t = src; src = dst; dst = t;
    }
    return src;
}

bool sort_carries_first(u64 n) { return n > 0 && n <= SS_MAX_N; }

// The sample-split sort of a lone PMA's batch of at most SS_MAX_N keys from
// inline sources, with the batch's first work (batch_begin without its
// sort-scratch clearing, which the sample-split sort does not need) in one
// more workgroup of the split kernel, pyr_chunks more for the pyramid's
// chunk levels the last batch left owing, and one more of the bucket kernel
// for the pyramid's top levels.
u64 *pack_and_sort_first(const Group &g, Status *status, u32 *wfirst, const InlineSegs &in, u64 n, u32 pshift,
                         u32 sbits, u32 hi_bit, u64 *k0, u64 *k1, u64 *k2, u32 *scratch, u32 pyr_chunks,
                         hipStream_t st) {
    const u32 nb = ss_buckets(n);
    const u32 ntiles = (u32)((n + SS_TILE - 1) / SS_TILE);
    u32 *bc = scratch + OS_GH_WORDS;
    const u32 ks = 1 + sbits;
    const u32 psh = hi_bit > 32 + ks ? hi_bit - 32 : ks;
    k_ss_split<true><<<ntiles + 1 + pyr_chunks, SS_TPB, 0, st>>>(nullptr, in.n, n, nullptr, pshift, sbits, psh, nb,
                                                                  k1, bc, in, g, status, wfirst);
    k_ss_bucket<<<nb + 1, SB_TPB, 0, st>>>(k1, k0, k2, n, nullptr, nb, bc, &status->flags, true, g.one);
    return k0;
}

}  // namespace pma
