// common.hpp -- shared types, device state and wave64 helpers for the
// MI355X (gfx950) packed-memory-array edge store.
//
// Layout in HBM (one flat logical array, see the reference's logical slot
// space GET_LINEAR_IDX, rpma.cuh:132-134):
//   keys[S << H]    u64 slots, KEY_NONE gaps; segment s = slots [s*S, s*S+S),
//                   its live keys front-packed and sorted.
//   cnt[nseg]       u32 live keys per segment (level 0 of the count pyramid).
//   pyr[nseg-1]     u32 live keys per window for levels 1..H
//                   (level l holds nseg>>l windows at offset nseg-(nseg>>(l-1))).
//   sep[nseg]       u64 separator per segment: every key of segment s lies in
//                   [sep[s], sep[s+1]); sep[0] = 0.  A key routes to the last
//                   segment whose separator is <= key.
//   owner[2nseg]    u64 window claims of the current batch, tagged with the
//                   batch epoch (epoch << 32 | ~run), so no reset is needed.
//
// A pipeline run updates a group of such arrays at once (one or many
// independent PMAs): the kernels read each PMA's geometry from a device
// table indexed by the PMA id carried in every batch element.
//   row_off[V+1]    u32, row_off[r+1] = slot of row r's marker key (r<<32|~0u)
//                   (GPMA convention, gpma.cuh:340-344), row_off[0] = 0.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace pma {

typedef uint64_t u64;
typedef uint32_t u32;
typedef int32_t i32;
typedef int64_t i64;
// a u64 in global memory: accesses through it compile to global (not flat)
// instructions, which only the vector memory counter tracks
typedef __attribute__((address_space(1))) u64 gu64;

static constexpr u64 KEY_NONE = ~0ull;   // empty slot (rpma.cuh:38)
// run key of a batch element: pma << 32 | segment, bit 63 the operation
static constexpr u64 RK_NONE = ~0ull;
static constexpr u64 RK_OP = 1ull << 63;

// status flags raised on the device during one update
static constexpr u32 FLAG_INVALID_KEY = 1u;   // src or dst >= V
static constexpr u32 FLAG_ROOT = 2u;          // a run chose the root window
static constexpr u32 FLAG_RESIZE = 4u;        // the root window does not fit
static constexpr u32 FLAG_SORT_STALL = 8u;    // a sort tile waited too long (the batch fails)

// most passes of the batch sort (11-bit digits over 64 bits)
static constexpr u32 SORT_MAX_PASSES = 6;

// rebalance work classes
static constexpr int CLASS_WAVE = 0;   // one wave64 per window
static constexpr int CLASS_BLOCK = 1;  // one 256-thread workgroup per window
static constexpr int CLASS_LARGE = 2;  // multi-workgroup global path
static constexpr int CLASS_SEG = 3;    // single-segment window, S/2 lanes (several per wave)
static constexpr int NUM_CLASSES = 4;

// LDS capacities of the two in-LDS classes: E = old live + updates, and the
// number of segments in the window
static constexpr u32 WAVE_E_CAP = 256;
static constexpr u32 WAVE_SEG_CAP = 64;
static constexpr u32 BLOCK_E_CAP = 2048;
static constexpr u32 BLOCK_SEG_CAP = 1024;
// the multi-workgroup path cuts windows into chunks of this many slots
static constexpr u32 LARGE_CHUNK_LOG2_SLOTS = 11;
// Update index of a lone PMA, rebuilt every batch: for each window of the
// index level (at most WFIRST_N windows), the first compacted update inside
// it.  It turns a window's update range above that level into two lookups.
static constexpr u32 WFIRST_LOG2 = 12;
static constexpr u32 WFIRST_N = 1u << WFIRST_LOG2;
__host__ __device__ __forceinline__ unsigned index_level(unsigned H) { return H > WFIRST_LOG2 + 8 ? H - WFIRST_LOG2 : 8; }

// device statistics slots (u64 counters)
enum StatSlot {
    ST_WINDOWS = 0,        // [0, 32): windows rebalanced per level
    ST_REBUILD_SLOTS = 32, // slots written by root rebalances and resizes
    ST_REBUILDS = 33,      // root rebalances and resizes
    ST_UPDATES = 64,       // effective updates applied by window rebalances
    ST_MARKERS = 65,       // marker (row offset) writes
    ST_ALG_BYTES_WAVE = 66,   // algorithmic bytes, per rebalance class
    ST_ALG_BYTES_BLOCK = 67,
    ST_ALG_BYTES_LARGE = 68,
    ST_ALG_BYTES_SEG = 69,
    ST_NUM = 80,
};
// statistics are spread over ST_ROWS copies (row = block index mod ST_ROWS)
// so that per-workgroup flushes do not serialise on one address
static constexpr u32 ST_ROWS = 64;

struct Status {
    u32 flags;
    u32 n_eff;        // effective updates after dedup / presence check
    i32 delta;        // inserts - deletes among them
    u32 nwork[NUM_CLASSES];
};

struct Work {
    u32 pid;  // PMA of the window
    u32 a;    // first segment of the window
    u32 L;    // level: the window holds 1 << L segments
    u32 u0;   // compacted update range [u0, u1)
    u32 u1;
    u32 C;    // live keys in the window before the batch
    u32 M;    // live keys after the batch
};

struct DevState {
    u64 *keys;
    u32 *cnt;
    u32 *pyr;
    u64 *sep;
    u64 *owner;
    u32 *row_off;
    u32 log2S;
    u32 H;
    u32 V;
    u32 vbits;
};

struct Density {
    u64 lower[32];
    u64 upper[32];
    u64 root_shrink_below;  // root also fits below lower[H] when m > this
};

// One PMA of a pipeline run, as the kernels see it.
struct PmaEntry {
    DevState d;
    Density dn;
    u32 pyr_blk0;  // first pyramid chunk of this PMA in the group's chunk list
    u32 pyr_nblk;  // its chunks
};

// Per-PMA outcome of a run, read back by the host with the Status.
struct PmaStat {
    u32 flags;     // FLAG_ROOT / FLAG_RESIZE raised by its runs
    u32 u0, u1;    // its effective updates: compacted range [u0, u1)
    i32 pd0, pd1;  // the update prefix at u0 and u1 (delta = pd1 - pd0)
};

// claimed-level masks: bit l set when some run of the PMA claimed a window
// of level l >= 1; LV_REPL copies per PMA (written by workgroup index) so
// the density check's atomics do not all meet on one word
static constexpr u32 LV_REPL = 32;

// One batch source: n keys of one operation for one PMA, placed at
// [off, off + n) of the packed batch.
struct BatchSeg {
    const u64 *keys;
    u64 n;
    u64 off;
    u32 pid;
    u32 op;
    u32 V;
    u32 vbits;
    u32 seq;   // request order inside a merged sequence (later wins)
    u32 pad_;
};

// Per-round values a graph replay reads from device memory (the host
// rewrites them before every launch).
struct RoundHdr {
    u64 n;        // batch elements
    u64 *merged;  // scratch of the multi-workgroup windows
    u32 epoch;
    u32 pad_;
};

// The PMAs one pipeline run updates.  Sort key of a batch element:
// pid << pshift | ((src << vbits | dst) << sbits | seq) << 1 | op,
// pshift = 2 * max vbits + sbits + 1; seq orders the requests of a merged
// sequence, so the last element of equal keys is the last request's.
struct Group {
    const PmaEntry *tab;
    PmaStat *pstat;
    u32 *lvmask;  // npma x LV_REPL claimed-level masks
    u64 *stats;   // ST_ROWS x ST_NUM counters of this pipeline
    u32 npma;
    u32 pshift;
    u32 sbits;    // request-order bits below the key (0: one request)
    u32 epoch;    // tags this run's window claims
    u32 log2S;    // segment size, the same for every PMA of the run
    u32 seg_log2G;  // lanes per window of the single-segment rebalance (log2)
    u32 inline_one; // one PMA whose geometry travels in `one` (no table load)
    u32 defer_rebuild;  // elect no windows when some PMA needs a root rebuild or resize
    const RoundHdr *rh;  // graph replays: n and epoch come from here
    DevState one;
};

__device__ __forceinline__ u32 group_epoch(const Group &g) { return g.rh ? g.rh->epoch : g.epoch; }

// geometry of PMA pid: from the kernel arguments for a lone PMA (no
// dependent load on the critical path), else from the device table
__device__ __forceinline__ DevState pma_of(const Group &g, u32 pid) {
    if (g.inline_one) return g.one;
    return g.tab[pid].d;
}

__host__ __device__ __forceinline__ u64 nseg_of(const DevState &d) { return (u64)1 << d.H; }

// offset of level l (1..H) inside pyr
__host__ __device__ __forceinline__ u64 pyr_off(u64 nseg, u32 l) { return nseg - (nseg >> (l - 1)); }

__host__ __device__ __forceinline__ u64 owner_idx(u64 nseg, u32 l, u64 w) {
    return l == 0 ? w : nseg + pyr_off(nseg, l) + w;
}

__device__ __forceinline__ u32 *count_ptr(const DevState &d, u32 l, u64 w) {
    return l == 0 ? d.cnt + w : d.pyr + pyr_off(nseg_of(d), l) + w;
}

__host__ __device__ __forceinline__ bool is_marker(u64 k) {
    return (u32)k == 0xFFFFFFFFu && (u32)(k >> 32) != 0xFFFFFFFFu;
}

__host__ __device__ __forceinline__ u64 marker_key(u32 r) { return ((u64)r << 32) | 0xFFFFFFFFull; }

// first rank placed in segment j of a window of nsegw = 1<<L segments that
// holds m keys: ceil(j*m / nsegw) (even spread, rpma.cuh:676-679)
__host__ __device__ __forceinline__ u64 spread_first(u64 j, u64 m, u32 L) {
    return (j * m + ((1ull << L) - 1)) >> L;
}

__device__ __forceinline__ bool window_fits(const Density &dn, u32 l, u32 H, u64 m) {
    if (l < H) return m >= dn.lower[l] && m <= dn.upper[l];
    if (m > dn.upper[H]) return false;
    return m >= dn.lower[H] || m > dn.root_shrink_below;
}

__device__ __forceinline__ bool lds_class_ok(u64 E, u32 L) {
    return E <= BLOCK_E_CAP && (1ull << L) <= BLOCK_SEG_CAP;
}

__device__ __forceinline__ int work_class(u64 E, u32 L, u32 log2S) {
    if (L == 0 && E <= (2ull << log2S)) return CLASS_SEG;
    if (E <= WAVE_E_CAP && (1ull << L) <= WAVE_SEG_CAP) return CLASS_WAVE;
    if (lds_class_ok(E, L)) return CLASS_BLOCK;
    return CLASS_LARGE;
}

// ---------------------------------------------------------------- searches

// first index in [lo, hi) with a[i] >= x
template <typename T>
__device__ __forceinline__ u64 lower_bound_dev(const T *a, u64 lo, u64 hi, T x) {
    while (lo < hi) {
        u64 mid = (lo + hi) >> 1;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// route a key: last segment s with sep[s] <= key (sep[0] = 0, nseg = 2^H).
// A 4-ary search: each step loads three pivots at once (sep is sorted, so
// the number of pivots <= key picks the quarter), which halves the chain of
// dependent loads; an odd H ends with one binary step.
__device__ __forceinline__ u32 locate_segment(const u64 *__restrict__ sep, u32 H, u64 key) {
    u32 base = 0;
    u32 q = 1u << H;
    for (; q >= 4; q >>= 2) {
        const u32 r = q >> 2;
        const u64 a = sep[base + r], b = sep[base + 2 * r], c = sep[base + 3 * r];
        base += (a <= key ? r : 0u) + (b <= key ? r : 0u) + (c <= key ? r : 0u);
    }
    if (q == 2 && sep[base + 1] <= key) base += 1;
    return base;
}

// is `key` among the live (front-packed) keys of segment s (a binary search:
// a 4-ary one here, three probes per step, measured slower)
__device__ __forceinline__ bool segment_holds(const DevState &d, u32 s, u64 key) {
    const u64 *seg = d.keys + ((u64)s << d.log2S);
    u32 c = d.cnt[s];
    u64 i = lower_bound_dev<u64>(seg, 0, c, key);
    return i < c && seg[i] == key;
}

// --------------------------------------------------------- wave64 scans

__device__ __forceinline__ u32 lane_id() { return __lane_id(); }

// Inclusive prefix sum of 32-bit lanes by DPP, with no LDS round trip:
// row shifts by 1, 2, 4, 8 scan each row of 16 lanes (a lane whose source
// lies outside its row adds 0), then lane 15's total is broadcast into rows 1
// and 3 and lane 31's into rows 2 and 3 (row_bcast:15 / row_bcast:31).
// Every lane of the wave must be active.
__device__ __forceinline__ u32 dpp_incl_scan32(u32 v) {
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, false);  // row_shr:8
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (u32)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// AND / OR of a 64-bit value over the wave, the same in every lane (the
// DPP scan pattern with the operation's identity for lanes out of a row;
// lane 63 ends with the whole wave's result)
template <bool AND>
__device__ __forceinline__ u32 dpp_reduce32(u32 v) {
    const int id = AND ? -1 : 0;
#define LPMA_DPP_STEP(ctrl, rows)                                                           \
    {                                                                                       \
        const u32 t = (u32)__builtin_amdgcn_update_dpp(id, (int)v, ctrl, rows, 0xF, false); \
        v = AND ? (v & t) : (v | t);                                                        \
    }
    LPMA_DPP_STEP(0x111, 0xF)
    LPMA_DPP_STEP(0x112, 0xF)
    LPMA_DPP_STEP(0x114, 0xF)
    LPMA_DPP_STEP(0x118, 0xF)
    LPMA_DPP_STEP(0x142, 0xA)
    LPMA_DPP_STEP(0x143, 0xC)
#undef LPMA_DPP_STEP
    return (u32)__builtin_amdgcn_readlane((int)v, 63);
}

__device__ __forceinline__ u64 wave_and64(u64 v) {
    return ((u64)dpp_reduce32<true>((u32)(v >> 32)) << 32) | dpp_reduce32<true>((u32)v);
}
__device__ __forceinline__ u64 wave_or64(u64 v) {
    return ((u64)dpp_reduce32<false>((u32)(v >> 32)) << 32) | dpp_reduce32<false>((u32)v);
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    if constexpr (sizeof(T) == 4 && std::is_integral<T>::value) {
        return (T)dpp_incl_scan32((u32)v);
    } else {
        const int lane = (int)lane_id();
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            T t = __shfl_up(v, o, 64);
            if (lane >= o) v += t;
        }
        return v;
    }
}

// sum over the wave, the same in every lane (every lane of the wave active)
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
    if constexpr (sizeof(T) == 4 && std::is_integral<T>::value) {
        return (T)__builtin_amdgcn_readlane((int)dpp_incl_scan32((u32)v), 63);
    } else {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        return v;
    }
}

// exclusive scan across a block of NT threads (NT multiple of 64).
// `tmp` needs NT/64 + 1 entries of LDS.  Returns the exclusive prefix and
// writes the block total to `total`.
template <int NT, typename T>
__device__ __forceinline__ T block_excl_scan(T v, T &total, T *tmp) {
    constexpr int NW = NT / 64;
    const int tid = threadIdx.x;
    const int w = tid >> 6;
    T inc = wave_incl_scan(v);
    if (NW == 1) {
        total = __shfl(inc, 63, 64);
        return inc - v;
    }
    if ((tid & 63) == 63) tmp[w] = inc;
    __syncthreads();
    // every thread adds up the wave totals (the same LDS words for all lanes)
    T pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const T t = tmp[i];
        tot += t;
        pre += i < w ? t : (T)0;
    }
    total = tot;
    __syncthreads();  // tmp is free for the next call
    return pre + inc - v;
}

// in-place exclusive scan of an LDS array of n entries (n may exceed NT);
// arr[n] receives the total.  Every thread of the block must call it.
template <int NT>
__device__ __forceinline__ void block_scan_array(u32 *arr, u32 n, u32 *tmp) {
    u32 carry = 0;
    for (u32 base = 0; base < n; base += NT) {
        u32 i = base + threadIdx.x;
        u32 v = i < n ? arr[i] : 0;
        u32 tot;
        u32 ex = block_excl_scan<NT, u32>(v, tot, tmp);
        if (i < n) arr[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) arr[n] = carry;
    __syncthreads();
}

inline u32 grid_for(u64 n, u32 tpb, u32 cap = 1u << 20) {
    u64 g = (n + tpb - 1) / tpb;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return (u32)g;
}

}  // namespace pma
