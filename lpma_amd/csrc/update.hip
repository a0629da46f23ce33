// update.hip -- the batched edge-update path on gfx950.
//
// Reference pipeline (update_rpma, rpma.cuh:1549-1718) and what replaces it:
//   sort (:1557)                      -> sort.hip (packed keys, LDS radix)
//   locate_leaf_kernel (:253-316)     -> k_locate: dedup + separator search
//                                        + presence check (handle_del_mod,
//                                        gpma.cuh:162-175)
//   remove_if / RLE / scan (:447-511) -> k_locate + k_compact: one block
//                                        scan of (count, delta) pairs
//   per-level host loop (:1611-1657)  -> k_density: every run picks the
//                                        lowest level whose window fits, in
//                                        one pass, from the count pyramid
//   window grouping per level          -> k_claim / k_leaders: the maximal
//                                        chosen windows are elected, no host
//                                        round trip per level
//   block_rebalancing_kernel (:765-847) and rebalancing_kernel (:1127-1211)
//                                     -> k_rebalance_{wave,block} (LDS) and
//                                        the multi-workgroup global path
//   resize / rebuild (:1301-1530)     -> rebuild_push + rebuild_spread
//   row-offset fixup (:652-657, gpma.cuh:340-344) -> fused into the spread.
//
// Every kernel works on a Group: one or many independent PMAs whose batch
// elements carry their PMA id, so a round over many PMAs costs one launch
// sequence.  A run (the updates of one segment) is keyed pid << 32 | seg.
#include <hip/hip_ext.h>

#include "common.hpp"
#include "kernels.hpp"
#include "begin.hpp"

namespace pma {

static constexpr int LOC_TPB = 256;
static constexpr int LOC_TILE = LOC_TPB;  // one batch element per thread

u32 locate_blocks(u64 n) { return (u32)((n + LOC_TILE - 1) / LOC_TILE); }

// One workgroup per PMA: zeroes the run status and the PMA's outcome,
// stores inline batch descriptors, and finishes the count pyramid left by
// the previous batch: the levels above the per-workgroup chunks of
// k_pyr_chunk (nseg >> 11 entries and fewer), so no separate launch is
// needed for them.
__global__ __launch_bounds__(1024) void k_batch_begin(Group g, Status *s, InlineSegs in, BatchSeg *segs, u32 *zero,
                                                      u32 nzero, u32 *wfirst);

hipError_t batch_begin(const Group &g, Status *status, const InlineSegs &in, BatchSeg *segs, u32 *zero, u32 nzero,
                       u32 *wfirst, hipStream_t st) {
    k_batch_begin<<<g.npma, 1024, 0, st>>>(g, status, in, segs, zero, nzero, wfirst);
    return hipGetLastError();
}

// ------------------------------------------------ dedup, locate, presence

// Dedup (keep the last of equal keys: later requests sort after earlier
// ones, and within a request deletes before inserts), route each key to
// its segment, and resolve it against the keys present there.
// ONE: a lone PMA whose geometry is a kernel argument (scalar registers);
// otherwise it is read from the table per element
template <bool ONE>
__device__ __forceinline__ DevState geom(const Group &g, u32 pid) {
    if (ONE) return g.one;
    return g.tab[pid].d;
}

template <bool ONE>
__global__ __launch_bounds__(LOC_TPB) void k_locate(Group g, const u64 *__restrict__ pk, u64 n,
                                                    u64 *__restrict__ ukey, u64 *__restrict__ useg,
                                                    i32 *__restrict__ part_cnt, i32 *__restrict__ part_dl) {
    if (g.rh) n = g.rh->n;
    const u64 pmask = (1ull << g.pshift) - 1;
    const u64 i = (u64)blockIdx.x * LOC_TILE + threadIdx.x;
    i32 eff = 0, dl = 0;
    if (i < n) {
        const u64 p = pk[i];
        const u32 ks = 1 + g.sbits;  // key (and pid) above the request order and op
        const bool dup = (i + 1 < n) && ((pk[i + 1] >> ks) == (p >> ks));
        // a PMA id out of range only follows a failed sort (FLAG_SORT_STALL)
        if (dup || (u32)(p >> g.pshift) >= g.npma) {
            useg[i] = RK_NONE;
        } else {
            const u32 pid = (u32)(p >> g.pshift);
            const DevState d = geom<ONE>(g, pid);
            const u64 kk = (p & pmask) >> ks;
            const u32 op = (u32)(p & 1);
            const u64 key = ((kk >> d.vbits) << 32) | (kk & ((1ull << d.vbits) - 1));
            const u32 s = locate_segment(d.sep, d.H, key);
            const u32 present = segment_holds(d, s, key) ? 1u : 0u;
            u64 rk = RK_NONE;
            if (op != present) {
                eff = 1;
                dl = op ? 1 : -1;
                rk = ((u64)pid << 32) | s | (op ? RK_OP : 0ull);
            }
            ukey[i] = key;
            useg[i] = rk;
        }
    }
    __shared__ i32 tmp[2][LOC_TPB / 64];
    i32 ce = wave_sum(eff), cd = wave_sum(dl);
    if (lane_id() == 0) { tmp[0][threadIdx.x >> 6] = ce; tmp[1][threadIdx.x >> 6] = cd; }
    __syncthreads();
    if (threadIdx.x == 0) {
        i32 a = 0, c = 0;
        for (int w = 0; w < LOC_TPB / 64; ++w) { a += tmp[0][w]; c += tmp[1][w]; }
        part_cnt[blockIdx.x] = a;
        part_dl[blockIdx.x] = c;
    }
}

__global__ __launch_bounds__(1024) void k_scan_parts(i32 *__restrict__ pc, i32 *__restrict__ pdl, u32 nb,
                                                     Status *__restrict__ status, i32 *__restrict__ pd) {
    __shared__ i32 tmp[1024 / 64 + 1];
    i32 carry_c = 0, carry_d = 0;
    for (u32 b = 0; b < nb; b += 1024) {
        u32 i = b + threadIdx.x;
        i32 vc = i < nb ? pc[i] : 0;
        i32 vd = i < nb ? pdl[i] : 0;
        i32 tc, td;
        i32 ec = block_excl_scan<1024, i32>(vc, tc, tmp);
        i32 ed = block_excl_scan<1024, i32>(vd, td, tmp);
        if (i < nb) { pc[i] = carry_c + ec; pdl[i] = carry_d + ed; }
        carry_c += tc;
        carry_d += td;
    }
    if (threadIdx.x == 0) {
        status->n_eff = (u32)carry_c;
        status->delta = carry_d;
        pd[carry_c] = carry_d;
    }
}

static constexpr u32 COMPACT_SELF_PREFIX_MAX = 2048;  // blocks

// Stream compaction of the effective updates.  With few blocks every block
// derives its prefix (and the totals) from the per-block partials itself;
// otherwise k_scan_parts has already scanned them.
__global__ __launch_bounds__(LOC_TPB) void k_compact(const Group g, const u64 *__restrict__ ukey,
                                                     const u64 *__restrict__ useg, u64 n, const i32 *__restrict__ pc,
                                                     const i32 *__restrict__ pdl, u32 nb,
                                                     int self_prefix, u64 *__restrict__ ckey, u64 *__restrict__ cseg,
                                                     i32 *__restrict__ pd, u32 *__restrict__ wfirst,
                                                     Status *__restrict__ status) {
    const RoundHdr *rh = g.rh;
    __shared__ i32 tmp[LOC_TPB / 64 + 1];
    __shared__ i32 red[4][LOC_TPB / 64];
    i32 base_c, base_d;
    if (rh) n = rh->n;
    if (self_prefix) {
        i32 a = 0, b = 0, ta = 0, tb = 0;
        for (u32 x = threadIdx.x; x < nb; x += LOC_TPB) {
            i32 v = pc[x], w = pdl[x];
            ta += v;
            tb += w;
            if (x < blockIdx.x) { a += v; b += w; }
        }
        a = wave_sum(a); b = wave_sum(b); ta = wave_sum(ta); tb = wave_sum(tb);
        if (lane_id() == 0) {
            red[0][threadIdx.x >> 6] = a; red[1][threadIdx.x >> 6] = b;
            red[2][threadIdx.x >> 6] = ta; red[3][threadIdx.x >> 6] = tb;
        }
        __syncthreads();
        base_c = 0; base_d = 0; ta = 0; tb = 0;
        for (int w = 0; w < LOC_TPB / 64; ++w) {
            base_c += red[0][w]; base_d += red[1][w]; ta += red[2][w]; tb += red[3][w];
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            status->n_eff = (u32)ta;
            status->delta = tb;
            pd[ta] = tb;
        }
    } else {
        base_c = pc[blockIdx.x];
        base_d = pdl[blockIdx.x];
    }
    const u64 i = (u64)blockIdx.x * LOC_TILE + threadIdx.x;
    const u64 sg = i < n ? useg[i] : RK_NONE;
    i32 ce = 0, cd = 0;
    if (sg != RK_NONE) { ce = 1; cd = (sg & RK_OP) ? 1 : -1; }
    i32 t1, t2;
    i32 ex = block_excl_scan<LOC_TPB, i32>(ce, t1, tmp) + base_c;
    i32 dx = block_excl_scan<LOC_TPB, i32>(cd, t2, tmp) + base_d;
    if (sg != RK_NONE) {
        ckey[ex] = ukey[i];
        cseg[ex] = sg & ~RK_OP;
        pd[ex] = dx;
    }
    if (g.npma == 1) {
        // update index: the first update of each index-level window.  The
        // wave's updates are sorted, so only the first of each window in
        // the wave competes (atomic min) with the other waves'.
        const u32 H = pma_of(g, 0).H, lb = index_level(H);
        if (lb <= H) {
            const bool e = sg != RK_NONE;
            const u32 w = e ? (u32)sg >> lb : 0u;  // low word of a run key: the segment
            const u64 below = __ballot(e) & ((1ull << lane_id()) - 1ull);
            const int p = below ? 63 - __clzll((long long)below) : (int)lane_id();
            const u32 wp = (u32)__shfl((int)w, p, 64);
            if (e && (below == 0 || wp != w)) atomicMin(&wfirst[w], (u32)ex);
        }
    }
}

hipError_t locate_and_compact(const Group &g, const u64 *pk, u64 n, const BatchBufs &b, hipStream_t st) {
    u32 nb = locate_blocks(n);
    i32 *pc = b.part, *pdl = b.part + nb + 1;
    if (g.inline_one) k_locate<true><<<nb, LOC_TPB, 0, st>>>(g, pk, n, b.ukey, b.useg, pc, pdl);
    else k_locate<false><<<nb, LOC_TPB, 0, st>>>(g, pk, n, b.ukey, b.useg, pc, pdl);
    int self_prefix = nb <= COMPACT_SELF_PREFIX_MAX;
    if (!self_prefix) k_scan_parts<<<1, 1024, 0, st>>>(pc, pdl, nb, b.status, b.pd);
    k_compact<<<nb, LOC_TPB, 0, st>>>(g, b.ukey, b.useg, n, pc, pdl, nb, self_prefix, b.ckey, b.cseg, b.pd,
                                      b.wfirst, b.status);
    return hipGetLastError();
}

// ------------------------------------------------------------ density check

// first index >= start with a[idx] >= target, given a[start-1] < target;
// the exponential search starts with strides of `step` (a guess of the
// distance)
__device__ __forceinline__ u32 gallop_fwd(const u64 *__restrict__ a, u32 start, u32 n, u64 target, u32 step) {
    u32 lo = start;
    if (lo >= n || a[lo] >= target) return lo;
    u32 hi = (n - lo > step) ? lo + step : n;
    while (hi < n && a[hi] < target) {
        lo = hi;
        step <<= 1;
        hi = (n - lo > step) ? lo + step : n;
    }
    // a[lo] < target, hi == n or a[hi] >= target
    u32 l = lo + 1, h = hi;
    while (l < h) {
        u32 m = (l + h) >> 1;
        if (a[m] < target) l = m + 1; else h = m;
    }
    return l;
}

// smallest index <= start whose suffix [idx, start] is >= target, given
// a[start] >= target; strides start at `step` as above
__device__ __forceinline__ u32 gallop_back(const u64 *__restrict__ a, u32 start, u64 target, u32 step) {
    if (start == 0 || a[start - 1] < target) return start;
    u32 hi = start - 1;  // a[hi] >= target
    u32 lo;
    while (true) {
        lo = hi >= step ? hi - step : 0;
        if (a[lo] < target) break;
        if (lo == 0) return 0;
        hi = lo;
        step <<= 1;
    }
    // a[lo] < target <= a[hi]
    u32 l = lo + 1, h = hi;
    while (l < h) {
        u32 m = (l + h) >> 1;
        if (a[m] < target) l = m + 1; else h = m;
    }
    return l;
}

// levels climbed with galloping searches before the lock-step searches (a
// group of PMAs; a lone PMA switches to its update index at index_level(H))
static constexpr u32 DENSITY_GALLOP_LEVELS = 8;
// updates (of the window below) from which a window is claimed on arrival
static constexpr u32 DENSITY_CLAIM_ON_ARRIVAL = 64;
// levels evaluated together above the gallops
static constexpr u32 DENSITY_LOCKSTEP = 8;

// Update-index lookup table of a block: F[x] = first compacted update at or
// after index-level window x (n_eff past the last), from the per-window
// firsts of k_compact by a suffix minimum.  nW <= WFIRST_N.
static constexpr u32 WF_PER = (WFIRST_N + 1 + 255) / 256 | 1;  // odd: LDS banks apart
__device__ __forceinline__ void fill_window_firsts(const u32 *__restrict__ wfirst, u32 nW, u32 n_eff, u32 *F,
                                                   u32 *wtot) {
    const u32 b0 = threadIdx.x * WF_PER;
    u32 v[WF_PER];
#pragma unroll
    for (u32 k = 0; k < WF_PER; ++k) v[k] = b0 + k < nW ? min(wfirst[b0 + k], n_eff) : n_eff;
    u32 m = n_eff;
#pragma unroll
    for (int k = WF_PER - 1; k >= 0; --k) {
        m = min(m, v[k]);
        v[k] = m;  // suffix minimum inside the thread's run
    }
    // suffix minimum over the later threads
    const u32 lane = lane_id(), wv = threadIdx.x >> 6;
    u32 x = m;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const u32 y = (u32)__shfl_down((int)x, o, 64);
        if (lane + o < 64) x = min(x, y);
    }
    if (lane == 0) wtot[wv] = x;
    const u32 nxt = (u32)__shfl_down((int)x, 1, 64);
    __syncthreads();
    u32 after = lane < 63 ? nxt : n_eff;
    for (u32 w = wv + 1; w < blockDim.x / 64; ++w) after = min(after, wtot[w]);
#pragma unroll
    for (u32 k = 0; k < WF_PER; ++k)
        if (b0 + k <= nW) F[b0 + k] = min(v[k], after);
    __syncthreads();
}

// Every run leader (first update of a run) climbs the count pyramid from its
// segment until a window fits after the batch, and claims it (atomic max of
// epoch << 32 | ~run: the smallest run index wins); the window's updates are
// counted by galloping over the sorted run keys.  Large windows (many
// updates below them) are evaluated once: a run arriving there claims it
// first, and only the first run to arrive climbs on; a later one stops there
// (its updates lie in that window, which the first run either takes or
// leaves for a larger window that contains it).  k_leaders keeps a claimed
// window only if no claimed window of a higher level contains it, so a
// window that was passed through never survives.
// Above the gallops, a lone PMA reads its windows' update ranges off the
// update index (two LDS lookups per level, every level's loads independent);
// one run per index-level window climbs on (the first in its wave, then the
// first to arrive).  A group searches the bounds of several levels in lock
// step.  A lone PMA that would hold more keys than it has slots (live + the
// batch's net delta) is resized without any climb.  The first and last
// update of each PMA record its compacted range.
template <bool ONE>
__global__ __launch_bounds__(256) void k_density(Group g, const u64 *__restrict__ cseg, const i32 *__restrict__ pd,
                                                 const u32 *__restrict__ wfirst, const Status *__restrict__ status_in,
                                                 u32 *__restrict__ rlevel, uint4 *__restrict__ rwin, Status *status) {
    __shared__ u32 F[WFIRST_N + 1];
    __shared__ u32 wtot[256 / 64];
    const u32 n_eff = status_in->n_eff;
    const u32 j = blockIdx.x * blockDim.x + threadIdx.x;
    const bool lone = g.npma == 1;
    u32 local_flags = 0;
    u32 pid = 0xFFFFFFFFu, lvbit = 0;
    bool leader = false, stopped = false, fit = false;
    u32 l = 0, u0 = 0, u1 = 0, s = 0;
    u32 fC = 0, fM = 0;  // the fitting window's count before and after the batch
    u64 pb = 0;
    bool grows = false;  // a lone PMA over its slots: resized, no climb
    if (j < n_eff) {
        const u64 rk = cseg[j];
        pid = (u32)(rk >> 32);
        const u64 prev = j > 0 ? cseg[j - 1] : RK_NONE;
        if (j == 0 || (u32)(prev >> 32) != pid) {
            g.pstat[pid].u0 = j;
            g.pstat[pid].pd0 = pd[j];
        }
        if (j + 1 == n_eff || (u32)(cseg[j + 1] >> 32) != pid) {
            g.pstat[pid].u1 = j + 1;
            g.pstat[pid].pd1 = pd[j + 1];
        }
        leader = j == 0 || prev != rk;
        if (leader) {
            // density bounds from the device table (a lone PMA's entry at a
            // fixed address): they stay out of the kernel arguments, which
            // every launch of the pipeline copies
            const Density &dn = g.tab[ONE ? 0 : pid].dn;
            const DevState d = geom<ONE>(g, pid);
            pb = (u64)pid << 32;
            s = (u32)rk;
            const u64 nseg = nseg_of(d);
            const u64 tag = (u64)group_epoch(g) << 32;
            u0 = j;
            u1 = j + 1;
            if (lone) {
                // more keys than slots: along some run's path every window
                // overflows its capacity (a full window has a child over its
                // own), so that run reaches the root and the PMA grows
                const i64 mroot = (i64)*count_ptr(d, d.H, 0) + (i64)status_in->delta;
                if (mroot > (i64)(nseg << d.log2S)) {
                    l = d.H + 1;
                    grows = true;
                }
            }
            const u32 lg = lone ? index_level(d.H) : DENSITY_GALLOP_LEVELS + 1;
            for (; l <= d.H && l < lg; ++l) {
                const u64 w = (u64)s >> l;
                const u64 a = w << l;
                // windows with many updates are claimed on arrival (only the
                // first run to arrive climbs on); small ones are evaluated by
                // every run that reaches them and claimed where they fit
                const bool on_arrival = l > 0 && u1 - u0 >= DENSITY_CLAIM_ON_ARRIVAL;
                if (on_arrival) {
                    const u64 old = atomicMax((unsigned long long *)&d.owner[owner_idx(nseg, l, w)],
                                              (unsigned long long)(tag | (u64)(~j)));
                    if ((old & ~0xFFFFFFFFull) == tag) {
                        stopped = true;
                        break;
                    }
                }
                // the window grows by its sibling, which holds about as many
                // runs as the window had: stride guess for the one boundary
                // that moves
                const u32 guess = max(1u, u1 - u0);
                const u64 C = *count_ptr(d, l, w);  // independent of the searches: in flight meanwhile
                u0 = gallop_back(cseg, u0, pb + a, guess);
                u1 = gallop_fwd(cseg, u1, n_eff, pb + a + (1ull << l), guess);
                const i64 m = (i64)C + (i64)(pd[u1] - pd[u0]);
                if (window_fits(dn, l, d.H, (u64)m)) {
                    if (l > 0 && l < d.H && !on_arrival)
                        atomicMax((unsigned long long *)&d.owner[owner_idx(nseg, l, w)],
                                  (unsigned long long)(tag | (u64)(~j)));
                    fit = true;
                    fC = (u32)C;
                    fM = (u32)m;
                    break;
                }
            }
        }
    }
    bool climbing = leader && !stopped && !fit;
    if (lone) {
        const DevState d = geom<ONE>(g, 0);
        const u32 lb = index_level(d.H);
        bool high = climbing && l == lb && lb <= d.H;
        // one run per index-level window climbs on: the first in its wave
        // (the wave's runs are sorted), then the first to arrive (claim)
        const u32 wl = (u32)((u64)s >> lb);
        const u64 below = __ballot(high) & ((1ull << lane_id()) - 1ull);
        const int pl = below ? 63 - __clzll((long long)below) : (int)lane_id();
        const u32 wpl = (u32)__shfl((int)wl, pl, 64);
        if (high && below != 0 && wpl == wl) {
            high = false;
            stopped = true;
        }
        const u64 nseg = nseg_of(d);
        const u64 tag = (u64)group_epoch(g) << 32;
        if (high && lb > 0 && lb < d.H) {
            const u64 old = atomicMax((unsigned long long *)&d.owner[owner_idx(nseg, lb, wl)],
                                      (unsigned long long)(tag | (u64)(~j)));
            if ((old & ~0xFFFFFFFFull) == tag) {
                high = false;
                stopped = true;
            }
        }
        if (__syncthreads_or(high)) fill_window_firsts(wfirst, 1u << (d.H - lb), n_eff, F, wtot);
        if (high) {
            const Density &dn = g.tab[0].dn;
            while (!fit && l <= d.H) {
                u64 Cl[DENSITY_LOCKSTEP];
                i32 p0[DENSITY_LOCKSTEP], p1[DENSITY_LOCKSTEP];
                u32 a0[DENSITY_LOCKSTEP], a1[DENSITY_LOCKSTEP];
                const u32 ng = min(DENSITY_LOCKSTEP, d.H + 1 - l);
#pragma unroll
                for (u32 q = 0; q < DENSITY_LOCKSTEP; ++q) {
                    const u32 lq = l + q;
                    if (q < ng) {
                        const u64 w = (u64)s >> lq;
                        Cl[q] = *count_ptr(d, lq, w);
                        a0[q] = F[(u32)(w << (lq - lb))];
                        a1[q] = F[(u32)((w + 1) << (lq - lb))];
                        p0[q] = pd[a0[q]];
                        p1[q] = pd[a1[q]];
                    }
                }
                for (u32 q = 0; q < ng; ++q, ++l) {
                    const i64 m = (i64)Cl[q] + (i64)(p1[q] - p0[q]);
                    if (window_fits(dn, l, d.H, (u64)m)) {
                        if (l > lb && l < d.H)
                            atomicMax((unsigned long long *)&d.owner[owner_idx(nseg, l, (u64)s >> l)],
                                      (unsigned long long)(tag | (u64)(~j)));
                        fit = true;
                        u0 = a0[q];
                        u1 = a1[q];
                        fC = (u32)Cl[q];
                        fM = (u32)m;
                        break;
                    }
                }
            }
        }
    } else if (climbing) {
        // Higher up (rare: a run still climbing there leads a large region
        // that is full or sparse), the bounds of several levels' windows are
        // found together: their binary searches step in lock step, so the
        // dependent loads are one search deep, not one per level; the
        // claims still go level by level.
        const DevState d = geom<ONE>(g, pid);
        const Density &dn = g.tab[ONE ? 0 : pid].dn;
        const u64 nseg = nseg_of(d);
        const u64 tag = (u64)group_epoch(g) << 32;
        while (!stopped && !fit && l <= d.H) {
            constexpr u32 G = DENSITY_LOCKSTEP;
            u32 lo0[G], hi0[G], lo1[G], hi1[G];
            u64 t0[G], t1[G], Cl[G];
            const u32 ng = min(G, d.H + 1 - l);
#pragma unroll
            for (u32 q = 0; q < G; ++q) {
                const u32 lq = l + q;
                const u64 a = ((u64)s >> lq) << lq;
                t0[q] = pb + a;
                t1[q] = pb + a + (1ull << lq);
                lo0[q] = 0; hi0[q] = q < ng ? u0 : 0u;      // the start lies at or before u0
                lo1[q] = u1; hi1[q] = q < ng ? n_eff : u1;  // the end at or after u1
                Cl[q] = q < ng ? *count_ptr(d, lq, (u64)s >> lq) : 0ull;
            }
            bool more = true;
            while (more) {
                more = false;
                u64 v0[G], v1[G];
#pragma unroll
                for (u32 q = 0; q < G; ++q) {
                    v0[q] = lo0[q] < hi0[q] ? cseg[(lo0[q] + hi0[q]) >> 1] : 0ull;
                    v1[q] = lo1[q] < hi1[q] ? cseg[(lo1[q] + hi1[q]) >> 1] : 0ull;
                }
#pragma unroll
                for (u32 q = 0; q < G; ++q) {
                    if (lo0[q] < hi0[q]) {
                        const u32 mid = (lo0[q] + hi0[q]) >> 1;
                        if (v0[q] < t0[q]) lo0[q] = mid + 1; else hi0[q] = mid;
                    }
                    if (lo1[q] < hi1[q]) {
                        const u32 mid = (lo1[q] + hi1[q]) >> 1;
                        if (v1[q] < t1[q]) lo1[q] = mid + 1; else hi1[q] = mid;
                    }
                    more |= lo0[q] < hi0[q] || lo1[q] < hi1[q];
                }
            }
            i32 p0[G], p1[G];
#pragma unroll
            for (u32 q = 0; q < G; ++q) {
                p0[q] = q < ng ? pd[lo0[q]] : 0;
                p1[q] = q < ng ? pd[lo1[q]] : 0;
            }
            for (u32 q = 0; q < ng; ++q, ++l) {
                const u64 w = (u64)s >> l;
                const u64 old = atomicMax((unsigned long long *)&d.owner[owner_idx(nseg, l, w)],
                                          (unsigned long long)(tag | (u64)(~j)));
                if ((old & ~0xFFFFFFFFull) == tag) {
                    stopped = true;
                    break;
                }
                u0 = lo0[q];
                u1 = lo1[q];
                const i64 m = (i64)Cl[q] + (i64)(p1[q] - p0[q]);
                if (window_fits(dn, l, d.H, (u64)m)) {
                    fit = true;
                    fC = (u32)Cl[q];
                    fM = (u32)m;
                    break;
                }
            }
        }
    }
    if (leader) {
        const u32 H = geom<ONE>(g, pid).H;
        rlevel[j] = l;
        // (every run of a growing PMA finds the same: the first raises it)
        if (!stopped && (!grows || j == 0)) {
            if (l >= H) local_flags = (l == H) ? FLAG_ROOT : FLAG_RESIZE;
            else if (l > 0) lvbit = 1u << l;
        }
        // the window the election may choose: its updates and counts.  A run
        // stopped by an earlier arrival can still own the window (the claim
        // keeps the smallest run index) without having evaluated it: its
        // record is marked absent and the election finds the range itself.
        rwin[j] = (!stopped && fit && l < H) ? make_uint4(u0, u1, fC, fM) : make_uint4(~0u, 0u, 0u, 0u);
    }
    // the flags: one atomic per workgroup for a lone PMA, per wave for a
    // group (a batch that resizes raises them in many runs, and atomics on
    // one word serialise: thousands of them cost tens of microseconds); per
    // PMA the flags and the claimed levels: lanes of the first PMA in the
    // wave combine theirs, the others (rare) go alone (the levels into the
    // workgroup's replica of the PMA's mask)
    const u32 f = dpp_reduce32<false>(local_flags);
    if (lone) {
        __shared__ u32 s_flags;
        if (threadIdx.x == 0) s_flags = 0;
        __syncthreads();
        if (lane_id() == 0 && f) atomicOr(&s_flags, f);
        __syncthreads();
        if (threadIdx.x == 0 && s_flags) {
            atomicOr(&status->flags, s_flags);
            atomicOr(&g.pstat[0].flags, s_flags);
        }
        local_flags = 0;  // raised
    } else if (lane_id() == 0 && f) {
        atomicOr(&status->flags, f);
    }
    const u32 pid0 = __shfl(pid, 0, 64);
    const u32 rep = blockIdx.x & (LV_REPL - 1);
    const u32 lv = dpp_reduce32<false>(pid == pid0 ? lvbit : 0u);
    const u32 pf = dpp_reduce32<false>(pid == pid0 ? local_flags : 0u);
    if (lane_id() == 0 && pid0 != 0xFFFFFFFFu) {
        if (lv) atomicOr(&g.lvmask[(u64)pid0 * LV_REPL + rep], lv);
        if (pf) atomicOr(&g.pstat[pid0].flags, pf);
    }
    if (pid != pid0 && pid != 0xFFFFFFFFu) {
        if (lvbit) atomicOr(&g.lvmask[(u64)pid * LV_REPL + rep], lvbit);
        if (local_flags) atomicOr(&g.pstat[pid].flags, local_flags);
    }
}

hipError_t density_check(const Group &g, u64 n, const BatchBufs &b, hipStream_t st) {
    // one thread per update (no grid-stride loop): the grid covers n
    if (g.inline_one)
        k_density<true><<<grid_for(n, 256, 1u << 30), 256, 0, st>>>(g, b.cseg, b.pd, b.wfirst, b.status, b.rlevel,
                                                                    b.rwin, b.status);
    else
        k_density<false><<<grid_for(n, 256, 1u << 30), 256, 0, st>>>(g, b.cseg, b.pd, b.wfirst, b.status, b.rlevel,
                                                                     b.rwin, b.status);
    return hipGetLastError();
}

// ---------------------------------------------------------- window election

template <bool ONE>
__global__ __launch_bounds__(1024) void k_leaders(Group g, const u64 *__restrict__ cseg, const i32 *__restrict__ pd,
                          const u32 *__restrict__ rlevel, const uint4 *__restrict__ rwin, Work *w0, Work *w1,
                          Work *w2, Work *w3, Status *status) {
    const u32 stop = FLAG_INVALID_KEY | FLAG_SORT_STALL | (g.defer_rebuild ? FLAG_ROOT | FLAG_RESIZE : 0u);
    if (status->flags & stop) return;
    const u32 n_eff = status->n_eff;
    u32 j = blockIdx.x * blockDim.x + threadIdx.x;
    int cls = -1;
    Work wk;
    // claimed levels of the workgroup's first PMA: its replicas or-ed once
    __shared__ u32 s_mask, s_pid;
    const u32 jb = blockIdx.x * blockDim.x;
    if (threadIdx.x < 64 && jb < n_eff) {
        const u32 pb = (u32)(cseg[jb] >> 32);
        u32 m = threadIdx.x < LV_REPL ? g.lvmask[(u64)pb * LV_REPL + threadIdx.x] : 0u;
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) m |= __shfl_xor(m, o, 64);
        if (threadIdx.x == 0) {
            s_mask = m;
            s_pid = pb;
        }
    }
    __syncthreads();
    if (j < n_eff) {
        const u64 rk = cseg[j];
        const u32 pid = (u32)(rk >> 32);
        bool leader = (j == 0 || cseg[j - 1] != rk);
        // a root rebalance or a resize redoes the whole PMA: no windows needed
        if (leader && (g.pstat[pid].flags & (FLAG_ROOT | FLAG_RESIZE))) leader = false;
        const DevState d = geom<ONE>(g, pid);
        const u32 s = (u32)rk;
        u32 L = leader ? rlevel[j] : 0;
        u64 nseg = nseg_of(d);
        u64 w = (u64)s >> L;
        const u32 epoch = group_epoch(g);
        const u64 tag = (u64)epoch << 32;
        if (leader && L > 0 && d.owner[owner_idx(nseg, L, w)] != (tag | (u64)(~j))) leader = false;
        if (leader) {
            // inside a larger chosen window?  only the levels some run of the
            // PMA claimed can hold one; their loads are issued together
            u32 lv = s_pid == pid ? s_mask : 0u;
            if (s_pid != pid)
                for (u32 r = 0; r < LV_REPL; ++r) lv |= g.lvmask[(u64)pid * LV_REPL + r];
            lv &= ~((2u << L) - 1u);
            bool dom = false;
            while (lv) {
                u32 l2 = (u32)__builtin_ctz(lv);
                lv &= lv - 1;
                dom |= (d.owner[owner_idx(nseg, l2, (u64)s >> l2)] >> 32) == (u64)epoch;
            }
            if (dom) leader = false;
        }
        if (leader) {
            // the window's updates and counts, as the density check found
            // them, or found here for a run that stopped on arrival
            uint4 rw = rwin[j];
            const u64 a = w << L;
            if (rw.x == ~0u) {
                const u64 pb = (u64)pid << 32;
                rw.x = gallop_back(cseg, j, pb + a, 1);
                rw.y = gallop_fwd(cseg, j + 1, n_eff, pb + a + (1ull << L), 1);
                rw.z = *count_ptr(d, L, w);
                rw.w = (u32)((i32)rw.z + pd[rw.y] - pd[rw.x]);
            }
            wk.pid = pid;
            wk.a = (u32)a;
            wk.L = L;
            wk.u0 = rw.x;
            wk.u1 = rw.y;
            wk.C = rw.z;
            wk.M = rw.w;
            cls = work_class((u64)rw.z + (rw.y - rw.x), L, d.log2S);
        }
    }
    // list appends: LDS counters per class, then one global atomic per
    // workgroup and class (global atomics on one address serialise in L2)
    __shared__ u32 ccount[NUM_CLASSES], cbase[NUM_CLASSES];
    if (threadIdx.x < NUM_CLASSES) ccount[threadIdx.x] = 0;
    __syncthreads();
    u32 off = 0;
    if (cls >= 0) off = atomicAdd(&ccount[cls], 1u);
    __syncthreads();
    if (threadIdx.x < NUM_CLASSES && ccount[threadIdx.x])
        cbase[threadIdx.x] = atomicAdd(&status->nwork[threadIdx.x], ccount[threadIdx.x]);
    __syncthreads();
    if (cls >= 0) {
        Work *dst = cls == CLASS_WAVE ? w0 : (cls == CLASS_BLOCK ? w1 : (cls == CLASS_LARGE ? w2 : w3));
        dst[cbase[cls] + off] = wk;
    }
}

__global__ void k_reset_work(Status *s) {
    if (threadIdx.x < NUM_CLASSES) s->nwork[threadIdx.x] = 0;
}

hipError_t reset_work_counts(Status *status, hipStream_t st) {
    k_reset_work<<<1, 64, 0, st>>>(status);
    return hipGetLastError();
}

// n bounds the number of runs (the batch size); the kernel reads the exact
// count from the status
hipError_t elect_windows(const Group &g, u64 n, const BatchBufs &b, hipStream_t st) {
    if (g.inline_one)
        k_leaders<true><<<grid_for(n, 1024, 1u << 30), 1024, 0, st>>>(g, b.cseg, b.pd, b.rlevel, b.rwin, b.work[0], b.work[1],
                                                                      b.work[2], b.work[3], b.status);
    else
        k_leaders<false><<<grid_for(n, 1024, 1u << 30), 1024, 0, st>>>(g, b.cseg, b.pd, b.rlevel, b.rwin, b.work[0], b.work[1],
                                                                       b.work[2], b.work[3], b.status);
    return hipGetLastError();
}

// ------------------------------------------------------- LDS window rebalance
//
// One group of NT threads rebalances one window entirely in LDS:
//   1. prefix of the window's segment counts,
//   2. coalesced load of the live keys (compacted) and of the updates,
//   3. merge by rank: each old key / insert finds its output rank with one
//      binary search over the other list (deleted keys drop out),
//   4. coalesced store of every slot of the window after the even spread
//      (rpma.cuh:676-679), KEY_NONE in the gaps, marker slots into row_off,
//   5. per-segment counts and separators.

struct BlockStats {
    u64 updates, markers, bytes;
    u32 *windows;  // LDS, per level
};

template <int NT, u32 ECAP, u32 SEGCAP>
__device__ void rebalance_window(const DevState &d, const Work &wk, const u64 *__restrict__ ckey,
                                 const i32 *__restrict__ pd, u64 *A, u64 *B, i32 *P, u32 *pre, u32 *stmp,
                                 BlockStats &bs) {
    const u32 tid = threadIdx.x;
    const u32 L = wk.L;
    const u64 a = wk.a;
    const u32 nsegw = 1u << L;
    const u32 C = wk.C;
    const u32 k = wk.u1 - wk.u0;
    const u32 log2S = d.log2S;
    const u32 Smask = (1u << log2S) - 1;
    const u64 base = a << log2S;
    const u32 W = nsegw << log2S;

    for (u32 x = tid; x < nsegw; x += NT) pre[x] = d.cnt[a + x];
    __syncthreads();
    block_scan_array<NT>(pre, nsegw, stmp);

    for (u32 t = tid; t < W; t += NT) {
        u32 seg = t >> log2S, lane = t & Smask;
        u32 p0 = pre[seg];
        if (lane < pre[seg + 1] - p0) A[p0 + lane] = d.keys[base + t];
    }
    for (u32 q = tid; q < k; q += NT) A[C + q] = ckey[wk.u0 + q];
    const i32 pbase = pd[wk.u0];
    for (u32 q = tid; q <= k; q += NT) P[q] = pd[wk.u0 + q] - pbase;
    __syncthreads();

    const u32 M = (u32)((i32)C + P[k]);
    const u64 *U = A + C;
    for (u32 r = tid; r < C; r += NT) {
        u64 x = A[r];
        u32 lo = 0, hi = k;
        while (lo < hi) {
            u32 m = (lo + hi) >> 1;
            if (U[m] < x) lo = m + 1; else hi = m;
        }
        if (lo < k && U[lo] == x) continue;  // deleted
        B[r + P[lo]] = x;
    }
    for (u32 q = tid; q < k; q += NT) {
        if (P[q + 1] - P[q] != 1) continue;  // deletes produce nothing
        u64 y = U[q];
        u32 lo = 0, hi = C;
        while (lo < hi) {
            u32 m = (lo + hi) >> 1;
            if (A[m] < y) lo = m + 1; else hi = m;
        }
        B[lo + P[q]] = y;
    }
    __syncthreads();

    u32 markers = 0;
    for (u32 t = tid; t < W; t += NT) {
        u32 seg = t >> log2S, lane = t & Smask;
        u64 f = spread_first(seg, M, L);
        u64 g = spread_first(seg + 1, M, L);
        u64 key = (lane < g - f) ? B[f + lane] : KEY_NONE;
        d.keys[base + t] = key;
        if (is_marker(key) && (u32)(key >> 32) < d.V) {
            d.row_off[(u32)(key >> 32) + 1] = (u32)(base + t);
            ++markers;
        }
    }
    const u64 sep_a = d.sep[a];
    for (u32 x = tid; x < nsegw; x += NT) {
        u64 f = spread_first(x, M, L);
        u64 g = spread_first(x + 1, M, L);
        d.cnt[a + x] = (u32)(g - f);
        if (x > 0) d.sep[a + x] = (g > f) ? B[f] : (f > 0 ? B[f - 1] + 1 : sep_a);
    }
    markers = wave_sum(markers);
    if (lane_id() == 0) bs.markers += markers;
    if (tid == 0) {
        bs.windows[L] += 1;
        bs.updates += k;
        bs.bytes += 16ull * W + 8ull * k;
    }
    __syncthreads();
}

__device__ void flush_stats(BlockStats &bs, u64 *stats_all, int cls_slot, u32 tid) {
    u64 *stats = stats_all + (u64)(blockIdx.x % ST_ROWS) * ST_NUM;
    // markers were accumulated by lane 0 of every wave
    if ((tid & 63) == 0 && bs.markers) {
        atomicAdd((unsigned long long *)&stats[ST_MARKERS], (unsigned long long)bs.markers);
        atomicAdd((unsigned long long *)&stats[cls_slot], (unsigned long long)(4ull * bs.markers));
    }
    if (tid == 0) {
        for (int l = 0; l < 32; ++l)
            if (bs.windows[l]) atomicAdd((unsigned long long *)&stats[ST_WINDOWS + l], (unsigned long long)bs.windows[l]);
        if (bs.bytes) {
            atomicAdd((unsigned long long *)&stats[ST_UPDATES], (unsigned long long)bs.updates);
            atomicAdd((unsigned long long *)&stats[cls_slot], (unsigned long long)bs.bytes);
        }
    }
}

template <int NT, u32 ECAP, u32 SEGCAP>
__device__ void rebalance_loop(const Group &g, const Work *__restrict__ work, const u32 *nwork,
                               const u64 *__restrict__ ckey, const i32 *__restrict__ pd, int cls_slot) {
    __shared__ u64 A[ECAP];
    __shared__ u64 B[ECAP];
    __shared__ i32 P[ECAP + 1];
    __shared__ u32 pre[SEGCAP + 1];
    __shared__ u32 stmp[NT / 64 + 1];
    __shared__ u32 wl[32];
    BlockStats bs;
    bs.updates = bs.markers = bs.bytes = 0;
    bs.windows = wl;
    if (threadIdx.x < 32) wl[threadIdx.x] = 0;
    __syncthreads();
    const u32 nw = *nwork;
    for (u32 i = blockIdx.x; i < nw; i += gridDim.x) {
        Work wk = work[i];
        const DevState d = pma_of(g, wk.pid);
        rebalance_window<NT, ECAP, SEGCAP>(d, wk, ckey, pd, A, B, P, pre, stmp, bs);
    }
    __syncthreads();
    flush_stats(bs, g.stats, cls_slot, threadIdx.x);
}

// One wave64 per window (E <= WAVE_E_CAP, <= WAVE_SEG_CAP segments).  The
// live keys are found with a ballot over the window's slots (segments are
// front-packed, gaps hold KEY_NONE), so neither the segment counts nor a scan
// sit on the dependency chain: the slot loads and the update loads are
// issued together, then one LDS merge, then the coalesced store.
// The lanes of one wave see each other's LDS stores: the barrier of code
// that each wave of a workgroup runs on its own (a one-wave window).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ void rebalance_window_wave(const DevState &d, const Work &wk, const u64 *__restrict__ ckey,
                                      const i32 *__restrict__ pd, u64 *A, u64 *B, i32 *P, BlockStats &bs) {
    const u32 lane = lane_id();
    const u64 lt = (1ull << lane) - 1ull;
    const u32 L = wk.L;
    const u64 a = wk.a;
    const u32 nsegw = 1u << L;
    const u32 C = wk.C;
    const u32 k = wk.u1 - wk.u0;
    const u32 log2S = d.log2S;
    const u32 Smask = (1u << log2S) - 1;
    const u64 base = a << log2S;
    const u32 W = nsegw << log2S;

    const i32 pbase = pd[wk.u0];
    for (u32 q = lane; q < k; q += 64) A[C + q] = ckey[wk.u0 + q];
    for (u32 q = lane; q <= k; q += 64) P[q] = pd[wk.u0 + q] - pbase;
    u32 off = 0;
    for (u32 t0 = 0; t0 < W; t0 += 256) {
        u64 v[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            u32 t = t0 + r * 64 + lane;
            v[r] = t < W ? d.keys[base + t] : KEY_NONE;
        }
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            bool live = v[r] != KEY_NONE;
            u64 m = __ballot(live);
            if (live) A[off + __popcll(m & lt)] = v[r];
            off += __popcll(m);
        }
    }
    wave_sync();

    const u32 M = (u32)((i32)C + P[k]);
    const u64 *U = A + C;
    for (u32 r = lane; r < C; r += 64) {
        u64 x = A[r];
        u32 lo = 0, hi = k;
        while (lo < hi) {
            u32 m = (lo + hi) >> 1;
            if (U[m] < x) lo = m + 1; else hi = m;
        }
        if (lo < k && U[lo] == x) continue;  // deleted
        B[r + P[lo]] = x;
    }
    for (u32 q = lane; q < k; q += 64) {
        if (P[q + 1] - P[q] != 1) continue;
        u64 y = U[q];
        u32 lo = 0, hi = C;
        while (lo < hi) {
            u32 m = (lo + hi) >> 1;
            if (A[m] < y) lo = m + 1; else hi = m;
        }
        B[lo + P[q]] = y;
    }
    wave_sync();

    u32 markers = 0;
    for (u32 t = lane; t < W; t += 64) {
        u32 seg = t >> log2S, ln = t & Smask;
        u64 f = spread_first(seg, M, L);
        u64 g = spread_first(seg + 1, M, L);
        u64 key = (ln < g - f) ? B[f + ln] : KEY_NONE;
        d.keys[base + t] = key;
        if (is_marker(key) && (u32)(key >> 32) < d.V) {
            d.row_off[(u32)(key >> 32) + 1] = (u32)(base + t);
            ++markers;
        }
    }
    const u64 sep_a = d.sep[a];
    for (u32 x = lane; x < nsegw; x += 64) {
        u64 f = spread_first(x, M, L);
        u64 g = spread_first(x + 1, M, L);
        d.cnt[a + x] = (u32)(g - f);
        if (x > 0) d.sep[a + x] = (g > f) ? B[f] : (f > 0 ? B[f - 1] + 1 : sep_a);
    }
    markers = wave_sum(markers);
    if (lane == 0) {
        bs.markers += markers;
        atomicAdd(&bs.windows[L], 1u);  // the waves of a workgroup share it
        bs.updates += k;
        bs.bytes += 16ull * W + 8ull * k;
    }
    wave_sync();
}

__global__ __launch_bounds__(64) void k_rebalance_wave(Group g, const Work *__restrict__ work,
                                                       const Status *__restrict__ status,
                                                       const u64 *__restrict__ ckey, const i32 *__restrict__ pd) {
    __shared__ u64 A[WAVE_E_CAP];
    __shared__ u64 B[WAVE_E_CAP];
    __shared__ i32 P[WAVE_E_CAP + 1];
    __shared__ u32 wl[32];
    BlockStats bs;
    bs.updates = bs.markers = bs.bytes = 0;
    bs.windows = wl;
    if (threadIdx.x < 32) wl[threadIdx.x] = 0;
    __syncthreads();
    const u32 nw = status->nwork[CLASS_WAVE];
    for (u32 i = blockIdx.x; i < nw; i += gridDim.x) {
        Work wk = work[i];
        const DevState d = pma_of(g, wk.pid);
        rebalance_window_wave(d, wk, ckey, pd, A, B, P, bs);
    }
    __syncthreads();
    flush_stats(bs, g.stats, ST_ALG_BYTES_WAVE, threadIdx.x);
}

__global__ __launch_bounds__(256) void k_rebalance_block(Group g, const Work *__restrict__ work,
                                                         const Status *__restrict__ status,
                                                         const u64 *__restrict__ ckey, const i32 *__restrict__ pd) {
    rebalance_loop<256, BLOCK_E_CAP, BLOCK_SEG_CAP>(g, work, &status->nwork[CLASS_BLOCK], ckey, pd,
                                                    ST_ALG_BYTES_BLOCK);
}

// Single-segment windows (level 0, E <= 2S): a group of G = S/SPL lanes (at
// most 64) rebalances one segment, SPL slots per lane, so a wave works on
// 64/G segments at once and no workgroup barrier is needed.  With one
// segment the even spread is the merged order itself, front-packed.  Every
// global load of a window (its S slots, its first G updates and their
// prefixes) is issued before the first of them is used, so a window costs
// one memory latency, and the old keys stay in registers: for each update
// in turn the group compares it with all of its keys at once, which gives
// every old key its shift (the net inserts before it) and the update its
// rank among the old keys (a popcount of the group's ballot), with no
// dependent LDS reads.  Only the updates and their prefixes go through the
// group's LDS slice.  Every key is loaded before the first store, so the
// in-place stores cannot overwrite a key another lane has yet to read.
static constexpr int SEG_TPB = 256;
// persistent grid: four workgroups per CU of 256 at most; each group loops
// over windows (fewer, longer-lived workgroups keep more of the device busy
// than one short workgroup per 32 windows)
static constexpr u32 SEG_GRID_MAX = 1024;
// workgroups of four waves for a lone PMA's one-wave windows in the same
// launch (resident beside the single-segment ones: 1536 workgroups of at
// most 21 KB of LDS)
static constexpr u32 EARLY_WAVE_GRID = 512;

// u32 words of a group's LDS slice: the updates (k <= 2S, u64) and their
// prefixes (k + 1 <= 2S + 1, i32), padded so that the slices of a wave's
// groups start on different banks
__host__ __device__ constexpr u32 seg_slice_words(u32 S) { return 4 * S + (2 * S + 2) + 16; }

// With wwork, a lone PMA's one-wave windows run in the same launch, each
// wave on its own (the LDS then holds one window's arrays per wave): the
// workgroups past the first sgrid take the first of them while the first
// sgrid take the single-segment windows, and these take the rest after
// theirs (wave w of the launch's one-wave order: the extra workgroups'
// waves, then the others').
static constexpr u32 WAVE_LDS_BYTES = 2 * WAVE_E_CAP * 8 + (WAVE_E_CAP + 1) * 4 + 4;  // A, B, P (8-aligned)

// one-wave windows i = first + wave, first + wave + stride, ... of a
// workgroup, each of its waves on its own
__device__ void rebalance_waves_of(const Group &g, const Work *__restrict__ wwork, const Status *__restrict__ status,
                                   const u64 *__restrict__ ckey, const i32 *__restrict__ pd, unsigned char *lds,
                                   u32 first, u32 stride) {
    __shared__ u32 wl[32];
    if (threadIdx.x < 32) wl[threadIdx.x] = 0;
    __syncthreads();
    BlockStats bs;
    bs.updates = bs.markers = bs.bytes = 0;
    bs.windows = wl;
    const u32 wv = threadIdx.x >> 6;
    u64 *A = (u64 *)(lds + (u64)wv * WAVE_LDS_BYTES);
    u64 *B = A + WAVE_E_CAP;
    i32 *P = (i32 *)(B + WAVE_E_CAP);
    const u32 nww = status->nwork[CLASS_WAVE];
    for (u32 i = first + wv; i < nww; i += stride) {
        const Work wk = wwork[i];
        const DevState d = pma_of(g, wk.pid);
        rebalance_window_wave(d, wk, ckey, pd, A, B, P, bs);
    }
    // per wave (lane 0 holds its wave's sums), the window counts per workgroup
    u64 *row = g.stats + (u64)(blockIdx.x % ST_ROWS) * ST_NUM;
    if (lane_id() == 0 && bs.bytes) {
        atomicAdd((unsigned long long *)&row[ST_UPDATES], (unsigned long long)bs.updates);
        atomicAdd((unsigned long long *)&row[ST_MARKERS], (unsigned long long)bs.markers);
        atomicAdd((unsigned long long *)&row[ST_ALG_BYTES_WAVE], (unsigned long long)(bs.bytes + 4ull * bs.markers));
    }
    __syncthreads();
    if (threadIdx.x < 32 && wl[threadIdx.x])
        atomicAdd((unsigned long long *)&row[ST_WINDOWS + threadIdx.x], (unsigned long long)wl[threadIdx.x]);
}

template <int SPL>
__global__ __launch_bounds__(SEG_TPB) void k_rebalance_seg(Group g, const Work *__restrict__ work,
                                                           const Status *__restrict__ status,
                                                           const u64 *__restrict__ ckey, const i32 *__restrict__ pd,
                                                           u32 log2G, u32 nw_host, const Work *__restrict__ wwork,
                                                           u32 sgrid) {
    extern __shared__ u64 lds[];
    __shared__ u64 red[4];
    const u32 G = 1u << log2G;
    const u32 S = G * SPL;
    const u32 lane_g = threadIdx.x & (G - 1);
    const u32 grp = threadIdx.x >> log2G;
    const u32 groups_per_block = SEG_TPB >> log2G;
    u64 *U = (u64 *)((u32 *)lds + (u64)grp * seg_slice_words(S));  // updates
    i32 *P = (i32 *)(U + 2 * S);                                      // their prefix
    // the group's lanes within the wave
    const u32 gshift = (threadIdx.x & 63) & ~(G - 1);
    const u64 gmask = (G == 64 ? ~0ull : ((1ull << G) - 1)) << gshift;
    constexpr u32 WPB = SEG_TPB / 64;
    if (wwork && blockIdx.x >= sgrid) {
        rebalance_waves_of(g, wwork, status, ckey, pd, (unsigned char *)lds, (blockIdx.x - sgrid) * WPB,
                           gridDim.x * WPB);
        return;
    }
    if (threadIdx.x < 4) red[threadIdx.x] = 0;
    // the host passes the window count when it knows it
    const u32 nw = nw_host != ~0u ? nw_host : status->nwork[CLASS_SEG];
    u64 bytes = 0, upd = 0, markers = 0, wins = 0;
    const u32 nblk = wwork ? sgrid : gridDim.x;
    for (u32 w = blockIdx.x * groups_per_block + grp; w < nw; w += nblk * groups_per_block) {
        const Work wk = work[w];
        const DevState d = pma_of(g, wk.pid);
        const u32 C = wk.C;
        const u32 k = wk.u1 - wk.u0;
        const u64 base = (u64)wk.a << g.log2S;
        gu64 *keys = (gu64 *)d.keys + base;
        // only the lines that hold live keys are read; the slots past C read
        // as the largest key, which never sorts before an update
        u64 v[SPL];
#pragma unroll
        for (int r = 0; r < SPL; ++r) {
            u64 x = KEY_NONE;
            if (r * G + lane_g < C) x = keys[r * G + lane_g];
            v[r] = x;
        }
        const i32 pbase = pd[wk.u0];
        const u64 uk = lane_g < k ? ckey[wk.u0 + lane_g] : 0ull;
        const i32 up = lane_g <= k ? pd[wk.u0 + lane_g] : 0;
        if (lane_g < k) U[lane_g] = uk;
        if (lane_g <= k) P[lane_g] = up - pbase;
        for (u32 q = lane_g + G; q < k; q += G) U[q] = ckey[wk.u0 + q];
        for (u32 q = lane_g + G; q <= k; q += G) P[q] = pd[wk.u0 + q] - pbase;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

        i32 shift[SPL];
        u32 gone = 0;  // bit r: old key r is deleted
#pragma unroll
        for (int r = 0; r < SPL; ++r) shift[r] = 0;
        i32 pq = 0;
        for (u32 q = 0; q < k; ++q) {
            const u64 y = U[q];
            const i32 pn = P[q + 1];
            const i32 sgn = pn - pq;  // +1 insert, -1 delete
            u32 below = 0;            // old keys of the window before y
#pragma unroll
            for (int r = 0; r < SPL; ++r) {
                shift[r] += y < v[r] ? sgn : 0;
                gone |= y == v[r] ? 1u << r : 0u;  // only deletes meet a present key
                below += __popcll(__ballot(v[r] < y) & gmask);
            }
            if (sgn == 1 && lane_g == (q & (G - 1))) {
                const u32 o = below + (u32)pq;
                if (o < S) keys[o] = y;  // inserts are never markers
            }
            pq = pn;
        }
        const u32 M = (u32)((i32)C + pq);
#pragma unroll
        for (int r = 0; r < SPL; ++r) {
            const u32 t = r * G + lane_g;
            if (t >= C || ((gone >> r) & 1u)) continue;
            const u32 o = (u32)((i32)t + shift[r]);
            // a key before every update keeps its slot: nothing to write
            if (o != t && o < S) {
                const u64 x = v[r];
                keys[o] = x;
                if (is_marker(x) && (u32)(x >> 32) < d.V) {
                    d.row_off[(u32)(x >> 32) + 1] = (u32)(base + o);
                    ++markers;
                }
            }
        }
        // slots emptied by deletes
        for (u32 t = M + lane_g; t < C; t += G) keys[t] = KEY_NONE;
        if (lane_g == 0) {
            d.cnt[wk.a] = M;
            bytes += 16ull * S + 8ull * k;
            upd += k;
            wins += 1;
        }
        __builtin_amdgcn_wave_barrier();
    }
    // statistics: one set of atomics per workgroup
    markers = wave_sum(markers);
    bytes = wave_sum(bytes);
    upd = wave_sum(upd);
    wins = wave_sum(wins);
    __syncthreads();
    if (lane_id() == 0 && wins) {
        atomicAdd((unsigned long long *)&red[0], (unsigned long long)wins);
        atomicAdd((unsigned long long *)&red[1], (unsigned long long)upd);
        atomicAdd((unsigned long long *)&red[2], (unsigned long long)markers);
        atomicAdd((unsigned long long *)&red[3], (unsigned long long)(bytes + 4ull * markers));
    }
    __syncthreads();
    u64 *row = g.stats + (u64)(blockIdx.x % ST_ROWS) * ST_NUM;
    if (threadIdx.x == 0 && red[0]) {
        atomicAdd((unsigned long long *)&row[ST_WINDOWS + 0], (unsigned long long)red[0]);
        atomicAdd((unsigned long long *)&row[ST_UPDATES], (unsigned long long)red[1]);
        atomicAdd((unsigned long long *)&row[ST_MARKERS], (unsigned long long)red[2]);
        atomicAdd((unsigned long long *)&row[ST_ALG_BYTES_SEG], (unsigned long long)red[3]);
    }
    if (wwork) {
        __syncthreads();  // every group is done with its LDS slice
        rebalance_waves_of(g, wwork, status, ckey, pd, (unsigned char *)lds, (gridDim.x - sgrid + blockIdx.x) * WPB,
                           gridDim.x * WPB);
    }
}

// lanes per window (log2) of the single-segment rebalance: `slots_log2`
// slots per lane, by default 4 (S/64 from 256-slot segments up), at most a
// wave, at most 16 slots per lane, and within the LDS of a workgroup
u32 seg_lanes_log2(u32 log2S, int slots_log2) {
    int g = slots_log2 < 0 ? (log2S >= 8 ? 6 : (int)log2S - 2) : (int)log2S - slots_log2;
    if (g < (int)log2S - 4) g = (int)log2S - 4;
    if (g < 0) g = 0;
    if (g > 6) g = 6;
    auto lds_bytes = [&](int gg) { return (size_t)(SEG_TPB >> gg) * seg_slice_words(1u << log2S) * 4; };
    while (g < 6 && lds_bytes(g) > 160u * 1024) ++g;
    return (u32)g;
}

// one single-segment launch; with events, the dispatch itself records its
// start and end in them (the kernel's execution time, no separate packets)
template <int SPL>
static void seg_dispatch(u32 grid, u32 wgrid, size_t lds, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                         const Group &g, const BatchBufs &b, u32 log2G, u32 nw, const Work *wwork) {
    if (e0)
        hipExtLaunchKernelGGL(k_rebalance_seg<SPL>, dim3(grid + wgrid), dim3(SEG_TPB), (u32)lds, st, e0, e1, 0u, g,
                              (const Work *)b.work[CLASS_SEG], (const Status *)b.status, (const u64 *)b.ckey,
                              (const i32 *)b.pd, log2G, nw, wwork, grid);
    else
        k_rebalance_seg<SPL><<<grid + wgrid, SEG_TPB, lds, st>>>(g, b.work[CLASS_SEG], b.status, b.ckey, b.pd, log2G,
                                                                 nw, wwork, grid);
}

// `windows`: the exact window count (eager rounds) or a bound that sizes the
// grid, the kernel then reading the count from the status (graph replays)
static hipError_t rebalance_seg_launch(const Group &g, const BatchBufs &b, u64 windows, bool exact, hipStream_t st,
                                       hipEvent_t e0, hipEvent_t e1, bool with_wave = false) {
    u32 S = 1u << g.log2S;
    u32 log2G = g.seg_log2G;
    u32 groups_per_block = SEG_TPB >> log2G;
    size_t lds = (size_t)groups_per_block * seg_slice_words(S) * sizeof(u32);
    if (with_wave) lds = std::max<size_t>(lds, (size_t)(SEG_TPB / 64) * WAVE_LDS_BYTES);
    const Work *wwork = with_wave ? b.work[CLASS_WAVE] : nullptr;
    // one-wave windows: up to EARLY_WAVE_GRID workgroups of their own (their
    // count is on the device), at most a wave per 64 batch keys
    const u32 wgrid = with_wave ? (u32)std::max<u64>(1, std::min<u64>(EARLY_WAVE_GRID, (windows + 255) / 256)) : 0u;
    u64 blocks = (windows + groups_per_block - 1) / groups_per_block;
    u32 grid = (u32)(blocks < 1 ? 1 : (blocks > SEG_GRID_MAX ? SEG_GRID_MAX : blocks));
    u32 nw = exact ? (u32)windows : ~0u;
    switch (S >> log2G) {
        case 1: seg_dispatch<1>(grid, wgrid, lds, st, e0, e1, g, b, log2G, nw, wwork); break;
        case 2: seg_dispatch<2>(grid, wgrid, lds, st, e0, e1, g, b, log2G, nw, wwork); break;
        case 4: seg_dispatch<4>(grid, wgrid, lds, st, e0, e1, g, b, log2G, nw, wwork); break;
        case 8: seg_dispatch<8>(grid, wgrid, lds, st, e0, e1, g, b, log2G, nw, wwork); break;
        case 16: seg_dispatch<16>(grid, wgrid, lds, st, e0, e1, g, b, log2G, nw, wwork); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t rebalance_seg(const Group &g, const BatchBufs &b, u64 windows, hipStream_t st, hipEvent_t e0,
                         hipEvent_t e1) {
    return rebalance_seg_launch(g, b, windows, true, st, e0, e1);
}

hipError_t rebalance_seg_wave_bound(const Group &g, const BatchBufs &b, u64 bound, hipStream_t st, hipEvent_t e0,
                                    hipEvent_t e1) {
    return rebalance_seg_launch(g, b, bound, false, st, e0, e1, true);
}

hipError_t rebalance_lds(const Group &g, const BatchBufs &b, int cls, u32 grid, hipStream_t st, hipEvent_t e0,
                         hipEvent_t e1) {
    const Work *w = b.work[cls == CLASS_WAVE ? CLASS_WAVE : CLASS_BLOCK];
    if (cls == CLASS_WAVE) {
        if (e0)
            hipExtLaunchKernelGGL(k_rebalance_wave, dim3(grid), dim3(64), 0u, st, e0, e1, 0u, g, w,
                                  (const Status *)b.status, (const u64 *)b.ckey, (const i32 *)b.pd);
        else
            k_rebalance_wave<<<grid, 64, 0, st>>>(g, w, b.status, b.ckey, b.pd);
    } else {
        if (e0)
            hipExtLaunchKernelGGL(k_rebalance_block, dim3(grid), dim3(256), 0u, st, e0, e1, 0u, g, w,
                                  (const Status *)b.status, (const u64 *)b.ckey, (const i32 *)b.pd);
        else
            k_rebalance_block<<<grid, 256, 0, st>>>(g, w, b.status, b.ckey, b.pd);
    }
    return hipGetLastError();
}

// Every rebalance class with grids fixed in advance (graph replays): the
// kernels loop over the work counts the election left in the status, and
// runs_cap bounds them.  merged must hold the live keys of every PMA of
// the group.
hipError_t rebalance_fixed(const Group &g, const BatchBufs &b, u32 runs_cap, u64 *merged, hipStream_t st) {
    (void)rebalance_seg_launch(g, b, runs_cap < 65536u ? runs_cap : 65536u, false, st, nullptr, nullptr);
    const u32 gw = runs_cap < 4096 ? (runs_cap ? runs_cap : 1) : 4096;
    k_rebalance_wave<<<gw, 64, 0, st>>>(g, b.work[CLASS_WAVE], b.status, b.ckey, b.pd);
    const u32 gb = runs_cap < 1024 ? (runs_cap ? runs_cap : 1) : 1024;
    k_rebalance_block<<<gb, 256, 0, st>>>(g, b.work[CLASS_BLOCK], b.status, b.ckey, b.pd);
    return rebalance_large(g, b, merged, 1024, st);
}

// ---------------------------------------------------- global window path
//
// Spread of merged keys (a bulk load's sorted keys, or a window's merged
// scratch): every destination slot reads the key the even spread assigns to
// it (k_pull), and every segment its count and separator (k_meta).

__global__ void k_pull(DevState t, u64 a, u32 L, u64 m, const u64 *__restrict__ merged, u64 *stats, int cls_slot) {
    const u64 W = (1ull << L) << t.log2S;
    const u64 base = a << t.log2S;
    const u32 Smask = (1u << t.log2S) - 1;
    u32 markers = 0;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < W; i += (u64)gridDim.x * blockDim.x) {
        u64 seg = i >> t.log2S;
        u32 lane = (u32)i & Smask;
        u64 f = spread_first(seg, m, L);
        u64 g = spread_first(seg + 1, m, L);
        u64 key = (lane < g - f) ? merged[f + lane] : KEY_NONE;
        t.keys[base + i] = key;
        if (is_marker(key) && (u32)(key >> 32) < t.V) {
            t.row_off[(u32)(key >> 32) + 1] = (u32)(base + i);
            ++markers;
        }
    }
    markers = wave_sum(markers);
    if (lane_id() == 0 && markers && stats) {
        u64 *row = stats + (u64)(blockIdx.x % ST_ROWS) * ST_NUM;
        atomicAdd((unsigned long long *)&row[ST_MARKERS], (unsigned long long)markers);
        atomicAdd((unsigned long long *)&row[cls_slot], (unsigned long long)(4ull * markers));
    }
}

__global__ void k_meta(DevState t, u64 a, u32 L, u64 m, const u64 *__restrict__ merged) {
    const u64 nsegw = 1ull << L;
    const u64 sep_a = (a == 0) ? 0 : t.sep[a];
    for (u64 x = (u64)blockIdx.x * blockDim.x + threadIdx.x; x < nsegw; x += (u64)gridDim.x * blockDim.x) {
        u64 f = spread_first(x, m, L);
        u64 g = spread_first(x + 1, m, L);
        t.cnt[a + x] = (u32)(g - f);
        if (x == 0) {
            if (a == 0) t.sep[0] = 0;
        } else {
            t.sep[a + x] = (g > f) ? merged[f] : (f > 0 ? merged[f - 1] + 1 : sep_a);
        }
    }
}

// ---------------------------------------------- batched large windows
//
// All windows beyond the LDS classes of one batch are handled by three
// launches, whatever their number and size.  A window of 2^L segments is cut
// into chunks of 2^cl segments (4096 slots); chunk j's rank offset inside
// its window is read off the count pyramid by a walk over the levels (the
// counts of the aligned blocks left of it), so no scan is needed.  Each
// window gets a private slice of the merged scratch (windows are disjoint,
// so the slices add up to at most the live key count).

__host__ __device__ __forceinline__ u32 chunk_level(u32 log2S, u32 L) {
    u32 cl = log2S >= LARGE_CHUNK_LOG2_SLOTS ? 0 : LARGE_CHUNK_LOG2_SLOTS - log2S;
    return L < cl ? L : cl;
}

// live keys of the window's segments [a, x), x - a a multiple of 2^cl
__device__ __forceinline__ u64 rank_in_window(const DevState &d, u64 a, u32 L, u64 x, u32 cl) {
    u64 rel = x - a, pos = a, off = 0;
    for (int l = (int)L - 1; l >= (int)cl; --l) {
        if ((rel >> l) & 1ull) {
            off += *count_ptr(d, (u32)l, pos >> l);
            pos += 1ull << l;
        }
    }
    return off;
}

__global__ __launch_bounds__(1024) void k_large_prep(const Group g, const Work *__restrict__ work,
                                                     const Status *__restrict__ status, u32 *__restrict__ chunk_base,
                                                     u64 *__restrict__ mbase, u32 *__restrict__ hdr) {
    __shared__ u32 tmp[1024 / 64 + 1];
    __shared__ u64 tmp64[1024 / 64 + 1];
    const u32 n = status->nwork[CLASS_LARGE];
    u32 carry = 0;
    u64 mcarry = 0;
    for (u32 b = 0; b < n; b += 1024) {
        u32 i = b + threadIdx.x;
        u32 nch = 0;
        u64 m = 0;
        if (i < n) {
            Work w = work[i];
            nch = 1u << (w.L - chunk_level(g.log2S, w.L));
            m = w.M;
        }
        u32 tot;
        u64 mtot;
        u32 ex = block_excl_scan<1024, u32>(nch, tot, tmp);
        u64 mex = block_excl_scan<1024, u64>(m, mtot, tmp64);
        if (i < n) {
            chunk_base[i] = carry + ex;
            mbase[i] = mcarry + mex;
        }
        carry += tot;
        mcarry += mtot;
    }
    if (threadIdx.x == 0) {
        chunk_base[n] = carry;
        hdr[0] = n;
        hdr[1] = carry;
    }
}

__device__ __forceinline__ u32 window_of_chunk(const u32 *chunk_base, u32 n, u32 c) {
    // last i with chunk_base[i] <= c
    u32 lo = 0, hi = n;
    while (hi - lo > 1) {
        u32 mid = (lo + hi) >> 1;
        if (chunk_base[mid] <= c) lo = mid; else hi = mid;
    }
    return lo;
}

static constexpr int LG_TPB = 256;     // pull: threads per block
static constexpr int LGP_TPB = 256;    // push: threads per block
static constexpr u32 LG_SLOTS = 1u << LARGE_CHUNK_LOG2_SLOTS;
static constexpr u32 LG_SEG_MAX = LG_SLOTS / 2;  // segments per chunk at S = 2
static constexpr int LGP_PER = (int)((LG_SLOTS + 2 + LGP_TPB - 1) / LGP_TPB);  // odd: 9
static_assert(LGP_PER & 1, "an odd run length keeps the scan's LDS accesses off shared banks");

// In-place exclusive scan of arr[0, n) (arr[n] receives the total), n <
// NT * PER: each thread sums PER consecutive entries, one block scan joins
// the runs.  Every thread of the block must call it.
template <int NT, int PER>
__device__ __forceinline__ void block_scan_runs(u32 *arr, u32 n, u32 *tmp) {
    const u32 b = threadIdx.x * PER;
    u32 v[PER];
    u32 s = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        v[k] = b + k < n ? arr[b + k] : 0u;
        s += v[k];
    }
    u32 tot;
    u32 ex = block_excl_scan<NT, u32>(s, tot, tmp);  // its barrier orders the reads before the writes
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        if (b + k < n) arr[b + k] = ex;
        ex += v[k];
    }
    if (threadIdx.x == 0) arr[n] = tot;
    __syncthreads();
}

// First index in [lo, hi) with a[idx] >= x (hi if none), found by one whole
// wave: each step tests 64 evenly spaced pivots at once, so the chain of
// dependent loads is log64 of the range, not log2.
__device__ __forceinline__ u32 wave_lower_bound(const u64 *__restrict__ a, u32 lo, u32 hi, u64 x) {
    const u32 lane = lane_id();
    while (lo < hi) {
        const u32 step = (hi - lo + 63) / 64;
        const u32 p = lo + lane * step;
        const u64 below = __ballot(p < hi && a[p] < x);  // a prefix of the lanes: a is sorted
        const u32 c = (u32)__popcll(below);
        if (c == 0) return lo;
        // a[lo + (c-1)*step] < x, and a[lo + c*step] >= x unless that is past hi
        const u32 nlo = lo + (c - 1) * step + 1;
        hi = min(lo + c * step, hi);
        lo = nlo;
    }
    return lo;
}

// Push of one chunk (2^cl segments, at most 4096 slots) of window (a, L):
// every key that survives the batch and every insert of the chunk gets its
// rank among the window's keys after the batch, and out(rank, key) places
// it.  The chunk's slots are staged in LDS; each update of the chunk finds
// its rank among the chunk's live keys there and records what it does to
// the keys after it (+1 at its rank for an insert, -1 after the key it
// deletes, which is flagged).  One scan over the ranks then gives every
// live key its shift, so each key costs a coalesced load and no search.
struct ChunkLds {
    u64 ks[LG_SLOTS];           // the chunk's slots
    u32 pre[LG_SEG_MAX + 1];    // live keys per segment, then their exclusive scan
    u32 sh[LG_SLOTS + 2];       // per rank: net updates that first move it, then their scan
    u32 del[LG_SLOTS / 32 + 1]; // ranks whose key is deleted (rank <= LG_SLOTS)
    u32 stmp[LGP_TPB / 64 + 1];
    u32 qr[2];
    u64 ob;
};

// A chunk's update range and window rank when computed beforehand
// (k_chunk_index): q[c], q[c + 1] and ob[c]; otherwise (q null) the chunk
// finds them itself.
struct ChunkPre {
    const u32 *q;
    const u64 *ob;
    u64 c;
};

// What ranking a chunk leaves: its live keys' exclusive prefix per segment
// (S.pre), their shifts (S.sh) and deletions (S.del) in LDS, and these.
struct ChunkRanks {
    u64 obase;  // the chunk's first live key's rank among the window's old keys
    i64 cbase;  // net updates of the window before the chunk
    u64 r0, r1; // the chunk's rank range in the window after the batch
    u32 nsc, W; // segments and slots of the chunk
};

// Phases of a chunk push up to the ranks: loads (the chunk's slots into kv
// and LDS, its counts, its update range and window rank), the counts' scan,
// the updates (on_insert(rank, key) for each insert), and the shifts' scan.
template <class OnInsert>
__device__ __forceinline__ ChunkRanks chunk_rank(const DevState &d, u64 a, u32 L, u32 wu0, u32 wu1, u64 pb, u64 s0,
                                                 u32 cl, const u64 *__restrict__ ckey, const u64 *__restrict__ cseg,
                                                 const i32 *__restrict__ pd, ChunkLds &S, ChunkPre pre,
                                                 u64 (&kv)[LG_SLOTS / LGP_TPB], OnInsert on_insert) {
    const u32 log2S = d.log2S;
    const u32 wv = threadIdx.x >> 6;
    ChunkRanks cr;
    cr.nsc = 1u << cl;
    cr.W = cr.nsc << log2S;  // <= LG_SLOTS (S <= 1024)
    const u32 nsc = cr.nsc, W = cr.W;
    const u64 slot0 = s0 << log2S;
    // the loads of this phase are independent of each other
#pragma unroll
    for (u32 k = 0; k < LG_SLOTS / LGP_TPB; ++k) {
        const u32 t = threadIdx.x + k * LGP_TPB;
        kv[k] = t < W ? d.keys[slot0 + t] : KEY_NONE;
    }
    for (u32 x = threadIdx.x; x < nsc; x += LGP_TPB) S.pre[x] = d.cnt[s0 + x];
    if (pre.q) {
        if (threadIdx.x < 2) S.qr[threadIdx.x] = pre.q[pre.c + threadIdx.x];
        if (threadIdx.x == 2) S.ob = pre.ob[pre.c];
    } else if (wv < 2) {
        // the chunk's updates: [first >= its first segment, first >= the next chunk's)
        const u32 q = wave_lower_bound(cseg, wu0, wu1, pb + s0 + (wv ? nsc : 0u));
        if (lane_id() == 0) S.qr[wv] = q;
    } else if (wv == 2) {
        // rank of the chunk's first key in its window: the live counts of
        // the aligned blocks left of it, one level per lane
        const u64 rel = s0 - a;
        const u32 l = cl + lane_id();
        u64 v = 0;
        if (l < L && ((rel >> l) & 1ull)) v = *count_ptr(d, l, (a + ((rel >> (l + 1)) << (l + 1))) >> l);
        v = wave_sum(v);
        if (lane_id() == 0) S.ob = v;
    }
    for (u32 t = threadIdx.x; t < W + 2; t += LGP_TPB) S.sh[t] = 0;
    for (u32 t = threadIdx.x; t < LG_SLOTS / 32 + 1; t += LGP_TPB) S.del[t] = 0;
#pragma unroll
    for (u32 k = 0; k < LG_SLOTS / LGP_TPB; ++k) {
        const u32 t = threadIdx.x + k * LGP_TPB;
        if (t < W) S.ks[t] = kv[k];
    }
    __syncthreads();
    block_scan_runs<LGP_TPB, LGP_PER>(S.pre, nsc, S.stmp);  // pre[nsc] = live keys of the chunk
    const u32 q0 = S.qr[0], q1 = S.qr[1];
    cr.obase = S.ob;
    const i32 pbase = pd[wu0];
    for (u32 q = q0 + threadIdx.x; q < q1; q += LGP_TPB) {
        const u64 y = ckey[q];
        const u32 ls = (u32)cseg[q] - (u32)s0;
        const i32 dl = pd[q + 1] - pd[q];  // +1 insert (y absent), -1 delete (y present)
        const u32 p0 = S.pre[ls];
        const u32 r = p0 + (u32)lower_bound_dev<u64>(S.ks + (ls << log2S), 0, S.pre[ls + 1] - p0, y);
        if (dl > 0) {
            atomicAdd(&S.sh[r], 1u);
            on_insert(cr.obase + r + (u64)(i64)(pd[q] - pbase), y);
        } else {
            atomicAdd(&S.sh[r + 1], 0xFFFFFFFFu);
            atomicOr(&S.del[r >> 5], 1u << (r & 31));
        }
    }
    __syncthreads();
    block_scan_runs<LGP_TPB, LGP_PER>(S.sh, S.pre[nsc] + 1, S.stmp);  // sh[r + 1]: net updates below rank r's key
    cr.cbase = (i64)(pd[q0] - pbase);
    cr.r0 = cr.obase + (u64)cr.cbase;
    cr.r1 = cr.r0 + S.pre[nsc] + (u64)(i64)(pd[q1] - pd[q0]);
    return cr;
}

// rank after the batch of the old key in chunk slot t, or ~0 (no live key
// there, or deleted)
__device__ __forceinline__ u64 chunk_key_rank(const ChunkLds &S, const ChunkRanks &cr, u32 t, u32 log2S) {
    const u32 seg = t >> log2S, lane = t & ((1u << log2S) - 1);
    const u32 p0 = S.pre[seg];
    if (lane >= S.pre[seg + 1] - p0) return ~0ull;
    const u32 r = p0 + lane;
    if ((S.del[r >> 5] >> (r & 31)) & 1u) return ~0ull;
    return cr.obase + r + (u64)(cr.cbase + (i64)(i32)S.sh[r + 1]);
}

// Push of one chunk (2^cl segments, at most 4096 slots) of window (a, L):
// every key that survives the batch and every insert of the chunk gets its
// rank among the window's keys after the batch, and out(rank, key) places
// it.  The chunk's slots are staged in LDS; each update of the chunk finds
// its rank among the chunk's live keys there and records what it does to
// the keys after it (+1 at its rank for an insert, -1 after the key it
// deletes, which is flagged).  One scan over the ranks then gives every
// live key its shift, so each key costs a coalesced load and no search.
template <class Out>
__device__ __forceinline__ void push_chunk(const DevState &d, u64 a, u32 L, u32 wu0, u32 wu1, u64 pb, u64 s0, u32 cl,
                                           const u64 *__restrict__ ckey, const u64 *__restrict__ cseg,
                                           const i32 *__restrict__ pd, ChunkLds &S, ChunkPre pre, Out out) {
    u64 kv[LG_SLOTS / LGP_TPB];
    const ChunkRanks cr = chunk_rank(d, a, L, wu0, wu1, pb, s0, cl, ckey, cseg, pd, S, pre, kv, out);
#pragma unroll
    for (u32 k = 0; k < LG_SLOTS / LGP_TPB; ++k) {
        const u32 t = threadIdx.x + k * LGP_TPB;
        if (t >= cr.W) break;
        const u64 o = chunk_key_rank(S, cr, t, d.log2S);
        if (o != ~0ull) out(o, kv[k]);
    }
    __syncthreads();  // the LDS arrays are reused by the next chunk
}

// Large windows: each chunk pushes into its window's slice of the merged
// scratch (k_large_pull then spreads it).
__global__ __launch_bounds__(LGP_TPB) void k_large_push(Group g, const Work *__restrict__ work,
                                                         const u32 *__restrict__ chunk_base,
                                                         const u64 *__restrict__ mbase, const u32 *__restrict__ hdr,
                                                         const u64 *__restrict__ ckey, const u64 *__restrict__ cseg,
                                                         const i32 *__restrict__ pd, u64 *__restrict__ merged) {
    if (g.rh) merged = g.rh->merged;
    __shared__ ChunkLds S;
    const u32 nwin = hdr[0], nch = hdr[1];
    for (u32 c = blockIdx.x; c < nch; c += gridDim.x) {
        const u32 i = window_of_chunk(chunk_base, nwin, c);
        const Work wk = work[i];
        const DevState d = pma_of(g, wk.pid);
        const u32 cl = chunk_level(g.log2S, wk.L);
        const u64 s0 = (u64)wk.a + ((u64)(c - chunk_base[i]) << cl);
        u64 *out = merged + mbase[i];
        const u64 M = wk.M;
        push_chunk(d, wk.a, wk.L, wk.u0, wk.u1, (u64)wk.pid << 32, s0, cl, ckey, cseg, pd, S,
                   ChunkPre{nullptr, nullptr, 0}, [=](u64 o, u64 key) {
                       if (o < M) out[o] = key;
                   });
    }
}

// Root rebuild and resize in one pass: every chunk of the old arrays s
// ranks its keys after the batch (chunk_rank) and writes them to new arrays
// t (2^t.H segments holding M keys, spread evenly).  The chunk's ranks
// [r0, r1) are contiguous, so its merged keys are assembled in LDS in rank
// order (its surviving keys from registers, its inserts from a stash; up to
// RB_MERGED of them, more than a full chunk holds), and
// then the destination segments that hold its ranks are written slot by
// slot, coalesced: a slot of segment x, lane j holds rank
// spread_first(x) + j if j is below the segment's count.  A chunk writes
// the slots of its own ranks, the count and separator of the segments whose
// first rank is its, and the empty tail of the segments whose last rank is
// its (with the segments left empty after them: fewer keys than segments);
// markers also write their row offsets.  A chunk whose keys outgrow the
// LDS (many inserts) places each key by the even spread's inverse instead.
static constexpr u32 RB_STASH = 256;  // inserts a chunk keeps in LDS
// merged keys a chunk assembles in LDS: the room of its slots, counts and
// shifts (ChunkLds up to del), read by then
static constexpr u32 RB_MERGED = (u32)(offsetof(ChunkLds, del) / sizeof(u64));
static_assert(offsetof(ChunkLds, ks) == 0, "the merged keys start at the chunk's slots");
struct RebuildLds {
    u64 iy[RB_STASH];
    u64 io[RB_STASH];
    u32 n;
};

// segment of the even spread of M keys over 2^H segments that holds rank o
// (o < M): the floating estimate of floor(o * 2^H / M) is within one of it
__device__ __forceinline__ u64 spread_segment(u64 o, u64 M, u32 H, double seg_per_key, u64 &f, u64 &g) {
    const u64 nseg = 1ull << H;
    u64 x = (u64)((double)o * seg_per_key);
    if (x >= nseg) x = nseg - 1;
    f = spread_first(x, M, H);
    while (f > o) f = spread_first(--x, M, H);
    g = spread_first(x + 1, M, H);
    while (g <= o) {
        f = g;
        g = spread_first(++x + 1, M, H);
    }
    return x;
}

__global__ __launch_bounds__(LGP_TPB) __attribute__((amdgpu_waves_per_eu(4))) void k_rebuild_scatter(DevState s, DevState t, u64 pb, u32 u0, u32 u1, u64 M,
                                                              double seg_per_key, const u64 *__restrict__ ckey,
                                                              const u64 *__restrict__ cseg,
                                                              const i32 *__restrict__ pd, const u32 *__restrict__ cq,
                                                              const u64 *__restrict__ cob) {
    __shared__ ChunkLds S;
    __shared__ RebuildLds R;
    const u32 cl = chunk_level(s.log2S, s.H);
    const u64 nch = nseg_of(s) >> cl;
    const u32 Ht = t.H, log2S = t.log2S;
    const u64 nseg_t = nseg_of(t);
    const u32 SL = 1u << log2S;
    const u64 Nm1 = nseg_t - 1;
    // one key by the even spread's inverse, with the segments left empty
    // after a segment's last key
    auto place = [=](u64 o, u64 key) {
        if (o >= M) return;
        u64 f, g;
        const u64 x = spread_segment(o, M, Ht, seg_per_key, f, g);
        const u32 lane = (u32)(o - f);
        const u64 base = x << log2S;
        t.keys[base + lane] = key;
        if (is_marker(key) && (u32)(key >> 32) < t.V) t.row_off[(u32)(key >> 32) + 1] = (u32)(base + lane);
        if (lane == 0) {
            t.cnt[x] = (u32)(g - f);
            t.sep[x] = x == 0 ? 0ull : key;
        }
        if (o + 1 == g) {
            for (u64 y = x + 1; y < nseg_t && spread_first(y + 1, M, Ht) == g; ++y) {
                t.cnt[y] = 0;
                t.sep[y] = key + 1;
            }
        }
    };
    for (u64 c = blockIdx.x; c < nch; c += gridDim.x) {
        if (threadIdx.x == 0) R.n = 0;
        __syncthreads();
        u64 kv[LG_SLOTS / LGP_TPB];
        const ChunkRanks cr = chunk_rank(s, 0, s.H, u0, u1, pb, c << cl, cl, ckey, cseg, pd, S,
                                         ChunkPre{cq, cob, c}, kv, [&](u64 o, u64 y) {
                                             const u32 i = atomicAdd(&R.n, 1u);
                                             if (i < RB_STASH) {
                                                 R.io[i] = o;
                                                 R.iy[i] = y;
                                             } else {
                                                 place(o, y);
                                             }
                                         });
        const u64 r0 = cr.r0, r1 = cr.r1;
        const bool in_lds = r1 - r0 <= RB_MERGED && R.n <= RB_STASH;
        if (!in_lds) {
            // keys by the inverse, then the empty tails of the segments whose
            // last key the chunk placed
            for (u32 i = threadIdx.x; i < min(R.n, RB_STASH); i += LGP_TPB) place(R.io[i], R.iy[i]);
#pragma unroll
            for (u32 k = 0; k < LG_SLOTS / LGP_TPB; ++k) {
                const u32 ts = threadIdx.x + k * LGP_TPB;
                if (ts >= cr.W) break;
                const u64 o = chunk_key_rank(S, cr, ts, s.log2S);
                if (o != ~0ull) place(o, kv[k]);
            }
            if (r0 < M && r1 > r0) {
                u64 f, g;
                const u64 xa = spread_segment(r0, M, Ht, seg_per_key, f, g);
                const u64 xb = r1 < M ? spread_segment(r1, M, Ht, seg_per_key, f, g) : nseg_t;
                const u64 nslots = xb > xa ? (xb - xa) << log2S : 0;
                for (u64 i = threadIdx.x; i < nslots; i += LGP_TPB) {
                    const u64 x = xa + (i >> log2S);
                    const u32 lane = (u32)i & (SL - 1);
                    const u64 tt = x * M + Nm1;
                    if (lane >= ((tt + M) >> Ht) - (tt >> Ht)) t.keys[(xa << log2S) + i] = KEY_NONE;
                }
            }
            __syncthreads();
            continue;
        }
        // merged keys of the chunk in rank order, into the chunk's LDS (its
        // slots, counts and shifts: every thread first takes its keys' ranks,
        // then the barrier, then the stores)
        u32 ov[LG_SLOTS / LGP_TPB];
#pragma unroll
        for (u32 k = 0; k < LG_SLOTS / LGP_TPB; ++k) {
            const u32 ts = threadIdx.x + k * LGP_TPB;
            const u64 o = ts < cr.W ? chunk_key_rank(S, cr, ts, s.log2S) : ~0ull;
            ov[k] = o != ~0ull ? (u32)(o - r0) : ~0u;
        }
        __syncthreads();
        u64 *mk = reinterpret_cast<u64 *>(&S);
#pragma unroll
        for (u32 k = 0; k < LG_SLOTS / LGP_TPB; ++k)
            if (ov[k] != ~0u) mk[ov[k]] = kv[k];
        for (u32 i = threadIdx.x; i < R.n; i += LGP_TPB) mk[R.io[i] - r0] = R.iy[i];
        __syncthreads();
        if (r1 > r0) {
            u64 f, g;
            const u64 x0 = spread_segment(r0, M, Ht, seg_per_key, f, g);
            const u64 xk = spread_segment(r1 - 1, M, Ht, seg_per_key, f, g);
            const u64 xe = r1 < M ? spread_segment(r1, M, Ht, seg_per_key, f, g) : nseg_t;
            const u64 x1 = xe > xk + 1 ? xe : xk + 1;
            const u64 nslots = (x1 - x0) << log2S;
            for (u64 i = threadIdx.x; i < nslots; i += LGP_TPB) {
                const u64 x = x0 + (i >> log2S);
                const u32 lane = (u32)i & (SL - 1);
                const u64 tt = x * M + Nm1;
                const u64 fx = tt >> Ht, gx = (tt + M) >> Ht;  // spread_first(x), spread_first(x + 1)
                const u64 slot = (x0 << log2S) + i;
                if (lane < gx - fx) {
                    const u64 o = fx + lane;
                    if (o >= r0 && o < r1) {
                        const u64 key = mk[o - r0];
                        t.keys[slot] = key;
                        if (is_marker(key) && (u32)(key >> 32) < t.V) t.row_off[(u32)(key >> 32) + 1] = (u32)slot;
                        if (lane == 0) {
                            t.cnt[x] = (u32)(gx - fx);
                            t.sep[x] = x == 0 ? 0ull : key;
                        }
                    }
                } else {
                    // the tail belongs to the chunk of the segment's last
                    // rank, an empty segment's to the chunk of the rank before
                    const u64 last = gx - 1;  // gx >= 1: fx == gx == 0 is impossible past segment 0
                    if (last >= r0 && last < r1) {
                        t.keys[slot] = KEY_NONE;
                        if (lane == 0 && gx == fx) {
                            t.cnt[x] = 0;
                            t.sep[x] = mk[last - r0] + 1;
                        }
                    }
                }
            }
        }
        __syncthreads();  // the LDS arrays are reused by the next chunk
    }
}

// For every chunk of a whole-PMA rebuild, one wave: the chunk's first
// update (a 64-pivot search) and its rank in the PMA (one pyramid count per
// level, one level per lane); q[nch] = u1.
__global__ __launch_bounds__(256) void k_chunk_index(DevState s, u64 pb, u32 u0, u32 u1, const u64 *__restrict__ cseg,
                                                     u32 *__restrict__ q, u64 *__restrict__ ob) {
    const u32 cl = chunk_level(s.log2S, s.H);
    const u64 nch = nseg_of(s) >> cl;
    const u64 c = ((u64)blockIdx.x * 256 + threadIdx.x) >> 6;
    if (c > nch) return;  // whole waves
    if (c == nch) {
        if (lane_id() == 0) q[nch] = u1;
        return;
    }
    const u64 s0 = c << cl;
    const u32 qc = wave_lower_bound(cseg, u0, u1, pb + s0);
    const u32 l = cl + lane_id();
    u64 v = 0;
    if (l < s.H && ((s0 >> l) & 1ull)) v = *count_ptr(s, l, ((s0 >> (l + 1)) << (l + 1)) >> l);
    v = wave_sum(v);
    if (lane_id() == 0) {
        q[c] = qc;
        ob[c] = v;
    }
}

u64 rebuild_index_words(const DevState &src) {
    const u64 nch = nseg_of(src) >> chunk_level(src.log2S, src.H);
    return 3 * (nch + 1) + 2;
}

hipError_t rebuild_scatter(const DevState &src, const DevState &dst, u32 pid, u32 u0, u32 u1, const BatchBufs &b,
                           u64 m, u32 *index, hipStream_t st) {
    const u64 nch = nseg_of(src) >> chunk_level(src.log2S, src.H);
    u32 *cq = index;                                                 // nch + 1 words
    u64 *cob = (u64 *)(((uintptr_t)(index + nch + 1) + 7) & ~(uintptr_t)7);  // nch u64, 8-byte aligned
    const u64 pb = (u64)pid << 32;
    k_chunk_index<<<(u32)((nch + 1 + 3) / 4), 256, 0, st>>>(src, pb, u0, u1, b.cseg, cq, cob);
    const u32 grid = (u32)std::min<u64>(nch, 1024);  // ~33 KB of LDS per block: four per CU
    k_rebuild_scatter<<<grid, LGP_TPB, 0, st>>>(src, dst, pb, u0, u1, m, (double)nseg_of(dst) / (double)m, b.ckey,
                                                b.cseg, b.pd, cq, cob);
    return hipGetLastError();
}

__global__ __launch_bounds__(LG_TPB) void k_large_pull(Group g, const Work *__restrict__ work,
                                                        const u32 *__restrict__ chunk_base, const u64 *__restrict__ mbase,
                                                        const u32 *__restrict__ hdr, const u64 *__restrict__ merged) {
    if (g.rh) merged = g.rh->merged;
    const u32 nwin = hdr[0], nch = hdr[1];
    const u32 log2S = g.log2S;
    const u32 Smask = (1u << log2S) - 1;
    u64 bytes = 0, markers = 0, upd = 0;
    __shared__ u32 windows[32];
    bool any = false;
    if (threadIdx.x < 32) windows[threadIdx.x] = 0;
    __syncthreads();
    for (u32 c = blockIdx.x; c < nch; c += gridDim.x) {
        const u32 i = window_of_chunk(chunk_base, nwin, c);
        const Work wk = work[i];
        const DevState d = pma_of(g, wk.pid);
        const u32 cl = chunk_level(log2S, wk.L);
        const u64 j = c - chunk_base[i];
        const u64 rel0 = j << cl;
        const u32 nsc = 1u << cl;
        const u64 mb = mbase[i];
        const u64 M = wk.M;
        const u64 slot0 = ((u64)wk.a + rel0) << log2S;
        const u32 W = nsc << log2S;  // <= LG_SLOTS
        // every load of the chunk is issued before the first store
        u64 kv[LG_SLOTS / LG_TPB];
#pragma unroll
        for (u32 k = 0; k < LG_SLOTS / LG_TPB; ++k) {
            const u32 t = threadIdx.x + k * LG_TPB;
            const u64 seg = rel0 + (t >> log2S);
            const u32 lane = t & Smask;
            const u64 f = spread_first(seg, M, wk.L);
            const u64 g = spread_first(seg + 1, M, wk.L);
            kv[k] = (t < W && lane < g - f) ? merged[mb + f + lane] : KEY_NONE;
        }
#pragma unroll
        for (u32 k = 0; k < LG_SLOTS / LG_TPB; ++k) {
            const u32 t = threadIdx.x + k * LG_TPB;
            if (t >= W) break;
            const u64 key = kv[k];
            d.keys[slot0 + t] = key;
            if (is_marker(key) && (u32)(key >> 32) < d.V) {
                d.row_off[(u32)(key >> 32) + 1] = (u32)(slot0 + t);
                ++markers;
            }
        }
        const u64 sep_a = d.sep[wk.a];
        for (u32 x = threadIdx.x; x < nsc; x += LG_TPB) {
            u64 seg = rel0 + x;
            u64 f = spread_first(seg, M, wk.L);
            u64 g = spread_first(seg + 1, M, wk.L);
            d.cnt[wk.a + seg] = (u32)(g - f);
            if (seg > 0) d.sep[wk.a + seg] = (g > f) ? merged[mb + f] : (f > 0 ? merged[mb + f - 1] + 1 : sep_a);
        }
        if (threadIdx.x == 0) {
            bytes += 16ull * W;
            if (j == 0) {
                windows[wk.L & 31] += 1u;
                upd += wk.u1 - wk.u0;
                bytes += 8ull * (wk.u1 - wk.u0);
            }
            any = true;
        }
    }
    markers = wave_sum(markers);
    u64 *stats = g.stats + (u64)(blockIdx.x % ST_ROWS) * ST_NUM;
    if (lane_id() == 0 && markers) {
        atomicAdd((unsigned long long *)&stats[ST_MARKERS], (unsigned long long)markers);
        atomicAdd((unsigned long long *)&stats[ST_ALG_BYTES_LARGE], (unsigned long long)(4ull * markers));
    }
    if (threadIdx.x == 0 && any) {
        atomicAdd((unsigned long long *)&stats[ST_ALG_BYTES_LARGE], (unsigned long long)bytes);
        atomicAdd((unsigned long long *)&stats[ST_UPDATES], (unsigned long long)upd);
        for (int l = 0; l < 32; ++l)
            if (windows[l]) atomicAdd((unsigned long long *)&stats[ST_WINDOWS + l], (unsigned long long)windows[l]);
    }
}

hipError_t rebalance_large(const Group &g, const BatchBufs &b, u64 *merged, u32 grid, hipStream_t st) {
    k_large_prep<<<1, 1024, 0, st>>>(g, b.work[CLASS_LARGE], b.status, b.lchunk, b.lmbase, b.lhdr);
    // push: ~29 KB of LDS per block, four blocks per CU
    k_large_push<<<min(grid, 1024u), LGP_TPB, 0, st>>>(g, b.work[CLASS_LARGE], b.lchunk, b.lmbase, b.lhdr, b.ckey, b.cseg, b.pd,
                                          merged);
    k_large_pull<<<grid, LG_TPB, 0, st>>>(g, b.work[CLASS_LARGE], b.lchunk, b.lmbase, b.lhdr, merged);
    return hipGetLastError();
}

hipError_t window_spread(const DevState &t, u64 a, u32 L, const u64 *merged, u64 m, u64 *stats, int cls_slot,
                         hipStream_t st);

// spread m merged keys over the window (a, L) of t
hipError_t window_spread(const DevState &t, u64 a, u32 L, const u64 *merged, u64 m, u64 *stats, int cls_slot,
                         hipStream_t st) {
    u64 nsegw = 1ull << L;
    u64 W = nsegw << t.log2S;
    k_pull<<<grid_for(W, 256, 16384), 256, 0, st>>>(t, a, L, m, merged, stats, cls_slot);
    k_meta<<<grid_for(nsegw, 256, 16384), 256, 0, st>>>(t, a, L, m, merged);
    return hipGetLastError();
}

hipError_t rebuild_spread(const DevState &dst, const u64 *merged, u64 m, hipStream_t st) {
    return window_spread(dst, 0, dst.H, merged, m, nullptr, ST_ALG_BYTES_LARGE, st);
}

// ------------------------------------------------------------- pyramid
//
// Window counts for levels 1..H are rebuilt from the segment counts after
// every batch that changed them: one workgroup reduces a chunk of 2^11
// segments through LDS (levels 1..11), a single workgroup finishes the
// levels above.  O(nseg) streaming work, no atomics.

__global__ __launch_bounds__(256) void k_pyr_chunk(DevState d) { pyr_chunk(d, blockIdx.x); }

// the chunks of every PMA of the run that took window rebalances
// the chunks of every PMA of the run that took window rebalances; the
// workgroups stride over the group's chunk list (its length is read from
// the table, so the grid does not depend on the members' heights)
__global__ __launch_bounds__(256) void k_pyr_chunk_group(Group g) {
    const PmaEntry &last = g.tab[g.npma - 1];
    const u32 total = last.pyr_blk0 + last.pyr_nblk;
    for (u32 c = blockIdx.x; c < total; c += gridDim.x) {
        u32 lo = 0, hi = g.npma;
        while (hi - lo > 1) {
            u32 mid = (lo + hi) >> 1;
            if (g.tab[mid].pyr_blk0 <= c) lo = mid; else hi = mid;
        }
        const PmaStat ps = g.pstat[lo];
        if (ps.u1 == ps.u0 || (ps.flags & (FLAG_ROOT | FLAG_RESIZE))) continue;
        const DevState d = pma_of(g, lo);
        if (d.H == 0) continue;
        pyr_chunk(d, c - g.tab[lo].pyr_blk0);
        __syncthreads();
    }
}

__global__ __launch_bounds__(1024) void k_pyr_top(DevState d) { pyr_top_levels(d); }

__global__ __launch_bounds__(1024) void k_batch_begin(Group g, Status *s, InlineSegs in, BatchSeg *segs, u32 *zero,
                                                      u32 nzero, u32 *wfirst) {
    if (blockIdx.x == 0 && threadIdx.x < in.n) segs[threadIdx.x] = in.s[threadIdx.x];
    batch_first_work(g, s, blockIdx.x, zero, nzero, wfirst);
}

// chunk levels of the PMAs of a run (their top levels follow in the next
// batch's first kernels); nblocks = the pyr_blk0 prefix total
hipError_t group_pyramid(const Group &g, u32 nblocks, hipStream_t st) {
    if (nblocks == 0) return hipSuccess;
    k_pyr_chunk_group<<<nblocks < 2048 ? nblocks : 2048, 256, 0, st>>>(g);
    return hipGetLastError();
}

u32 pyramid_blocks(u32 H) { return H == 0 ? 0 : (H <= PYR_CHL ? 1 : 1u << (H - PYR_CHL)); }

// full: also the levels above the chunks (otherwise the next batch's first
// kernels compute them before anything reads them)
hipError_t rebuild_pyramid(const DevState &d, bool full, hipStream_t st) {
    if (d.H == 0) return hipSuccess;
    u64 nseg = nseg_of(d);
    u32 nb = (u32)(nseg < PYR_CH ? 1 : nseg / PYR_CH);
    k_pyr_chunk<<<nb, 256, 0, st>>>(d);
    if (full && d.H > PYR_CHL) k_pyr_top<<<1, 1024, 0, st>>>(d);
    return hipGetLastError();
}

// --------------------------------------------------------------- fills

__global__ void k_markers(u64 *out, u32 V) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < V; i += (u64)gridDim.x * blockDim.x)
        out[i] = marker_key((u32)i);
}

hipError_t make_markers(u64 *out, u32 V, hipStream_t st) {
    k_markers<<<grid_for(V, 256, 8192), 256, 0, st>>>(out, V);
    return hipGetLastError();
}

__global__ void k_fill_u32(u32 *p, u64 n, u32 v) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(u64 *p, u64 n, u64 v) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) p[i] = v;
}

// device-to-device copy as a kernel at HBM rate (16-byte accesses; the
// runtime's copy of a multi-GB array can go through a DMA engine far below
// it and hold up the stream behind)
__global__ void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, u64 n16) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (u64)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

__global__ void k_copy_tail(unsigned char *__restrict__ dst, const unsigned char *__restrict__ src, u64 n) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) dst[i] = src[i];
}

hipError_t copy_device(void *dst, const void *src, u64 bytes, hipStream_t st) {
    if (bytes == 0) return hipSuccess;
    const bool aligned = ((((uintptr_t)dst) | ((uintptr_t)src)) & 15) == 0;
    u64 n16 = aligned ? bytes / 16 : 0;
    if (n16) k_copy16<<<grid_for(n16, 256, 16384), 256, 0, st>>>((uint4 *)dst, (const uint4 *)src, n16);
    u64 done = n16 * 16;
    if (done < bytes)
        k_copy_tail<<<grid_for(bytes - done, 256, 4096), 256, 0, st>>>((unsigned char *)dst + done,
                                                                       (const unsigned char *)src + done, bytes - done);
    return hipGetLastError();
}

hipError_t fill_u32(u32 *p, u64 n, u32 v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_fill_u32<<<grid_for(n, 256, 8192), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t fill_u64(u64 *p, u64 n, u64 v, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_fill_u64<<<grid_for(n, 256, 8192), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

}  // namespace pma
