// begin.hpp -- a batch's first device work (status, update index, the count
// pyramid's top levels), shared by the batch-begin kernel and the sort
// kernel that carries it for a lone PMA.
#pragma once
#include "common.hpp"

namespace pma {

// the pyramid's chunk levels: one workgroup reduces a chunk of 2^PYR_CHL
// segments (levels 1..PYR_CHL); the levels above are summed at the next
// batch's start
static constexpr u32 PYR_CHL = 11;

static constexpr u32 PYR_CH = 1u << PYR_CHL;

// Levels 1..PYR_CHL of one chunk of 2^PYR_CHL segments (all levels of a
// smaller PMA), by one workgroup of 256 threads, through LDS.
__device__ inline void pyr_chunk(const DevState &d, u64 chunk) {
    __shared__ u32 v0[PYR_CH];
    __shared__ u32 v1[PYR_CH / 2];
    const u64 nseg = nseg_of(d);
    const u32 levels = d.H < PYR_CHL ? d.H : PYR_CHL;
    const u32 n = (u32)(nseg < PYR_CH ? nseg : PYR_CH);
    const u64 c0 = chunk * n;
    for (u32 i = threadIdx.x; i < n; i += 256) v0[i] = d.cnt[c0 + i];
    __syncthreads();
    u32 *in = v0, *tmp = v1;
    for (u32 l = 1; l <= levels; ++l) {
        u32 nw = n >> l;
        u32 *out = d.pyr + pyr_off(nseg, l) + (c0 >> l);
        for (u32 x = threadIdx.x; x < nw; x += 256) {
            u32 s = in[2 * x] + in[2 * x + 1];
            tmp[x] = s;
            out[x] = s;
        }
        __syncthreads();
        u32 *sw = in; in = tmp; tmp = sw;
    }
}

// Levels above the chunk levels, by one workgroup.  Up to PYR_TOP_LDS
// chunk-level counts (H <= 23) are loaded once into LDS and every level
// above is summed there (no global round trip per level), laid out as in
// the pyramid (each level right after the one below), so the levels are
// written out together at the end.  Taller pyramids go level by level
// through global memory.
static constexpr u32 PYR_TOP_LDS = 4096;

__device__ inline void pyr_top_levels(const DevState &d) {
    __shared__ u32 v[2 * PYR_TOP_LDS];
    const u64 nseg = nseg_of(d);
    if (d.H <= PYR_CHL) return;
    const u64 nc = nseg >> PYR_CHL;
    if (nc <= PYR_TOP_LDS) {
        const u32 *in = d.pyr + pyr_off(nseg, PYR_CHL);
        for (u32 x = threadIdx.x; x < (u32)nc; x += blockDim.x) v[x] = in[x];
        __syncthreads();
        u32 lo = 0, n = (u32)nc;  // the level below: v[lo, lo + n)
        for (u32 l = PYR_CHL + 1; l <= d.H; ++l) {
            for (u32 x = threadIdx.x; x < n / 2; x += blockDim.x) v[lo + n + x] = v[lo + 2 * x] + v[lo + 2 * x + 1];
            lo += n;
            n /= 2;
            __syncthreads();
        }
        // levels PYR_CHL + 1 .. H: v[nc, 2 nc - 1), contiguous in the pyramid too
        u32 *out = d.pyr + pyr_off(nseg, PYR_CHL + 1);
        for (u32 x = threadIdx.x; x < (u32)nc - 1; x += blockDim.x) out[x] = v[nc + x];
        return;
    }
    for (u32 l = PYR_CHL + 1; l <= d.H; ++l) {
        u64 nw = nseg >> l;
        const u32 *in = d.pyr + pyr_off(nseg, l - 1);
        u32 *out = d.pyr + pyr_off(nseg, l);
        for (u64 x = threadIdx.x; x < nw; x += blockDim.x) out[x] = in[2 * x] + in[2 * x + 1];
        __syncthreads();
    }
}

// One workgroup's share of a batch's first work for PMA pid: its per-batch
// record and claimed-level masks cleared, its pyramid's top levels summed
// (unless top is false: then a later kernel sums them); the workgroup of
// PMA 0 also clears the status, zero[0, nzero) and, for a lone PMA, resets
// the update index wfirst.
__device__ inline void batch_first_work(const Group &g, Status *s, u32 pid, u32 *zero, u32 nzero, u32 *wfirst,
                                        bool top = true) {
    if (pid == 0) {
        for (u32 i = threadIdx.x; i < nzero; i += blockDim.x) zero[i] = 0;
        if (g.npma == 1)
            for (u32 i = threadIdx.x; i <= WFIRST_N; i += blockDim.x) wfirst[i] = 0xFFFFFFFFu;
        if (threadIdx.x == 0) {
            s->flags = 0;
            s->n_eff = 0;
            s->delta = 0;
            for (int c = 0; c < NUM_CLASSES; ++c) s->nwork[c] = 0;
        }
    }
    if (threadIdx.x == 0) {
        PmaStat z;
        z.flags = 0; z.u0 = 0; z.u1 = 0; z.pd0 = 0; z.pd1 = 0;
        g.pstat[pid] = z;
    }
    if (threadIdx.x < LV_REPL) g.lvmask[(u64)pid * LV_REPL + threadIdx.x] = 0;
    if (top) pyr_top_levels(pma_of(g, pid));
}

}  // namespace pma
