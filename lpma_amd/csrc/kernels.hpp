// kernels.hpp -- host-side launchers of the HIP kernels (one per phase).
#pragma once
#include "common.hpp"

namespace pma {

// sort.hip
u64 scan_scratch_words(u64 n);
hipError_t device_excl_scan(const u32 *in, u64 n, u32 *out, u32 *scratch, hipStream_t st);
u64 sort_scratch_words(u64 n);
u32 sort_zero_words();
// k0, k1, k2: n keys each (k2 only for batches of up to 2^17 keys, as the
// second buffer of an oversized bucket)
u64 *pack_and_sort(const BatchSeg *segs, u32 nsegs, u64 n, const u64 *n_dev, u32 pshift, u32 sbits, u32 lo_bit,
                   u32 hi_bit, u64 *k0, u64 *k1, u64 *k2, u32 *scratch, u32 *flags, hipStream_t st);

// update.hip
struct BatchBufs {
    u64 *ukey;     // n: decoded keys of the sorted batch
    u64 *useg;     // n: run key pid << 32 | segment (| RK_OP for inserts) of effective updates, else RK_NONE
    u64 *ckey;     // n: compacted effective updates (sorted by pid, key)
    u64 *cseg;     // n: their run keys
    i32 *pd;       // n+1: exclusive prefix of +1 (insert) / -1 (delete)
    u32 *rlevel;   // n: chosen level per run start
    uint4 *rwin;   // n: per run start, its chosen window's update range [u0, u1) and counts C, M
    i32 *part;     // 2 * nblocks + 2: block partials (count, delta)
    Work *work[NUM_CLASSES];  // n each
    u32 *lchunk;   // n+1: chunk prefix over large windows
    u64 *lmbase;   // n: merged-scratch base per large window
    u32 *lhdr;     // 4: large windows, chunks
    u32 *wfirst;   // WFIRST_N + 1: first compacted update per index-level window (lone PMA)
    Status *status;
};

// batch descriptors small enough to travel as kernel arguments
static constexpr u32 INLINE_SEGS = 4;
struct InlineSegs {
    BatchSeg s[INLINE_SEGS];
    u32 n;
};

// sort.hip: a lone PMA's batch of at most 2^17 keys from inline sources is
// sorted with the batch's first work (batch_begin's, without clearing the
// sort scratch) carried by the sort kernels: no batch_begin launch, and the
// pyramid's pyr_chunks chunk levels owed by the last batch summed there too
bool sort_carries_first(u64 n);
u64 *pack_and_sort_first(const Group &g, Status *status, u32 *wfirst, const InlineSegs &in, u64 n, u32 pshift,
                         u32 sbits, u32 hi_bit, u64 *k0, u64 *k1, u64 *k2, u32 *scratch, u32 pyr_chunks,
                         hipStream_t st);

u32 locate_blocks(u64 n);
// also clears zero[0, nzero) (the sort's digit totals)
// and resets the update index wfirst
hipError_t batch_begin(const Group &g, Status *status, const InlineSegs &in, BatchSeg *segs, u32 *zero, u32 nzero,
                       u32 *wfirst, hipStream_t st);
hipError_t locate_and_compact(const Group &g, const u64 *pk, u64 n, const BatchBufs &b, hipStream_t st);
// fixed grids for graph replays (work counts read on the device)
hipError_t rebalance_fixed(const Group &g, const BatchBufs &b, u32 runs_cap, u64 *merged, hipStream_t st);
hipError_t density_check(const Group &g, u64 n, const BatchBufs &b, hipStream_t st);
hipError_t elect_windows(const Group &g, u64 n, const BatchBufs &b, hipStream_t st);
// zeroes the work counts of the status (before a repeated election)
hipError_t reset_work_counts(Status *status, hipStream_t st);
hipError_t rebalance_lds(const Group &g, const BatchBufs &b, int cls, u32 grid, hipStream_t st, hipEvent_t e0 = nullptr,
                         hipEvent_t e1 = nullptr);
// with events (timing): the launch records its own start and end in them
hipError_t rebalance_seg(const Group &g, const BatchBufs &b, u64 windows, hipStream_t st, hipEvent_t e0 = nullptr,
                         hipEvent_t e1 = nullptr);
// single-segment and one-wave windows in one launch (a lone PMA's early
// rebalance), its grid sized for at most `bound` windows, the counts read
// from the status on the device (queued before the host knows them)
hipError_t rebalance_seg_wave_bound(const Group &g, const BatchBufs &b, u64 bound, hipStream_t st,
                                    hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr);
u32 seg_lanes_log2(u32 log2S, int slots_log2);
hipError_t group_pyramid(const Group &g, u32 nblocks, hipStream_t st);
u32 pyramid_blocks(u32 H);

// global (multi-workgroup) window path, in place, and the full rebuild
hipError_t rebalance_large(const Group &g, const BatchBufs &b, u64 *merged, u32 grid, hipStream_t st);
// ubeg: nseg + 1 words of scratch
// root rebuild / resize of PMA pid: its keys after the batch (updates [u0, u1)
// of the compacted batch) from src's arrays straight into dst's, spread evenly
// (index: rebuild_index_words(src) words of scratch for the per-chunk
// update starts and ranks)
hipError_t rebuild_scatter(const DevState &src, const DevState &dst, u32 pid, u32 u0, u32 u1, const BatchBufs &b,
                           u64 m, u32 *index, hipStream_t st);
u64 rebuild_index_words(const DevState &src);
hipError_t rebuild_spread(const DevState &dst, const u64 *merged, u64 m, hipStream_t st);
hipError_t window_spread(const DevState &t, u64 a, u32 L, const u64 *merged, u64 m, u64 *stats, int cls_slot,
                         hipStream_t st);
hipError_t rebuild_pyramid(const DevState &d, bool full, hipStream_t st);
hipError_t make_markers(u64 *out, u32 V, hipStream_t st);
hipError_t fill_u32(u32 *p, u64 n, u32 v, hipStream_t st);
hipError_t fill_u64(u64 *p, u64 n, u64 v, hipStream_t st);
hipError_t copy_device(void *dst, const void *src, u64 bytes, hipStream_t st);

// gen_dev.hip
hipError_t gen_rmat_device(u32 scale, u64 n, double a, double b, double c, u64 seed, u64 *out, hipStream_t st);

// views.hip
hipError_t export_compacted(const DevState &d, u32 *gpre, u32 *scan_scratch, u64 *edge_keys, u32 *row_std,
                            u64 *all_keys, u32 *csr_row, u32 *csr_col, hipStream_t st);
// the export with its scratch in a static device array (no allocation);
// setup finds the array's address on the current device
hipError_t export_scratch_setup(int device);
u64 export_static_max_segments();
hipError_t export_compacted_static(int device, const DevState &d, u64 *edge_keys, u32 *row_std, hipStream_t st);
hipError_t edge_query(const DevState &d, const u64 *q, u64 n, u32 *res, hipStream_t st);
hipError_t neighbor_count(const DevState &d, const u32 *srcs, u64 n, u32 *len, hipStream_t st);
hipError_t neighbor_gather(const DevState &d, const u32 *srcs, u64 n, const u32 *offs, u32 *out, hipStream_t st);
hipError_t check_invariants(const DevState &d, u32 *err, hipStream_t st);

// analytics.hip
__host__ __device__ __forceinline__ u64 bfs_visited_words(u64 V) { return (V + 31) / 32; }
hipError_t bfs_init(u32 *levels, u32 V, u32 src, u32 *queue, u32 *counters, u32 *visited, hipStream_t st);
u64 bfs_run_words(u64 slots);
hipError_t bfs_level(const DevState &d, const u32 *q, u32 qn, u32 *len, u64 *first, u32 *runs, u32 *scan_scratch,
                     u32 *levels, u32 *visited, u32 level, u32 *qnext, u32 *counters, u32 expand_grid,
                     hipStream_t st);
hipError_t pr_outdeg(const DevState &d, u32 *gpre, u32 *scan_scratch, u32 *outdeg, hipStream_t st);
hipError_t pr_init(float *pr, u32 V, float *red, hipStream_t st);
u64 pr_acc_floats(u32 V);
u32 pr_red_floats();
hipError_t pr_iterate(const DevState &d, const float *pr, const u32 *outdeg, float *c, float *acc, float *out,
                      float alpha, float *red, float *red_next, hipStream_t st);
hipError_t pr_scale(float *pr, u32 V, const float *red, hipStream_t st);
// binned push: geometry of one PageRank call
struct PrBins {
    u32 part_log;    // slots per part (log2)
    u32 parts;       // parts of the slot array (<= 1024)
    u32 bins;        // destination bins of 8192 vertices
    u32 chunk;       // pairs per accumulation chunk
    u32 max_chunks;  // bound on the chunks of all bins (the acc grid)
    u32 staged;      // bin pass stages whole 32-byte chunks in LDS (runs padded)
    u64 pairs;       // live edges
};
bool pr_bins_fit(u32 V);
PrBins pr_bins_plan(const DevState &d);
u64 pr_bins_count_words(const PrBins &b);
hipError_t pr_bins_count(const DevState &d, const PrBins &b, u32 *off, u32 *scan_scratch, hipStream_t st);
void pr_bins_chunking(PrBins &b, u64 pairs);
u64 pr_bins_partial_floats(const PrBins &b);
hipError_t pr_bins_chunks(const PrBins &b, const u32 *off, u32 *cb, u32 *scan_scratch, hipStream_t st);
hipError_t pr_iterate_binned(const DevState &d, const PrBins &b, const float *pr, const u32 *outdeg, float *c,
                             const u32 *off, const u32 *cb, u64 *pairs, float *partial, float *out, float alpha,
                             float *red, float *red_next, hipStream_t st);

}  // namespace pma
