// views.hip -- read paths over the gapped layout: compacted export (the
// parity surface), the reference's CSR convention, edge-membership queries,
// neighbor-list reads, and an on-device invariant checker.
#include <mutex>

#include "common.hpp"
#include "kernels.hpp"

namespace pma {

// Compacted export.  With p = rank of a live slot among live slots (a scan of
// the segment counts):
//   all_keys[p] = key                          (get_data_rpma, rpma.cuh:1735-1761)
//   edge_keys[p - src] = key, row_std[src+1] = p - src for src's marker
//                                              (standard CSR, markers excluded)
//   csr_col[p] = dst, or src for a marker; csr_row[src+1] = p + 1
//                                              (to_csr_rpma, rpma.cuh:2192-2295)
__global__ void k_export(DevState d, const u32 *__restrict__ gpre, u64 *__restrict__ edge_keys,
                         u32 *__restrict__ row_std, u64 *__restrict__ all_keys, u32 *__restrict__ csr_row,
                         u32 *__restrict__ csr_col) {
    const u64 W = nseg_of(d) << d.log2S;
    const u32 Smask = (1u << d.log2S) - 1;
    u64 t0 = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (t0 == 0) {
        if (row_std) row_std[0] = 0;
        if (csr_row) csr_row[0] = 0;
    }
    for (u64 t = t0; t < W; t += (u64)gridDim.x * blockDim.x) {
        u64 seg = t >> d.log2S;
        u32 lane = (u32)t & Smask;
        if (lane >= d.cnt[seg]) continue;
        u64 key = d.keys[t];
        u64 p = (u64)gpre[seg] + lane;
        u32 src = (u32)(key >> 32);
        if (all_keys) all_keys[p] = key;
        if (is_marker(key) && src < d.V) {
            if (row_std) row_std[src + 1] = (u32)(p - src);
            if (csr_col) csr_col[p] = src;
            if (csr_row) csr_row[src + 1] = (u32)(p + 1);
        } else {
            if (edge_keys && p >= src) edge_keys[p - src] = key;
            if (csr_col) csr_col[p] = (u32)key;
        }
    }
}

hipError_t export_compacted(const DevState &d, u32 *gpre, u32 *scan_scratch, u64 *edge_keys, u32 *row_std,
                            u64 *all_keys, u32 *csr_row, u32 *csr_col, hipStream_t st) {
    u64 nseg = nseg_of(d);
    (void)device_excl_scan(d.cnt, nseg, gpre, scan_scratch, st);
    k_export<<<grid_for(nseg << d.log2S, 256, 16384), 256, 0, st>>>(d, gpre, edge_keys, row_std, all_keys, csr_row,
                                                                    csr_col);
    return hipGetLastError();
}

// ------------------------------------------- export without allocation
//
// The compacted export's scratch (the segment-count prefix, nseg + 1 words,
// and the scan's partials) in a static device array, for callers that may
// not allocate (graph capture, serving loops).  Every call on a device
// shares that device's copy: calls that overlap on one device race.  Nothing
// in it needs clearing: the scan writes every word before it is read.

static constexpr u64 EXPORT_STATIC_MAX_SEGS = 1ull << 21;
static constexpr u64 EXPORT_STATIC_WORDS =
    (EXPORT_STATIC_MAX_SEGS + 1 + (EXPORT_STATIC_MAX_SEGS + 2047) / 2048 + 2 + 8 + 63) / 64 * 64;
static __device__ __align__(256) u32 scratch[EXPORT_STATIC_WORDS];

static constexpr int EXPORT_STATIC_MAX_DEVICES = 64;
static u32 *g_scratch_ptr[EXPORT_STATIC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// The device address of `scratch` on `device` (the current device), found
// once per device; the host-side name of the array is not a device address.
hipError_t export_scratch_setup(int device) {
    if (device < 0 || device >= EXPORT_STATIC_MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[device] = (u32 *)p;
    return hipSuccess;
}

u64 export_static_max_segments() { return EXPORT_STATIC_MAX_SEGS; }

hipError_t export_compacted_static(int device, const DevState &d, u64 *edge_keys, u32 *row_std, hipStream_t st) {
    if (nseg_of(d) > EXPORT_STATIC_MAX_SEGS) return hipErrorInvalidValue;
    hipError_t e = export_scratch_setup(device);
    if (e != hipSuccess) return e;
    u32 *gpre = g_scratch_ptr[device];
    u32 *scan = gpre + EXPORT_STATIC_MAX_SEGS + 1;
    return export_compacted(d, gpre, scan, edge_keys, row_std, nullptr, nullptr, nullptr, st);
}

// Exact edge membership (edge_query_rpma / edge_query_kernel,
// rpma.cuh:2090-2190): 1 when the key is a live edge, 0 otherwise
// (row markers and out-of-range keys are not edges).
__global__ __launch_bounds__(256) void k_edge_query(DevState d, const u64 *__restrict__ q, u64 n,
                                                     u32 *__restrict__ res) {
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x) {
        u64 key = q[i];
        u32 src = (u32)(key >> 32), dst = (u32)key;
        u32 r = 0;
        if (src < d.V && dst < d.V) {
            u32 s = locate_segment(d.sep, d.H, key);
            r = segment_holds(d, s, key) ? 1u : 0u;
        }
        res[i] = r;
    }
}

hipError_t edge_query(const DevState &d, const u64 *q, u64 n, u32 *res, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_edge_query<<<grid_for(n, 256, 16384), 256, 0, st>>>(d, q, n, res);
    return hipGetLastError();
}

// Neighbor lists (query_rpma, rpma.cuh:1827-2036, without the time window):
// row r's edges are the live slots strictly between the markers of rows
// r-1 and r, i.e. slots (row_off[r], row_off[r+1]) (from slot 0 for r = 0).
// One wave per queried vertex walks that slot range.
__device__ __forceinline__ void row_range(const DevState &d, u32 r, u64 &lo, u64 &hi) {
    lo = r == 0 ? 0 : (u64)d.row_off[r] + 1;
    hi = d.row_off[r + 1];
}

__global__ void k_nbr_count(DevState d, const u32 *__restrict__ srcs, u64 n, u32 *__restrict__ len) {
    const u64 waves = ((u64)gridDim.x * blockDim.x) >> 6;
    const u32 lane = lane_id();
    for (u64 i = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += waves) {
        u32 r = srcs[i];
        u32 c = 0;
        if (r < d.V) {
            u64 lo, hi;
            row_range(d, r, lo, hi);
            for (u64 t = lo + lane; t < hi; t += 64) c += d.keys[t] != KEY_NONE ? 1u : 0u;
        }
        c = wave_sum(c);
        if (lane == 0) len[i] = c;
    }
}

__global__ void k_nbr_gather(DevState d, const u32 *__restrict__ srcs, u64 n, const u32 *__restrict__ offs,
                             u32 *__restrict__ out) {
    const u64 waves = ((u64)gridDim.x * blockDim.x) >> 6;
    const u32 lane = lane_id();
    for (u64 i = ((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6; i < n; i += waves) {
        u32 r = srcs[i];
        if (r >= d.V) continue;
        u64 lo, hi;
        row_range(d, r, lo, hi);
        u64 o = offs[i];
        for (u64 t0 = lo; t0 < hi; t0 += 64) {
            u64 t = t0 + lane;
            u64 key = t < hi ? d.keys[t] : KEY_NONE;
            bool live = key != KEY_NONE;
            u64 m = __ballot(live);
            if (live) out[o + __popcll(m & ((1ull << lane) - 1ull))] = (u32)key;
            o += __popcll(m);
        }
    }
}

hipError_t neighbor_count(const DevState &d, const u32 *srcs, u64 n, u32 *len, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_nbr_count<<<grid_for(n * 64, 256, 16384), 256, 0, st>>>(d, srcs, n, len);
    return hipGetLastError();
}

hipError_t neighbor_gather(const DevState &d, const u32 *srcs, u64 n, const u32 *offs, u32 *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    k_nbr_gather<<<grid_for(n * 64, 256, 16384), 256, 0, st>>>(d, srcs, n, offs, out);
    return hipGetLastError();
}

// Invariants of the layout, one thread per segment; err[0] counts
// violations: a segment's live keys are front-packed, strictly increasing,
// inside [sep[s], sep[s+1]); its count matches; a marker's slot is its row
// offset.  (The reference has no such checker; its only checks are the
// host-side comparisons in test_update_rpma.cu:131-165.)
__global__ void k_check(DevState d, u32 *err) {
    const u64 nseg = nseg_of(d);
    const u32 S = 1u << d.log2S;
    u32 bad = 0;
    for (u64 s = (u64)blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += (u64)gridDim.x * blockDim.x) {
        const u64 *seg = d.keys + (s << d.log2S);
        u32 c = d.cnt[s];
        if (c > S) { ++bad; continue; }
        u64 lo = d.sep[s];
        u64 hi = s + 1 < nseg ? d.sep[s + 1] : KEY_NONE;
        if (s == 0 && lo != 0) ++bad;
        if (hi < lo) ++bad;
        for (u32 l = 0; l < S; ++l) {
            u64 k = seg[l];
            if (l < c) {
                if (k == KEY_NONE) ++bad;
                if (k < lo || (s + 1 < nseg && k >= hi)) ++bad;
                if (l > 0 && k <= seg[l - 1]) ++bad;
                if (is_marker(k)) {
                    u32 r = (u32)(k >> 32);
                    if (r >= d.V || d.row_off[r + 1] != (u32)((s << d.log2S) + l)) ++bad;
                }
            } else if (k != KEY_NONE) {
                ++bad;
            }
        }
    }
    if (bad) atomicAdd(err, bad);
}

hipError_t check_invariants(const DevState &d, u32 *err, hipStream_t st) {
    k_check<<<grid_for(nseg_of(d), 256, 16384), 256, 0, st>>>(d, err);
    return hipGetLastError();
}

}  // namespace pma
