/*
 * cabi_client.c -- a C caller of include/lpma.h, the way a maintainer binds
 * the library from C: every entry point through dlopen/dlsym with the
 * header's own prototypes, device buffers from the HIP runtime's C API.
 *
 *   cabi_client layout          struct sizes and field offsets, one per line
 *   cabi_client symbols LIB     resolves every entry point (no device work)
 *   cabi_client stream LIB      an update stream on device 0, checked against
 *                               a set reference kept by this program
 *
 * The test that builds it compares `layout` with the ctypes mirror of the
 * header, so a struct that drifts between the two is caught.
 */
#define __HIP_PLATFORM_AMD__
#include <dlfcn.h>
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../include/lpma.h"

/* The ABI the header fixes on x86-64 (LP64): a change of any of these is a
 * break for every binary caller. */
_Static_assert(sizeof(lpma_config_t) == 48, "lpma_config_t");
_Static_assert(offsetof(lpma_config_t, device) == 40, "lpma_config_t.device");
_Static_assert(sizeof(lpma_view_t) == 64, "lpma_view_t");
_Static_assert(offsetof(lpma_view_t, version) == 56, "lpma_view_t.version");
_Static_assert(sizeof(lpma_stats_t) == 584, "lpma_stats_t");
_Static_assert(sizeof(lpma_profile_t) == 120, "lpma_profile_t");
_Static_assert(sizeof(lpma_server_result_t) == 80, "lpma_server_result_t");
_Static_assert(sizeof(int) == 4 && sizeof(void *) == 8, "LP64");

#define F(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
#define S(T) printf(#T " %zu\n", sizeof(T))

static void layout(void) {
    S(lpma_config_t);
    F(lpma_config_t, segment_slots); F(lpma_config_t, density_lower_leaf); F(lpma_config_t, density_lower_root);
    F(lpma_config_t, density_upper_leaf); F(lpma_config_t, density_upper_root); F(lpma_config_t, device);
    S(lpma_view_t);
    F(lpma_view_t, keys); F(lpma_view_t, row_offsets); F(lpma_view_t, capacity); F(lpma_view_t, live_keys);
    F(lpma_view_t, edges); F(lpma_view_t, num_vertices); F(lpma_view_t, segment_slots);
    F(lpma_view_t, tree_height); F(lpma_view_t, version);
    S(lpma_stats_t);
    F(lpma_stats_t, updates_calls); F(lpma_stats_t, effective_updates); F(lpma_stats_t, windows_per_level);
    F(lpma_stats_t, slots_per_level); F(lpma_stats_t, rebuilds); F(lpma_stats_t, rebuild_slots);
    F(lpma_stats_t, marker_moves); F(lpma_stats_t, alg_bytes_wave); F(lpma_stats_t, alg_bytes_block);
    F(lpma_stats_t, alg_bytes_large); F(lpma_stats_t, alg_bytes_seg);
    S(lpma_profile_t);
    F(lpma_profile_t, sort_ms); F(lpma_profile_t, locate_ms); F(lpma_profile_t, elect_ms);
    F(lpma_profile_t, rebalance_wave_ms); F(lpma_profile_t, rebalance_block_ms); F(lpma_profile_t, large_ms);
    F(lpma_profile_t, rebuild_ms); F(lpma_profile_t, pyramid_ms); F(lpma_profile_t, rebalance_seg_ms);
    F(lpma_profile_t, launches_wave); F(lpma_profile_t, launches_block); F(lpma_profile_t, launches_seg);
    F(lpma_profile_t, graph_ms); F(lpma_profile_t, graph_launches); F(lpma_profile_t, graph_captures);
    S(lpma_server_config_t);
    F(lpma_server_config_t, bfs_workers); F(lpma_server_config_t, hbm_budget_bytes);
    F(lpma_server_config_t, max_merge_keys); F(lpma_server_config_t, max_merge_requests);
    F(lpma_server_config_t, max_merge_vertices); F(lpma_server_config_t, snapshot_max_age_us);
    F(lpma_server_config_t, search_workgroups);
    S(lpma_server_result_t);
    F(lpma_server_result_t, status); F(lpma_server_result_t, kind); F(lpma_server_result_t, applied_updates);
    F(lpma_server_result_t, version); F(lpma_server_result_t, queue_us); F(lpma_server_result_t, latency_us);
    F(lpma_server_result_t, count); F(lpma_server_result_t, offsets); F(lpma_server_result_t, values);
    F(lpma_server_result_t, depth); F(lpma_server_result_t, edges_scanned);
    S(lpma_server_stats_t);
    F(lpma_server_stats_t, submitted); F(lpma_server_stats_t, completed); F(lpma_server_stats_t, update_runs);
    F(lpma_server_stats_t, read_runs); F(lpma_server_stats_t, snapshots);
    F(lpma_server_stats_t, max_concurrent_bfs); F(lpma_server_stats_t, admission_waits);
    F(lpma_server_stats_t, budget_bytes); F(lpma_server_stats_t, in_use_bytes); F(lpma_server_stats_t, peak_bytes);
    F(lpma_server_stats_t, applied_updates); F(lpma_server_stats_t, update_us); F(lpma_server_stats_t, read_us);
    F(lpma_server_stats_t, dispatch_us); F(lpma_server_stats_t, stage_us); F(lpma_server_stats_t, pipeline_us);
    F(lpma_server_stats_t, finish_us);
}

/* entry points resolved with the header's prototypes */
static struct {
    __typeof__(&lpma_config_default) config_default;
    __typeof__(&lpma_create) create;
    __typeof__(&lpma_destroy) destroy;
    __typeof__(&lpma_update) update;
    __typeof__(&lpma_update_mixed) update_mixed;
    __typeof__(&lpma_bulk_load) bulk_load;
    __typeof__(&lpma_get_view) get_view;
    __typeof__(&lpma_export_compacted) export_compacted;
    __typeof__(&lpma_export_data) export_data;
    __typeof__(&lpma_edge_query) edge_query;
    __typeof__(&lpma_check) check;
    __typeof__(&lpma_strerror) strerror_;
} api;

static void *resolve(void *lib, const char *name) {
    void *p = dlsym(lib, name);
    if (!p) {
        fprintf(stderr, "missing symbol %s\n", name);
        exit(3);
    }
    return p;
}

static void *load(const char *path) {
    void *lib = dlopen(path, RTLD_NOW | RTLD_LOCAL);
    if (!lib) {
        fprintf(stderr, "dlopen %s: %s\n", path, dlerror());
        exit(3);
    }
    api.config_default = (__typeof__(api.config_default))resolve(lib, "lpma_config_default");
    api.create = (__typeof__(api.create))resolve(lib, "lpma_create");
    api.destroy = (__typeof__(api.destroy))resolve(lib, "lpma_destroy");
    api.update = (__typeof__(api.update))resolve(lib, "lpma_update");
    api.update_mixed = (__typeof__(api.update_mixed))resolve(lib, "lpma_update_mixed");
    api.bulk_load = (__typeof__(api.bulk_load))resolve(lib, "lpma_bulk_load");
    api.get_view = (__typeof__(api.get_view))resolve(lib, "lpma_get_view");
    api.export_compacted = (__typeof__(api.export_compacted))resolve(lib, "lpma_export_compacted");
    api.export_data = (__typeof__(api.export_data))resolve(lib, "lpma_export_data");
    api.edge_query = (__typeof__(api.edge_query))resolve(lib, "lpma_edge_query");
    api.check = (__typeof__(api.check))resolve(lib, "lpma_check");
    api.strerror_ = (__typeof__(api.strerror_))resolve(lib, "lpma_strerror");
    return lib;
}

/* ---------------------------------------------------------- set reference */

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

static size_t sort_unique(uint64_t *a, size_t n) {
    if (!n) return 0;
    qsort(a, n, sizeof(uint64_t), cmp_u64);
    size_t m = 1;
    for (size_t i = 1; i < n; ++i)
        if (a[i] != a[m - 1]) a[m++] = a[i];
    return m;
}

static int has(const uint64_t *a, size_t n, uint64_t k) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = (lo + hi) / 2;
        if (a[mid] < k) lo = mid + 1; else hi = mid;
    }
    return lo < n && a[lo] == k;
}

/* ref = sort(unique((ref \ del) | ins)); ref has room for n + nins */
static size_t ref_apply(uint64_t *ref, size_t n, const uint64_t *ins, size_t nins, const uint64_t *del, size_t ndel) {
    uint64_t *d = malloc((ndel + 1) * sizeof(uint64_t));
    memcpy(d, del, ndel * sizeof(uint64_t));
    ndel = sort_unique(d, ndel);
    size_t m = 0;
    for (size_t i = 0; i < n; ++i)
        if (!has(d, ndel, ref[i])) ref[m++] = ref[i];
    free(d);
    memcpy(ref + m, ins, nins * sizeof(uint64_t));
    return sort_unique(ref, m + nins);
}

static uint64_t lcg = 0x9E3779B97F4A7C15ull;
static uint32_t next32(void) {
    lcg = lcg * 6364136223846793005ull + 1442695040888963407ull;
    return (uint32_t)(lcg >> 33);
}

#define HIPOK(x)                                                              \
    do {                                                                      \
        hipError_t e_ = (x);                                                  \
        if (e_ != hipSuccess) {                                               \
            fprintf(stderr, "%s:%d %s: HIP error %d\n", __FILE__, __LINE__, #x, (int)e_); \
            exit(4);                                                          \
        }                                                                     \
    } while (0)
#define LPOK(x)                                                               \
    do {                                                                      \
        int r_ = (x);                                                         \
        if (r_ != LPMA_OK) {                                                  \
            fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, api.strerror_(r_)); \
            exit(5);                                                          \
        }                                                                     \
    } while (0)

static int stream(const char *path) {
    load(path);
    const uint32_t V = 3000;
    const size_t B = 20000, ROUNDS = 8;
    size_t cap = B * (ROUNDS + 1) + 16;
    uint64_t *ref = calloc(cap, sizeof(uint64_t)), *ins = calloc(B, 8), *del = calloc(B, 8), *got = calloc(cap, 8);
    uint32_t *ro = calloc(V + 1, 4);
    uint64_t *d_ins, *d_del, *d_out;
    uint32_t *d_ro;
    HIPOK(hipSetDevice(0));
    HIPOK(hipMalloc((void **)&d_ins, B * 8));
    HIPOK(hipMalloc((void **)&d_del, B * 8));
    HIPOK(hipMalloc((void **)&d_out, cap * 8));
    HIPOK(hipMalloc((void **)&d_ro, (V + 1) * 4));
    hipStream_t st;
    HIPOK(hipStreamCreate(&st));

    lpma_config_t cfg;
    api.config_default(&cfg);
    lpma_t *h = NULL;
    LPOK(api.create(&h, V, &cfg, st));
    size_t n = 0;
    for (size_t r = 0; r < ROUNDS; ++r) {
        size_t ni = B, nd = r ? B / 2 : 0;
        for (size_t i = 0; i < ni; ++i) ins[i] = ((uint64_t)(next32() % V) << 32) | (next32() % V);
        /* deletes: half live keys, half random (absent ones are no-ops) */
        for (size_t i = 0; i < nd; ++i)
            del[i] = (i & 1) && n ? ref[next32() % n] : (((uint64_t)(next32() % V) << 32) | (next32() % V));
        HIPOK(hipMemcpy(d_ins, ins, ni * 8, hipMemcpyHostToDevice));
        HIPOK(hipMemcpy(d_del, del, nd * 8, hipMemcpyHostToDevice));
        if (r == 0) {
            LPOK(api.bulk_load(h, d_ins, ni, st));
        } else if (r % 3 == 1) {
            LPOK(api.update(h, d_del, nd, LPMA_OP_DELETE, st));
            LPOK(api.update(h, d_ins, ni, LPMA_OP_INSERT, st));
        } else {
            LPOK(api.update_mixed(h, d_ins, ni, d_del, nd, st));
        }
        n = ref_apply(ref, n, ins, ni, del, nd);

        lpma_view_t v;
        LPOK(api.get_view(h, &v));
        if (v.edges != n || v.num_vertices != V || v.live_keys != n + V || v.segment_slots != 32) {
            fprintf(stderr, "round %zu: view edges %llu live %llu, want %zu\n", r, (unsigned long long)v.edges,
                    (unsigned long long)v.live_keys, n);
            return 6;
        }
        LPOK(api.export_compacted(h, d_out, d_ro, st));
        HIPOK(hipStreamSynchronize(st));
        HIPOK(hipMemcpy(got, d_out, n * 8, hipMemcpyDeviceToHost));
        HIPOK(hipMemcpy(ro, d_ro, (V + 1) * 4, hipMemcpyDeviceToHost));
        if (memcmp(got, ref, n * 8) != 0) {
            fprintf(stderr, "round %zu: exported edges differ\n", r);
            return 7;
        }
        size_t e = 0;
        for (uint32_t s = 0; s <= V; ++s) {
            while (e < n && (uint32_t)(ref[e] >> 32) < s) ++e;
            if (ro[s] != (uint32_t)e) {
                fprintf(stderr, "round %zu: row offset %u = %u, want %zu\n", r, s, ro[s], e);
                return 8;
            }
        }
        /* get_data_rpma: edges and row markers, in key order */
        LPOK(api.export_data(h, d_out, st));
        HIPOK(hipStreamSynchronize(st));
        HIPOK(hipMemcpy(got, d_out, (n + V) * 8, hipMemcpyDeviceToHost));
        size_t a = 0;
        for (uint32_t s = 0; s < V; ++s) {
            while (a < n && (uint32_t)(ref[a] >> 32) == s && got[a + s] == ref[a]) ++a;
            if (got[a + s] != (((uint64_t)s << 32) | LPMA_MARKER_DST)) {
                fprintf(stderr, "round %zu: export_data differs near row %u\n", r, s);
                return 9;
            }
        }
        uint64_t bad = 1;
        LPOK(api.check(h, &bad, st));
        if (bad) return 10;
    }
    LPOK(api.destroy(h));
    HIPOK(hipStreamDestroy(st));
    hipFree(d_ins); hipFree(d_del); hipFree(d_out); hipFree(d_ro);
    free(ref); free(ins); free(del); free(got); free(ro);
    printf("cabi stream ok: %zu rounds, %zu edges\n", ROUNDS, n);
    return 0;
}

int main(int argc, char **argv) {
    if (argc >= 2 && !strcmp(argv[1], "layout")) {
        layout();
        return 0;
    }
    if (argc >= 3 && !strcmp(argv[1], "symbols")) {
        load(argv[2]);
        printf("symbols ok\n");
        return 0;
    }
    if (argc >= 3 && !strcmp(argv[1], "stream")) return stream(argv[2]);
    fprintf(stderr, "usage: %s layout | symbols LIB | stream LIB\n", argv[0]);
    return 2;
}
