/* The library's host code under AddressSanitizer and UBSan: built by
 * tests/test_sanitizers.py against a copy of liblpma whose host code is
 * instrumented (hipcc -Xarch_host -fsanitize=...), it runs every argument
 * check of the C ABI with invalid arguments, the host generators and the
 * edgelist reader on valid and malformed files, and creation without a
 * usable device (an error, not a fault).  No GPU is needed.
 *
 *   san_host TMPDIR          prints "san host ok" on success
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lpma.h"

#define CHECK(c)                                                           \
    do {                                                                   \
        if (!(c)) {                                                        \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            return 1;                                                      \
        }                                                                  \
    } while (0)

static int write_file(const char *path, const char *text) {
    FILE *f = fopen(path, "w");
    if (!f) return 1;
    fputs(text, f);
    fclose(f);
    return 0;
}

static int check_arguments(void) {
    const int INV = LPMA_ERR_INVALID_ARGUMENT;
    uint64_t k = 0, rem = 0;
    uint32_t u = 0, depth = 0;
    lpma_view_t view;
    lpma_stats_t stats;
    lpma_profile_t prof;
    lpma_config_t cfg;
    lpma_config_default(&cfg);
    CHECK(cfg.segment_slots == 32);
    for (int c = -2; c < 16; ++c) CHECK(lpma_strerror(c) != NULL);

    CHECK(lpma_create(NULL, 10, NULL, NULL) == INV);
    lpma_t *h = NULL;
    CHECK(lpma_create(&h, 0, NULL, NULL) == INV && h == NULL);
    cfg.segment_slots = 3;  /* not a power of two */
    CHECK(lpma_create(&h, 10, &cfg, NULL) == INV && h == NULL);
    lpma_config_default(&cfg);
    cfg.density_lower_leaf = 0.9;
    cfg.density_upper_leaf = 0.5;  /* lower above upper */
    CHECK(lpma_create(&h, 10, &cfg, NULL) == INV && h == NULL);

    CHECK(lpma_destroy(NULL) == INV);
    CHECK(lpma_update(NULL, &k, 1, LPMA_OP_INSERT, NULL) == INV);
    CHECK(lpma_update_mixed(NULL, &k, 1, &k, 1, NULL) == INV);
    CHECK(lpma_update_sequence(NULL, NULL, NULL, NULL, NULL, 0, NULL) == INV);
    CHECK(lpma_bulk_load(NULL, &k, 1, NULL) == INV);
    CHECK(lpma_get_view(NULL, &view) == INV);
    CHECK(lpma_copy_view(NULL, &k, &u, NULL) == INV);
    CHECK(lpma_export_compacted(NULL, &k, &u, NULL) == INV);
    CHECK(lpma_export_compacted_noalloc(NULL, &k, &u, NULL) == INV);
    CHECK(lpma_export_data(NULL, &k, NULL) == INV);
    CHECK(lpma_to_csr(NULL, &u, &u, NULL) == INV);
    CHECK(lpma_edge_query(NULL, &k, 1, &u, NULL) == INV);
    CHECK(lpma_neighbor_count(NULL, &u, 1, &u, NULL) == INV);
    CHECK(lpma_bfs(NULL, 0, &u, &depth, &k, NULL) == INV);
    CHECK(lpma_check(NULL, &k, NULL) == INV);
    CHECK(lpma_get_stats(NULL, &stats) == INV);
    CHECK(lpma_get_profile(NULL, &prof) == INV);
    CHECK(lpma_reset_stats(NULL) == INV);
    CHECK(lpma_set_profiling(NULL, 1) == INV);
    CHECK(lpma_set_graph(NULL, 1) == INV);
    CHECK(lpma_reserve(NULL, 10, NULL) == INV);
    CHECK(lpma_reserve_capacity(NULL, 10, NULL) == INV);
    CHECK(lpma_set_launch(NULL, 4) == INV);
    CHECK(lpma_snapshot_create(NULL, NULL, NULL) == INV);
    CHECK(lpma_snapshot_destroy(NULL) == INV);
    CHECK(lpma_snapshot_get_view(NULL, &view) == INV);
    CHECK(lpma_group_create(NULL, NULL, 0) == INV);
    lpma_group_t *g = NULL;
    CHECK(lpma_group_create(&g, NULL, 1) == INV && g == NULL);
    CHECK(lpma_server_create(NULL, NULL, NULL) == INV);

    /* the pool controls and the fault injection are host state */
    CHECK(lpma_set_pool_budget(0) == LPMA_OK);
    CHECK(lpma_inject_alloc_failure(5, &rem) == LPMA_OK);
    CHECK(lpma_inject_alloc_failure(0, &rem) == LPMA_OK && rem == 5);

    /* a valid creation: without a usable device it fails cleanly */
    int r = lpma_create(&h, 100, NULL, NULL);
    if (r == LPMA_OK) {
        CHECK(lpma_destroy(h) == LPMA_OK);
    } else {
        CHECK(h == NULL);
    }
    return 0;
}

static int check_generators(void) {
    const uint64_t n = 20000;
    uint64_t *a = (uint64_t *)malloc(n * sizeof(uint64_t));
    uint64_t *b = (uint64_t *)malloc(n * sizeof(uint64_t));
    CHECK(lpma_gen_uniform(1000, n, 7, a) == LPMA_OK);
    CHECK(lpma_gen_uniform(1000, n, 7, b) == LPMA_OK && memcmp(a, b, n * sizeof(uint64_t)) == 0);
    for (uint64_t i = 0; i < n; ++i) CHECK((a[i] >> 32) < 1000 && (uint32_t)a[i] < 1000 && (a[i] >> 32) != (uint32_t)a[i]);
    CHECK(lpma_gen_uniform(3, 7, 1, a) != LPMA_OK);  /* more distinct edges than 3 vertices hold */
    CHECK(lpma_gen_rmat(12, n, 0.57, 0.19, 0.19, 2, a) == LPMA_OK);
    CHECK(lpma_gen_rmat_samples(12, n, 0.57, 0.19, 0.19, 3, b) == LPMA_OK);
    CHECK(lpma_gen_rmat(12, n, 0.9, 0.2, 0.2, 2, a) != LPMA_OK);  /* probabilities beyond 1 */
    CHECK(lpma_shuffle(a, n, 4) == LPMA_OK);
    CHECK(lpma_shuffle(a, 0, 4) == LPMA_OK);
    /* absent edges against a sorted present set */
    for (uint64_t i = 1; i < n; ++i) {
        uint64_t x = b[i], j = i;
        while (j > 0 && b[j - 1] > x) { b[j] = b[j - 1]; --j; }
        b[j] = x;
    }
    CHECK(lpma_gen_absent(1u << 12, b, n, 5000, 6, a) == LPMA_OK);
    CHECK(lpma_mt64_fill(5489, 1000, a) == LPMA_OK);
    free(a);
    free(b);
    return 0;
}

static int check_edgelist(const char *dir) {
    char path[4096];
    uint32_t V = 0;
    uint64_t E = 0;
    uint64_t out[8];
    snprintf(path, sizeof path, "%s/ok.txt", dir);
    CHECK(write_file(path, "5 4\n0 1\n1 2\n2 3\n4 0\n") == 0);
    CHECK(lpma_read_edgelist(path, &V, &E, 0, NULL) == LPMA_OK && V == 5 && E == 4);
    CHECK(lpma_read_edgelist(path, &V, &E, 0, out) == LPMA_OK && out[0] == 1ull && out[3] == (4ull << 32));
    CHECK(lpma_read_edgelist(path, &V, &E, 2, out) == LPMA_OK && E == 2);
    /* malformed: fewer edges than announced, a vertex out of range, text,
     * an empty file, a missing file */
    snprintf(path, sizeof path, "%s/short.txt", dir);
    CHECK(write_file(path, "5 4\n0 1\n1 2\n") == 0);
    CHECK(lpma_read_edgelist(path, &V, &E, 0, out) != LPMA_OK);
    snprintf(path, sizeof path, "%s/range.txt", dir);
    CHECK(write_file(path, "3 2\n0 1\n1 7\n") == 0);
    CHECK(lpma_read_edgelist(path, &V, &E, 0, out) != LPMA_OK);
    snprintf(path, sizeof path, "%s/text.txt", dir);
    CHECK(write_file(path, "three two\nzero one\n") == 0);
    CHECK(lpma_read_edgelist(path, &V, &E, 0, out) != LPMA_OK);
    snprintf(path, sizeof path, "%s/empty.txt", dir);
    CHECK(write_file(path, "") == 0);
    CHECK(lpma_read_edgelist(path, &V, &E, 0, out) != LPMA_OK);
    snprintf(path, sizeof path, "%s/missing.txt", dir);
    CHECK(lpma_read_edgelist(path, &V, &E, 0, out) != LPMA_OK);
    CHECK(lpma_read_edgelist(NULL, &V, &E, 0, out) != LPMA_OK);
    return 0;
}

int main(int argc, char **argv) {
    const char *dir = argc > 1 ? argv[1] : "/tmp";
    if (check_arguments() || check_generators() || check_edgelist(dir)) return 1;
    printf("san host ok\n");
    return 0;
}
