/* The CPU oracle under the sanitizers: built by tests/test_sanitizers.py with
 * gcc -fsanitize=address,undefined (and a second time with
 * -fsanitize=thread for its threaded merge and row offsets), it drives every
 * oracle function over seeded random cases and checks them against each
 * other.  Any sanitizer report aborts with a nonzero status.
 *
 *   san_oracle [threads]     prints "san oracle ok" on success
 */
#include "../oracle/oracle.c"

#include <stdio.h>

static u64 rng_state = 0x9E3779B97F4A7C15ull;
static u64 rnd(void) {
    /* SplitMix64 */
    u64 z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static u64 *rand_keys(u32 V, size_t n) {
    u64 *k = (u64 *)malloc((n ? n : 1) * sizeof(u64));
    for (size_t i = 0; i < n; ++i) k[i] = ((rnd() % V) << 32) | (rnd() % V);
    return k;
}

#define CHECK(c)                                                           \
    do {                                                                   \
        if (!(c)) {                                                        \
            fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
            return 1;                                                      \
        }                                                                  \
    } while (0)

static int check_apply(int threads) {
    for (int round = 0; round < 40; ++round) {
        const u32 V = 1 + (u32)(rnd() % 3000);
        const size_t ns = rnd() % 20000, ni = rnd() % 5000, nd = rnd() % 5000;
        u64 *S = rand_keys(V, ns);
        const size_t m = or_sort_unique(S, ns);
        u64 *I = rand_keys(V, ni);
        u64 *D = rand_keys(V, nd);
        for (size_t i = 0; i < nd / 2 && m; ++i) D[i] = S[rnd() % m]; /* deletes of present keys too */
        u64 *a = (u64 *)malloc((m + ni + 1) * sizeof(u64));
        u64 *b = (u64 *)malloc((m + ni + 1) * sizeof(u64));
        const size_t na = or_apply(S, m, I, ni, D, nd, a, 1);
        const size_t nb = or_apply(S, m, I, ni, D, nd, b, threads);
        CHECK(na == nb && (na == 0 || memcmp(a, b, na * sizeof(u64)) == 0));
        CHECK(or_is_sorted_unique(a, na));
        u32 *r1 = (u32 *)malloc((V + 1) * sizeof(u32));
        u32 *r2 = (u32 *)malloc((V + 1) * sizeof(u32));
        CHECK(or_row_offsets(a, na, V, r1) == 0);
        CHECK(or_row_offsets_mt(a, na, V, r2, threads) == 0);
        CHECK(memcmp(r1, r2, (V + 1) * sizeof(u32)) == 0 && r1[V] == na);
        free(S); free(I); free(D); free(a); free(b); free(r1); free(r2);
    }
    return 0;
}

static int check_pma(void) {
    const u32 slots[] = {2, 4, 32, 128};
    for (int c = 0; c < 4; ++c) {
        const u32 V = 500 + (u32)(rnd() % 1500);
        or_pma *p = or_pma_create(V, slots[c], 0.42, 0.42, 1.0, 1.0);
        CHECK(p != NULL);
        u64 *live = (u64 *)malloc(sizeof(u64));
        size_t nlive = 0;
        for (int b = 0; b < 30; ++b) {
            /* growth first, then mostly deletes: resizes both ways */
            const size_t ni = b < 15 ? 1 + rnd() % 4000 : rnd() % 300;
            const size_t nd = b < 15 ? rnd() % 300 : 1 + rnd() % 4000;
            u64 *I = rand_keys(V, ni);
            u64 *D = rand_keys(V, nd);
            for (size_t i = 0; i < nd && nlive; ++i)
                if (rnd() & 1) D[i] = live[rnd() % nlive];
            CHECK(or_pma_update_mixed(p, I, ni, D, nd) == 0);
            u64 *next = (u64 *)malloc((nlive + ni + 1) * sizeof(u64));
            nlive = or_apply(live, nlive, I, ni, D, nd, next, 1);
            free(live);
            live = next;
            /* the PMA's keys: the live edges plus one marker per row */
            u64 *exp = (u64 *)malloc((nlive + V + 1) * sizeof(u64));
            u64 *mk = (u64 *)malloc((size_t)V * sizeof(u64));
            for (u32 r = 0; r < V; ++r) mk[r] = ((u64)r << 32) | 0xFFFFFFFFull;
            const size_t ne = or_apply(live, nlive, mk, V, NULL, 0, exp, 1);
            u64 *got = (u64 *)malloc((or_pma_capacity(p) + 1) * sizeof(u64));
            const u64 ng = or_pma_export(p, got);
            CHECK(ng == ne && memcmp(got, exp, ne * sizeof(u64)) == 0);
            CHECK(or_pma_live(p) == ne);
            free(I); free(D); free(exp); free(mk); free(got);
        }
        or_pma_destroy(p);
        free(live);
    }
    return 0;
}

static int check_analytics(void) {
    for (int round = 0; round < 10; ++round) {
        const u32 V = 1 + (u32)(rnd() % 2000);
        const size_t n = rnd() % 20000;
        u64 *k = rand_keys(V, n);
        const size_t m = or_sort_unique(k, n);
        u32 *ro = (u32 *)malloc((V + 1) * sizeof(u32));
        u32 *col = (u32 *)malloc((m ? m : 1) * sizeof(u32));
        CHECK(or_row_offsets(k, m, V, ro) == 0);
        for (size_t i = 0; i < m; ++i) col[i] = (u32)k[i];
        u32 *lv = (u32 *)malloc((size_t)V * sizeof(u32));
        const u32 depth = or_bfs(ro, col, V, (u32)(rnd() % V), lv);
        CHECK(depth <= V);
        double *pr = (double *)malloc((size_t)V * sizeof(double));
        const double l1 = or_pagerank(ro, col, V, 0.85, 20, pr);
        double sum = 0;
        for (u32 v = 0; v < V; ++v) sum += pr[v];
        CHECK(l1 >= 0 && sum > 0.999 && sum < 1.001);
        free(k); free(ro); free(col); free(lv); free(pr);
    }
    return 0;
}

int main(int argc, char **argv) {
    const int threads = argc > 1 ? atoi(argv[1]) : 8;
    if (check_apply(threads) || check_pma() || check_analytics()) return 1;
    printf("san oracle ok\n");
    return 0;
}
