/*
 * oracle.c -- CPU oracle for the batched edge-update path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (lpma_amd/) links, loads
 * or calls this file; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only as the checker / the timed CPU baseline.
 *
 * Two parts:
 *
 *  1. The set-semantics contract (or_apply, or_row_offsets).  After every
 *     batch the live key set must equal sort(unique((S \ D) u I)) and the
 *     row offsets must count keys per source vertex.  This generalises the
 *     reference's update-correctness check, which compares the compacted
 *     structure against a host sort of every inserted key
 *     (test_update_correctness_rpma.cu:47-67, test_update_rpma.cu:131-165),
 *     and its CSR check (test_csr.cu:14-39).  Delete semantics follow the
 *     flat GPMA path (gpma.cuh:162-175): a delete of a present key removes
 *     it, a delete of an absent key and an insert of a present key are
 *     no-ops.  Within one batch deletes apply before inserts.
 *     or_apply is multithreaded (pthreads, key-range partition); the same
 *     code is the reported CPU baseline.
 *
 *  2. A CPU model of the packed memory array in its logical slot space
 *     (or_pma_*).  It restates the reference's per-level orchestration of
 *     update_rpma (rpma.cuh:1549-1718): sort the batch (:1557), route each
 *     key to a segment (:253-316), then walk the levels 0..H and rebalance
 *     every window that holds pending updates and fits the density bounds
 *     (:1611-1657, block_rebalancing_kernel :765-847), spreading the merged
 *     keys evenly with the reference's formula (:676-679, :961-967), moving
 *     row-marker positions into the row offsets (gpma.cuh:340-344), and
 *     resizing when the root does not fit (:1301-1354, :1662-1712).
 *     Deliberate differences from rpma.cuh, shared with the GPU product:
 *       - slots are one flat array in logical (in-order) order with a
 *         power-of-two number of segments, so capacity is S*2^H rather than
 *         32*(2^(H+1)-1), and a resize is a full even re-spread;
 *       - a key routes to the segment whose separator interval holds it
 *         (separator = first key of the segment after each rebalance),
 *         instead of the first/last-key tree walk;
 *       - the batch is deduplicated and resolved against present keys
 *         first (gpma.cuh:162-175), so no key is stored twice (the
 *         reference stores duplicates, rpma.cuh:303-308, :739-748);
 *       - the 'skip' heuristic for middle levels (rpma.cuh:1609-1619) is
 *         not modelled; every level is tried in order.
 *     The GPU product selects the same windows in one pass, so its gapped
 *     slot array, per-segment counts, separators and row offsets must match
 *     this model slot for slot.
 */
#define _GNU_SOURCE
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

typedef uint64_t u64;
typedef uint32_t u32;

#define OR_KEY_NONE 0xFFFFFFFFFFFFFFFFull

/* ------------------------------------------------------------------------- */
/* Sorting: LSD radix sort of u64, 8-bit digits, skipping constant digits.   */
/* ------------------------------------------------------------------------- */

static void radix_sort_u64(u64 *a, size_t n, u64 *tmp) {
    if (n < 2) return;
    u64 *src = a, *dst = tmp;
    for (int pass = 0; pass < 8; ++pass) {
        size_t count[256];
        memset(count, 0, sizeof(count));
        int shift = pass * 8;
        for (size_t i = 0; i < n; ++i) count[(src[i] >> shift) & 0xFF]++;
        int constant = 0;
        for (int d = 0; d < 256; ++d)
            if (count[d] == n) { constant = 1; break; }
        if (constant) continue;
        size_t sum = 0;
        for (int d = 0; d < 256; ++d) { size_t c = count[d]; count[d] = sum; sum += c; }
        for (size_t i = 0; i < n; ++i) dst[count[(src[i] >> shift) & 0xFF]++] = src[i];
        u64 *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, n * sizeof(u64));
}

/* Sort ascending and drop repeats; returns the new length. */
size_t or_sort_unique(u64 *a, size_t n) {
    if (n == 0) return 0;
    u64 *tmp = (u64 *)malloc(n * sizeof(u64));
    radix_sort_u64(a, n, tmp);
    free(tmp);
    size_t m = 1;
    for (size_t i = 1; i < n; ++i)
        if (a[i] != a[m - 1]) a[m++] = a[i];
    return m;
}

static size_t lower_bound_u64(const u64 *a, size_t n, u64 x) {
    size_t lo = 0, hi = n;
    while (lo < hi) {
        size_t mid = lo + (hi - lo) / 2;
        if (a[mid] < x) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* ------------------------------------------------------------------------- */
/* Part 1: the set contract.                                                 */
/* ------------------------------------------------------------------------- */

/* Sequential kernel of or_apply: merge sorted-unique S, I, D restricted to
 * the given index ranges; out = (S \ D) u I.  Returns the output length;
 * out == NULL only counts. */
static size_t merge_range(const u64 *S, size_t s0, size_t s1, const u64 *I, size_t i0, size_t i1,
                          const u64 *D, size_t d0, size_t d1, u64 *out) {
    size_t o = 0, s = s0, i = i0, d = d0;
    while (s < s1 || i < i1) {
        u64 x;
        if (i >= i1 || (s < s1 && S[s] < I[i])) {
            x = S[s++];
            while (d < d1 && D[d] < x) ++d;
            if (d < d1 && D[d] == x) continue; /* deleted */
        } else if (s >= s1 || I[i] < S[s]) {
            x = I[i++]; /* inserted (absent from S) */
        } else {
            x = S[s++]; ++i; /* present and re-inserted: delete-then-insert keeps it */
        }
        if (out) out[o] = x;
        ++o;
    }
    return o;
}

typedef struct {
    const u64 *S, *I, *D;
    size_t s0, s1, i0, i1, d0, d1;
    u64 *out;  /* NULL: count only */
    size_t n;
} merge_job;

static void *merge_worker(void *p) {
    merge_job *j = (merge_job *)p;
    j->n = merge_range(j->S, j->s0, j->s1, j->I, j->i0, j->i1, j->D, j->d0, j->d1, j->out);
    return NULL;
}

static void run_jobs(merge_job *jobs, int T) {
    pthread_t *tid = (pthread_t *)calloc(T, sizeof(pthread_t));
    for (int t = 0; t < T; ++t) pthread_create(&tid[t], NULL, merge_worker, &jobs[t]);
    for (int t = 0; t < T; ++t) pthread_join(tid[t], NULL);
    free(tid);
}

/*
 * or_apply: out = sort(unique((S \ D) u I)).  S must be sorted and unique
 * (the previous output); I and D are arbitrary and are sorted/deduplicated
 * in private copies.  `out` must hold ns + ni keys.  threads <= 1 runs the
 * sequential merge; otherwise the key range is split at equally spaced
 * keys of S and each thread merges one range.  Returns the output length.
 */
size_t or_apply(const u64 *S, size_t ns, const u64 *I_in, size_t ni, const u64 *D_in, size_t nd,
                u64 *out, int threads) {
    u64 *I = (u64 *)malloc((ni ? ni : 1) * sizeof(u64));
    u64 *D = (u64 *)malloc((nd ? nd : 1) * sizeof(u64));
    if (ni) memcpy(I, I_in, ni * sizeof(u64));
    if (nd) memcpy(D, D_in, nd * sizeof(u64));
    ni = or_sort_unique(I, ni);
    nd = or_sort_unique(D, nd);
    size_t total;
    if (threads <= 1 || ns < (size_t)threads * 1024) {
        total = merge_range(S, 0, ns, I, 0, ni, D, 0, nd, out);
    } else {
        /* key-range partition at equally spaced keys of S; each thread counts
         * its range's output, then merges straight into its place in out */
        int T = threads;
        merge_job *jobs = (merge_job *)calloc(T, sizeof(merge_job));
        size_t prev_s = 0, prev_i = 0, prev_d = 0;
        for (int t = 0; t < T; ++t) {
            size_t s1, i1, d1;
            if (t == T - 1) {
                s1 = ns; i1 = ni; d1 = nd;
            } else {
                u64 split = S[(ns / T) * (t + 1)];
                s1 = lower_bound_u64(S, ns, split);
                i1 = lower_bound_u64(I, ni, split);
                d1 = lower_bound_u64(D, nd, split);
            }
            jobs[t].S = S; jobs[t].I = I; jobs[t].D = D;
            jobs[t].s0 = prev_s; jobs[t].s1 = s1;
            jobs[t].i0 = prev_i; jobs[t].i1 = i1;
            jobs[t].d0 = prev_d; jobs[t].d1 = d1;
            jobs[t].out = NULL;
            prev_s = s1; prev_i = i1; prev_d = d1;
        }
        run_jobs(jobs, T);
        total = 0;
        for (int t = 0; t < T; ++t) {
            jobs[t].out = out + total;
            total += jobs[t].n;
        }
        run_jobs(jobs, T);
        free(jobs);
    }
    free(I);
    free(D);
    return total;
}

/*
 * Standard CSR row offsets over a sorted key list (src<<32|dst):
 * ro[v] = number of keys whose src < v, for v in [0, V].  Keys whose low
 * 32 bits are all ones (row markers) are not edges and are skipped.
 * Returns 0, or -1 if a key's src is >= V.
 */
int or_row_offsets(const u64 *keys, size_t n, u32 V, u32 *ro) {
    memset(ro, 0, (size_t)(V + 1) * sizeof(u32));
    for (size_t i = 0; i < n; ++i) {
        u64 k = keys[i];
        if ((u32)k == 0xFFFFFFFFu) continue;
        u32 src = (u32)(k >> 32);
        if (src >= V) return -1;
        ro[src + 1]++;
    }
    for (u32 v = 0; v < V; ++v) ro[v + 1] += ro[v];
    return 0;
}

typedef struct {
    const u64 *keys;
    size_t n;
    u32 v0, v1;
    u32 *ro;
} ro_job;

/* ro[v + 1] = edges of row v for v in [v0, v1) (rows are disjoint per job) */
static void *ro_worker(void *p) {
    ro_job *j = (ro_job *)p;
    size_t i = lower_bound_u64(j->keys, j->n, (u64)j->v0 << 32);
    const size_t e = lower_bound_u64(j->keys, j->n, (u64)j->v1 << 32);
    for (u32 v = j->v0; v < j->v1; ++v) j->ro[v + 1] = 0;
    for (; i < e; ++i) {
        u64 k = j->keys[i];
        if ((u32)k == 0xFFFFFFFFu) continue;
        j->ro[(u32)(k >> 32) + 1]++;
    }
    return NULL;
}

/* or_row_offsets on `threads` threads: rows split into equal vertex ranges,
 * each counted by one thread (binary search for its key range), then one
 * prefix sum.  Same result and error as or_row_offsets. */
int or_row_offsets_mt(const u64 *keys, size_t n, u32 V, u32 *ro, int threads) {
    if (threads <= 1 || V < (u32)threads) return or_row_offsets(keys, n, V, ro);
    size_t m = n; /* sorted: the last edge key (markers skipped) has the largest src */
    while (m && (u32)keys[m - 1] == 0xFFFFFFFFu) --m;
    if (m && (u32)(keys[m - 1] >> 32) >= V) return -1;
    int T = threads;
    ro_job *jobs = (ro_job *)calloc(T, sizeof(ro_job));
    pthread_t *tid = (pthread_t *)calloc(T, sizeof(pthread_t));
    for (int t = 0; t < T; ++t) {
        jobs[t].keys = keys;
        jobs[t].n = n;
        jobs[t].v0 = (u32)((u64)V * t / T);
        jobs[t].v1 = (u32)((u64)V * (t + 1) / T);
        jobs[t].ro = ro;
        pthread_create(&tid[t], NULL, ro_worker, &jobs[t]);
    }
    for (int t = 0; t < T; ++t) pthread_join(tid[t], NULL);
    free(jobs);
    free(tid);
    ro[0] = 0;
    for (u32 v = 0; v < V; ++v) ro[v + 1] += ro[v];
    return 0;
}

/* Returns 1 when keys are strictly increasing. */
int or_is_sorted_unique(const u64 *keys, size_t n) {
    for (size_t i = 1; i < n; ++i)
        if (keys[i] <= keys[i - 1]) return 0;
    return 1;
}

/* ------------------------------------------------------------------------- */
/* Part 2: the packed-memory-array model in logical slot space.              */
/* ------------------------------------------------------------------------- */

typedef struct or_pma {
    u32 V;
    u32 log2S, S;
    u32 H;          /* nseg = 1 << H */
    u64 cap;        /* S << H */
    u64 *keys;      /* cap slots, OR_KEY_NONE gaps, each segment front-packed */
    u32 *cnt;       /* live keys per segment */
    u64 *sep;       /* separator (smallest key routed) per segment, sep[0] = 0 */
    u32 *row_off;   /* V+1: row_off[r+1] = slot of row r's marker */
    u64 live;
    double lo_leaf, lo_root, up_leaf, up_root;
    u64 lower[64], upper[64];
    /* statistics */
    u64 windows_per_level[64];
    u64 rebuilds;
} or_pma;

static u64 marker_key(u32 r) { return ((u64)r << 32) | 0xFFFFFFFFull; }
static int is_marker(u64 k) { return k != OR_KEY_NONE && (u32)k == 0xFFFFFFFFu; }

/* Density table, rpma.cuh:136-168.  With lo_leaf == lo_root == 0.42 and
 * up_leaf == up_root == 1 this is the reference's active code
 * (lower = (u32)(0.42 * len), upper = len, :164-165); otherwise the bounds are
 * interpolated linearly from leaf to root as in its commented-out variant. */
static void pma_density(or_pma *p) {
    for (u32 l = 0; l < 64; ++l) {
        double len = (double)((u64)p->S << (l < 40 ? l : 40));
        double t = (p->H == 0 || l >= p->H) ? 0.0 : (double)(p->H - l) / (double)p->H;
        double lo = p->lo_root + (p->lo_leaf - p->lo_root) * t;
        double up = p->up_root + (p->up_leaf - p->up_root) * t;
        p->lower[l] = (u64)(lo * len);
        p->upper[l] = (u64)(up * len);
    }
}

/* Smallest height whose root window holds m keys under the root upper bound. */
static u32 pma_fit_height(const or_pma *p, u64 m) {
    u32 h = 0;
    while ((u64)(p->up_root * (double)((u64)p->S << h)) < m) ++h;
    return h;
}

static int pma_root_fits(const or_pma *p, u64 m) {
    if (m > p->upper[p->H]) return 0;
    if (m >= p->lower[p->H]) return 1;
    return pma_fit_height(p, m) == p->H; /* cannot shrink further */
}

static int pma_alloc(or_pma *p, u32 H) {
    u64 nseg = (u64)1 << H;
    p->H = H;
    p->cap = (u64)p->S << H;
    p->keys = (u64 *)malloc(p->cap * sizeof(u64));
    p->cnt = (u32 *)calloc(nseg, sizeof(u32));
    p->sep = (u64 *)calloc(nseg, sizeof(u64));
    if (!p->keys || !p->cnt || !p->sep) return -1;
    for (u64 i = 0; i < p->cap; ++i) p->keys[i] = OR_KEY_NONE;
    pma_density(p);
    return 0;
}

/* Even spread of m sorted keys over the window of nsegw segments starting
 * at segment a (rpma.cuh:676-679): key i goes to segment
 * floor(nsegw*i/m), lane (nsegw*i mod m)/nsegw, i.e. segment j receives the
 * keys [ceil(j*m/nsegw), ceil((j+1)*m/nsegw)) front-packed.  Updates counts,
 * separators and marker row offsets; sep[a] is kept. */
static void pma_spread(or_pma *p, u64 a, u64 nsegw, const u64 *merged, u64 m) {
    u64 base = a << p->log2S;
    u64 sep_a = p->sep[a];
    for (u64 t = 0; t < (nsegw << p->log2S); ++t) p->keys[base + t] = OR_KEY_NONE;
    for (u64 i = 0; i < m; ++i) {
        u64 seg = (nsegw * i) / m;
        u64 lane = ((nsegw * i) % m) / nsegw;
        u64 slot = base + (seg << p->log2S) + lane;
        p->keys[slot] = merged[i];
        if (is_marker(merged[i])) p->row_off[(u32)(merged[i] >> 32) + 1] = (u32)slot;
    }
    for (u64 j = 0; j < nsegw; ++j) {
        u64 f = (j * m + nsegw - 1) / nsegw;
        u64 g = ((j + 1) * m + nsegw - 1) / nsegw;
        p->cnt[a + j] = (u32)(g - f);
        if (j == 0) continue;
        if (g > f) p->sep[a + j] = merged[f];
        else if (f > 0) p->sep[a + j] = merged[f - 1] + 1;
        else p->sep[a + j] = sep_a;
    }
}

/* Live keys of segments [a, a+nsegw) in order. */
static u64 pma_gather(const or_pma *p, u64 a, u64 nsegw, u64 *out) {
    u64 o = 0;
    for (u64 s = a; s < a + nsegw; ++s) {
        u64 b = s << p->log2S;
        for (u32 l = 0; l < p->cnt[s]; ++l) out[o++] = p->keys[b + l];
    }
    return o;
}

/* Route a key to its segment: the last s with sep[s] <= key. */
static u64 pma_locate(const or_pma *p, u64 key) {
    u64 lo = 0, hi = (u64)1 << p->H; /* find first s with sep[s] > key */
    while (lo < hi) {
        u64 mid = (lo + hi) / 2;
        if (p->sep[mid] <= key) lo = mid + 1; else hi = mid;
    }
    return lo - 1;
}

static int pma_present(const or_pma *p, u64 s, u64 key) {
    const u64 *seg = p->keys + (s << p->log2S);
    size_t i = lower_bound_u64(seg, p->cnt[s], key);
    return i < p->cnt[s] && seg[i] == key;
}

/* merge old live keys with sorted effective updates (op 1 insert / 0 delete) */
static u64 pma_merge(const u64 *old, u64 c, const u64 *uk, const unsigned char *uop, u64 k, u64 *out) {
    u64 o = 0, i = 0, q = 0;
    while (i < c || q < k) {
        if (q >= k || (i < c && old[i] < uk[q])) {
            out[o++] = old[i++];
        } else if (i >= c || uk[q] < old[i]) {
            if (uop[q]) out[o++] = uk[q];
            ++q;
        } else { /* equal: an effective update on a present key is a delete */
            ++i; ++q;
        }
    }
    return o;
}

/* Full re-spread into height H' with all live keys and the given effective
 * updates (rebuild_rpma, rpma.cuh:1417-1472). */
static int pma_rebuild(or_pma *p, u32 newH, const u64 *uk, const unsigned char *uop, u64 k) {
    u64 nseg_old = (u64)1 << p->H;
    u64 *old = (u64 *)malloc((p->live + 1) * sizeof(u64));
    u64 c = pma_gather(p, 0, nseg_old, old);
    u64 *merged = (u64 *)malloc((c + k + 1) * sizeof(u64));
    u64 m = pma_merge(old, c, uk, uop, k, merged);
    free(old);
    free(p->keys); free(p->cnt); free(p->sep);
    if (pma_alloc(p, newH)) { free(merged); return -1; }
    p->sep[0] = 0;
    pma_spread(p, 0, (u64)1 << newH, merged, m);
    p->live = m;
    p->rebuilds++;
    free(merged);
    return 0;
}

/*
 * One batch: `keys` are full 64-bit keys, `ops` 1 = insert / 0 = delete.
 * Returns 0, or -1 on an invalid key (src or dst >= V) with no change.
 */
static int pma_apply_batch(or_pma *p, const u64 *keys, const unsigned char *ops, size_t n, int internal) {
    if (n == 0) return 0;
    /* 1. sort by (key, op) and keep the last of equal keys: delete-then-insert */
    u64 *pk = (u64 *)malloc(n * sizeof(u64));
    for (size_t i = 0; i < n; ++i) {
        u64 k = keys[i];
        u32 src = (u32)(k >> 32), dst = (u32)k;
        if (!internal && (src >= p->V || dst >= p->V)) { free(pk); return -1; }
        pk[i] = k; /* full keys: sort pairs (key, op) via index below */
    }
    /* sort indices by (key, op): stable split by op (deletes first), then a
     * stable LSD radix pass over the key bytes */
    size_t *idx = (size_t *)malloc(n * sizeof(size_t));
    {
        size_t *t2 = (size_t *)malloc(n * sizeof(size_t));
        size_t *t3 = (size_t *)malloc(n * sizeof(size_t));
        size_t z = 0;
        for (size_t i = 0; i < n; ++i) if (!ops[i]) t2[z++] = i;
        for (size_t i = 0; i < n; ++i) if (ops[i]) t2[z++] = i;
        for (int pass = 0; pass < 8; ++pass) {
            size_t count[256] = {0};
            int shift = pass * 8;
            for (size_t i = 0; i < n; ++i) count[(pk[t2[i]] >> shift) & 0xFF]++;
            size_t sum = 0;
            for (int d = 0; d < 256; ++d) { size_t c = count[d]; count[d] = sum; sum += c; }
            for (size_t i = 0; i < n; ++i) t3[count[(pk[t2[i]] >> shift) & 0xFF]++] = t2[i];
            size_t *sw = t2; t2 = t3; t3 = sw;
        }
        memcpy(idx, t2, n * sizeof(size_t));
        free(t2); free(t3);
    }
    /* 2-3. dedup (keep last), route, resolve against present keys */
    u64 *uk = (u64 *)malloc(n * sizeof(u64));
    unsigned char *uop = (unsigned char *)malloc(n);
    u64 *useg = (u64 *)malloc(n * sizeof(u64));
    u64 k = 0;
    for (size_t j = 0; j < n; ++j) {
        u64 key = pk[idx[j]];
        if (j + 1 < n && pk[idx[j + 1]] == key) continue;
        int op = ops[idx[j]];
        u64 s = pma_locate(p, key);
        int present = pma_present(p, s, key);
        if (op == present) continue; /* insert-present / delete-absent */
        uk[k] = key; uop[k] = (unsigned char)op; useg[k] = s; ++k;
    }
    free(pk); free(idx);
    if (k == 0) { free(uk); free(uop); free(useg); return 0; }

    /* 4. level loop (rpma.cuh:1611-1657): windows with pending updates are
     * rebalanced at the lowest level whose bounds they fit. */
    unsigned char *done = (unsigned char *)calloc(k, 1);
    u64 pending = k;
    u64 *old = (u64 *)malloc(((u64)p->S << p->H) * sizeof(u64) + 8);
    u64 *merged = (u64 *)malloc((((u64)p->S << p->H) + k) * sizeof(u64) + 8);
    u64 *wk = (u64 *)malloc(k * sizeof(u64));
    unsigned char *wop = (unsigned char *)malloc(k);
    int need_resize = 0;
    for (u32 level = 0; level <= p->H && pending; ++level) {
        u64 q = 0;
        while (q < k) {
            if (done[q]) { ++q; continue; }
            u64 w = useg[q] >> level;
            u64 a = w << level, nsegw = (u64)1 << level;
            /* all updates (pending or applied) in the window */
            u64 q0 = q, q1 = q;
            while (q0 > 0 && (useg[q0 - 1] >> level) == w) --q0;
            while (q1 < k && (useg[q1] >> level) == w) ++q1;
            u64 live = 0;
            for (u64 s = a; s < a + nsegw; ++s) live += p->cnt[s];
            /* count = live (including updates applied at lower levels) + pending */
            int64_t delta = 0;
            for (u64 r = q0; r < q1; ++r)
                if (!done[r]) delta += uop[r] ? 1 : -1;
            u64 m = (u64)((int64_t)live + delta);
            int fits;
            if (level == p->H) fits = pma_root_fits(p, m);
            else fits = (m >= p->lower[level] && m <= p->upper[level]);
            if (fits) {
                u64 c = pma_gather(p, a, nsegw, old);
                u64 kk = 0;
                for (u64 r = q0; r < q1; ++r)
                    if (!done[r]) { wk[kk] = uk[r]; wop[kk] = uop[r]; ++kk; done[r] = 1; --pending; }
                u64 mm = pma_merge(old, c, wk, wop, kk, merged);
                pma_spread(p, a, nsegw, merged, mm);
                p->live = p->live - c + mm;
                p->windows_per_level[level]++;
            } else if (level == p->H) {
                need_resize = 1;
            }
            q = q1;
        }
    }
    if (pending) {
        /* root does not fit: resize to the smallest fitting height and
         * re-spread everything together with the pending updates */
        u64 kk = 0;
        for (u64 r = 0; r < k; ++r)
            if (!done[r]) { wk[kk] = uk[r]; wop[kk] = uop[r]; ++kk; }
        int64_t delta = 0;
        for (u64 r = 0; r < kk; ++r) delta += wop[r] ? 1 : -1;
        u64 m = (u64)((int64_t)p->live + delta);
        (void)need_resize;
        pma_rebuild(p, pma_fit_height(p, m), wk, wop, kk);
    }
    free(done); free(old); free(merged); free(wk); free(wop);
    free(uk); free(uop); free(useg);
    return 0;
}

or_pma *or_pma_create(u32 V, u32 seg_slots, double lo_leaf, double lo_root, double up_leaf, double up_root) {
    if (seg_slots < 2 || (seg_slots & (seg_slots - 1))) return NULL;
    or_pma *p = (or_pma *)calloc(1, sizeof(or_pma));
    p->V = V;
    p->S = seg_slots;
    p->log2S = (u32)__builtin_ctz(seg_slots);
    p->lo_leaf = lo_leaf; p->lo_root = lo_root; p->up_leaf = up_leaf; p->up_root = up_root;
    p->row_off = (u32 *)calloc((size_t)V + 1, sizeof(u32));
    if (pma_alloc(p, 0)) return NULL;
    p->sep[0] = 0;
    p->live = 0;
    /* one marker per row, inserted through the update path (init_csr_rpma,
     * rpma.cuh:1807-1824; marker convention gpma.cuh:834-856) */
    if (V) {
        u64 *mk = (u64 *)malloc((size_t)V * sizeof(u64));
        unsigned char *op = (unsigned char *)malloc(V);
        for (u32 r = 0; r < V; ++r) { mk[r] = marker_key(r); op[r] = 1; }
        pma_apply_batch(p, mk, op, V, 1);
        free(mk); free(op);
    }
    return p;
}

void or_pma_destroy(or_pma *p) {
    if (!p) return;
    free(p->keys); free(p->cnt); free(p->sep); free(p->row_off);
    free(p);
}

int or_pma_update(or_pma *p, const u64 *keys, size_t n, int op) {
    unsigned char *ops = (unsigned char *)malloc(n ? n : 1);
    memset(ops, op ? 1 : 0, n);
    int r = pma_apply_batch(p, keys, ops, n, 0);
    free(ops);
    return r;
}

int or_pma_update_mixed(or_pma *p, const u64 *ins, size_t ni, const u64 *del, size_t nd) {
    size_t n = ni + nd;
    u64 *k = (u64 *)malloc((n ? n : 1) * sizeof(u64));
    unsigned char *ops = (unsigned char *)malloc(n ? n : 1);
    for (size_t i = 0; i < nd; ++i) { k[i] = del[i]; ops[i] = 0; }
    for (size_t i = 0; i < ni; ++i) { k[nd + i] = ins[i]; ops[nd + i] = 1; }
    int r = pma_apply_batch(p, k, ops, n, 0);
    free(k); free(ops);
    return r;
}

u64 or_pma_capacity(const or_pma *p) { return p->cap; }
u32 or_pma_height(const or_pma *p) { return p->H; }
u64 or_pma_live(const or_pma *p) { return p->live; }
const u64 *or_pma_keys(const or_pma *p) { return p->keys; }
const u32 *or_pma_cnt(const or_pma *p) { return p->cnt; }
const u64 *or_pma_sep(const or_pma *p) { return p->sep; }
const u32 *or_pma_row_offsets(const or_pma *p) { return p->row_off; }
u64 or_pma_rebuilds(const or_pma *p) { return p->rebuilds; }
u64 or_pma_windows_at_level(const or_pma *p, u32 l) { return l < 64 ? p->windows_per_level[l] : 0; }

/* Compacted live keys (including row markers), like get_data_rpma
 * (rpma.cuh:1735-1761).  Returns the count. */
u64 or_pma_export(const or_pma *p, u64 *out) { return pma_gather(p, 0, (u64)1 << p->H, out); }

/* ------------------------------------------------------------------------- */
/* Analytics oracles over a standard CSR (ro[V+1], col = dst ids).            */
/* ------------------------------------------------------------------------- */

/* BFS levels from src (0xFFFFFFFF = unreached), level-synchronous queue;
 * the reference calls nvGraph's BFS on a CSR (test_algos.cu:136-186,
 * include/bfs.hxx).  Levels do not depend on the traversal order.  Returns the
 * number of levels (depth + 1). */
u32 or_bfs(const u32 *ro, const u32 *col, u32 V, u32 src, u32 *levels) {
    for (u32 v = 0; v < V; ++v) levels[v] = 0xFFFFFFFFu;
    if (src >= V) return 0;
    u32 *q = (u32 *)malloc((size_t)V * sizeof(u32));
    u32 head = 0, tail = 0;
    q[tail++] = src;
    levels[src] = 0;
    u32 depth = 0;
    while (head < tail) {
        u32 u = q[head++];
        for (u32 e = ro[u]; e < ro[u + 1]; ++e) {
            u32 v = col[e];
            if (levels[v] == 0xFFFFFFFFu) {
                levels[v] = levels[u] + 1;
                if (levels[v] > depth) depth = levels[v];
                q[tail++] = v;
            }
        }
    }
    free(q);
    return depth + 1;
}

/* PageRank in FP64 with nvGraph's formulation (include/pagerank.hxx:66-88):
 * pr' = alpha * A^T (pr / outdeg) + (alpha * dangling(pr) + (1-alpha) * sum(pr)) / V,
 * from the uniform vector, exactly `iters` iterations, then divided by its L1
 * norm.  Returns the L1 change of the last iteration. */
double or_pagerank(const u32 *ro, const u32 *col, u32 V, double alpha, u32 iters, double *pr) {
    double *acc = (double *)malloc((size_t)V * sizeof(double));
    for (u32 v = 0; v < V; ++v) pr[v] = 1.0 / (double)V;
    double delta = 0.0;
    for (u32 it = 0; it < iters; ++it) {
        double dang = 0.0, sum = 0.0;
        for (u32 v = 0; v < V; ++v) acc[v] = 0.0;
        for (u32 u = 0; u < V; ++u) {
            u32 od = ro[u + 1] - ro[u];
            sum += pr[u];
            if (!od) { dang += pr[u]; continue; }
            double c = pr[u] / (double)od;
            for (u32 e = ro[u]; e < ro[u + 1]; ++e) acc[col[e]] += c;
        }
        double gamma = (alpha * dang + (1.0 - alpha) * sum) / (double)V;
        delta = 0.0;
        for (u32 v = 0; v < V; ++v) {
            double p = alpha * acc[v] + gamma;
            delta += fabs(p - pr[v]);
            pr[v] = p;
        }
    }
    double l1 = 0.0;
    for (u32 v = 0; v < V; ++v) l1 += pr[v];
    for (u32 v = 0; v < V; ++v) pr[v] /= l1;
    free(acc);
    return delta;
}
