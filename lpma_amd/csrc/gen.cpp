// gen.cpp -- input formats: seeded edge generators and the edgelist reader.
//
// The reference reads edgelists (README.md:7-17, test_update_rpma.cu:174-192);
// its datasets are not shipped, so the workloads here are synthetic.  Every
// generator draws from MT19937-64 (Matsumoto & Nishimura; the engine C++11
// names std::mt19937_64) with rejection range reduction, not
// std::uniform_*_distribution, so a seed fixes the output everywhere.
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/lpma.h"

namespace {

typedef uint64_t u64;
typedef uint32_t u32;

struct Mt64 {
    static constexpr int NN = 312, MM = 156;
    static constexpr u64 MATRIX_A = 0xB5026F5AA96619E9ull;
    static constexpr u64 UM = 0xFFFFFFFF80000000ull;
    static constexpr u64 LM = 0x7FFFFFFFull;
    u64 mt[NN];
    int mti;

    explicit Mt64(u64 seed) {
        mt[0] = seed;
        for (int i = 1; i < NN; ++i) mt[i] = 6364136223846793005ull * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (u64)i;
        mti = NN;
    }

    u64 next() {
        if (mti >= NN) {
            int i;
            for (i = 0; i < NN - MM; ++i) {
                u64 x = (mt[i] & UM) | (mt[i + 1] & LM);
                mt[i] = mt[i + MM] ^ (x >> 1) ^ ((x & 1ull) ? MATRIX_A : 0ull);
            }
            for (; i < NN - 1; ++i) {
                u64 x = (mt[i] & UM) | (mt[i + 1] & LM);
                mt[i] = mt[i + (MM - NN)] ^ (x >> 1) ^ ((x & 1ull) ? MATRIX_A : 0ull);
            }
            u64 x = (mt[NN - 1] & UM) | (mt[0] & LM);
            mt[NN - 1] = mt[MM - 1] ^ (x >> 1) ^ ((x & 1ull) ? MATRIX_A : 0ull);
            mti = 0;
        }
        u64 x = mt[mti++];
        x ^= (x >> 29) & 0x5555555555555555ull;
        x ^= (x << 17) & 0x71D67FFFEDA60000ull;
        x ^= (x << 37) & 0xFFF7EEE000000000ull;
        x ^= (x >> 43);
        return x;
    }

    // uniform in [0, n): top ceil(log2 n) bits, rejecting values >= n
    u64 below(u64 n) {
        if (n <= 1) return 0;
        int bits = 64 - __builtin_clzll(n - 1);
        for (;;) {
            u64 x = next() >> (64 - bits);
            if (x < n) return x;
        }
    }

    // uniform in [0, 1) with 53 bits
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// open-addressing set of keys (keys never equal ~0)
struct KeySet {
    std::vector<u64> slots;
    u64 mask;
    explicit KeySet(u64 n) {
        u64 cap = 16;
        while (cap < 2 * n + 16) cap <<= 1;
        slots.assign(cap, ~0ull);
        mask = cap - 1;
    }
    static u64 mix(u64 x) {
        x ^= x >> 30;
        x *= 0xbf58476d1ce4e5b9ull;
        x ^= x >> 27;
        x *= 0x94d049bb133111ebull;
        x ^= x >> 31;
        return x;
    }
    bool insert(u64 k) {
        u64 i = mix(k) & mask;
        while (slots[i] != ~0ull) {
            if (slots[i] == k) return false;
            i = (i + 1) & mask;
        }
        slots[i] = k;
        return true;
    }
};

// one R-MAT edge (Chakrabarti, Zhan & Faloutsos 2004): per level, the
// quadrant (a | b / c | d) picks one bit of src and of dst
u64 rmat_draw(Mt64 &g, u32 scale, double a, double ab, double abc) {
    u64 src = 0, dst = 0;
    for (u32 l = 0; l < scale; ++l) {
        double u = g.unit();
        u64 bs, bd;
        if (u < a) { bs = 0; bd = 0; }
        else if (u < ab) { bs = 0; bd = 1; }
        else if (u < abc) { bs = 1; bd = 0; }
        else { bs = 1; bd = 1; }
        src = (src << 1) | bs;
        dst = (dst << 1) | bd;
    }
    return (src << 32) | dst;
}

bool sorted_contains(const u64 *a, u64 n, u64 x) {
    u64 lo = 0, hi = n;
    while (lo < hi) {
        u64 m = (lo + hi) >> 1;
        if (a[m] < x) lo = m + 1; else hi = m;
    }
    return lo < n && a[lo] == x;
}

}  // namespace

extern "C" int lpma_mt64_fill(uint64_t seed, uint64_t n, uint64_t *out) {
    if (!out && n) return LPMA_ERR_INVALID_ARGUMENT;
    Mt64 g(seed);
    for (u64 i = 0; i < n; ++i) out[i] = g.next();
    return LPMA_OK;
}

extern "C" int lpma_gen_uniform(uint32_t V, uint64_t n, uint64_t seed, uint64_t *out) {
    if ((!out && n) || V < 2) return n ? LPMA_ERR_INVALID_ARGUMENT : LPMA_OK;
    if (n > (u64)V * (V - 1)) return LPMA_ERR_INVALID_ARGUMENT;
    Mt64 g(seed);
    KeySet set(n);
    u64 k = 0;
    while (k < n) {
        u64 src = g.below(V), dst = g.below(V);
        if (src == dst) continue;
        u64 key = (src << 32) | dst;
        if (set.insert(key)) out[k++] = key;
    }
    return LPMA_OK;
}

extern "C" int lpma_gen_rmat(uint32_t scale, uint64_t n, double a, double b, double c, uint64_t seed,
                             uint64_t *out) {
    if ((!out && n) || scale < 1 || scale > 31) return LPMA_ERR_INVALID_ARGUMENT;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return LPMA_ERR_INVALID_ARGUMENT;
    u64 V = 1ull << scale;
    if (n > V * (V - 1)) return LPMA_ERR_INVALID_ARGUMENT;
    Mt64 g(seed);
    KeySet set(n);
    const double ab = a + b, abc = a + b + c;
    u64 k = 0;
    while (k < n) {
        u64 key = rmat_draw(g, scale, a, ab, abc);
        if ((key >> 32) == (key & 0xFFFFFFFFull)) continue;
        if (set.insert(key)) out[k++] = key;
    }
    return LPMA_OK;
}

extern "C" int lpma_gen_rmat_samples(uint32_t scale, uint64_t n, double a, double b, double c, uint64_t seed,
                                     uint64_t *out) {
    if ((!out && n) || scale < 1 || scale > 31) return LPMA_ERR_INVALID_ARGUMENT;
    if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0) return LPMA_ERR_INVALID_ARGUMENT;
    Mt64 g(seed);
    const double ab = a + b, abc = a + b + c;
    u64 k = 0;
    while (k < n) {
        u64 key = rmat_draw(g, scale, a, ab, abc);
        if ((key >> 32) == (key & 0xFFFFFFFFull)) continue;
        out[k++] = key;
    }
    return LPMA_OK;
}

extern "C" int lpma_shuffle(uint64_t *keys, uint64_t n, uint64_t seed) {
    if (!keys && n) return LPMA_ERR_INVALID_ARGUMENT;
    Mt64 g(seed);
    for (u64 i = n; i > 1; --i) {
        u64 j = g.below(i);
        u64 t = keys[i - 1];
        keys[i - 1] = keys[j];
        keys[j] = t;
    }
    return LPMA_OK;
}

extern "C" int lpma_gen_absent(uint32_t V, const uint64_t *present, uint64_t np, uint64_t n, uint64_t seed,
                               uint64_t *out) {
    if ((!out && n) || (!present && np) || V < 2) return n ? LPMA_ERR_INVALID_ARGUMENT : LPMA_OK;
    if (n + np > (u64)V * (V - 1)) return LPMA_ERR_INVALID_ARGUMENT;
    Mt64 g(seed);
    KeySet set(n);
    u64 k = 0;
    while (k < n) {
        u64 src = g.below(V), dst = g.below(V);
        if (src == dst) continue;
        u64 key = (src << 32) | dst;
        if (sorted_contains(present, np, key)) continue;
        if (set.insert(key)) out[k++] = key;
    }
    return LPMA_OK;
}

extern "C" int lpma_read_edgelist(const char *path, uint32_t *V, uint64_t *E, uint64_t max_edges, uint64_t *out) {
    if (!path || !V || !E) return LPMA_ERR_INVALID_ARGUMENT;
    FILE *f = fopen(path, "r");
    if (!f) return LPMA_ERR_INVALID_ARGUMENT;
    unsigned long long v = 0, e = 0;
    if (fscanf(f, "%llu %llu", &v, &e) != 2 || v == 0 || v >= (1ull << 31)) {
        fclose(f);
        return LPMA_ERR_INVALID_ARGUMENT;
    }
    *V = (uint32_t)v;
    u64 want = e;
    if (max_edges && max_edges < want) want = max_edges;
    *E = want;
    if (!out) {
        fclose(f);
        return LPMA_OK;
    }
    for (u64 i = 0; i < want; ++i) {
        unsigned long long s, d;
        if (fscanf(f, "%llu %llu", &s, &d) != 2 || s >= v || d >= v) {
            fclose(f);
            return LPMA_ERR_INVALID_KEY;
        }
        out[i] = ((u64)s << 32) | (u64)d;
    }
    fclose(f);
    return LPMA_OK;
}
