// gen.cpp -- seeded synthetic workloads (graphs and change batches).
//
// Every random value comes from rng_at(seed, stream, index) (splitmix64 of a
// counter), so outputs are identical on any machine and thread count.  The
// reference's own generator (createChangedEdges.cpp:49-125) uses unseeded
// rand() and samples deletions with replacement; these generators follow the
// workload definitions instead: deletions without replacement, insertions of
// distinct absent pairs, weights in a given range.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/dsssp.h"
#include "host.h"

namespace dsssp {
namespace {

// Open-addressing set of non-zero uint64 keys (key 0 is reserved as empty).
struct KeySet {
    std::vector<uint64_t> t;
    uint64_t mask = 0;
    explicit KeySet(size_t expect)
    {
        size_t cap = 16;
        while (cap < expect * 2 + 16) cap <<= 1;
        t.assign(cap, 0);
        mask = cap - 1;
    }
    bool insert(uint64_t k)  // true if newly inserted
    {
        size_t h = mix64(k) & mask;
        for (;;) {
            if (t[h] == k) return false;
            if (t[h] == 0) {
                t[h] = k;
                return true;
            }
            h = (h + 1) & mask;
        }
    }
};

inline uint64_t pair_key(uint32_t x, uint32_t y)
{
    const uint32_t lo = x < y ? x : y, hi = x < y ? y : x;
    return ((uint64_t)lo << 32) | hi;
}

int bits_for(uint64_t v)
{
    int b = 1;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

int check_weights(int32_t wlo, int32_t whi)
{
    if (wlo < 0 || whi < wlo) return fail(DSSSP_ERR_WEIGHT, "weight range [%d,%d] invalid", wlo, whi);
    return DSSSP_OK;
}

inline int32_t draw_weight(uint64_t r, int32_t wlo, int32_t whi)
{
    return wlo + (int32_t)bounded(r, (uint64_t)(whi - wlo) + 1);
}

struct DSU {
    std::vector<int32_t> p;
    explicit DSU(int64_t n) : p((size_t)n) { std::iota(p.begin(), p.end(), 0); }
    int32_t find(int32_t x)
    {
        while (p[x] != x) {
            p[x] = p[p[x]];
            x = p[x];
        }
        return x;
    }
    void unite(int32_t a, int32_t b)
    {
        a = find(a), b = find(b);
        if (a != b) p[a < b ? b : a] = a < b ? a : b;
    }
};

}  // namespace
}  // namespace dsssp

using namespace dsssp;

extern "C" int dsssp_gen_tree_plus_random(int64_t n, int64_t m, int32_t wlo, int32_t whi, uint64_t seed, int32_t *a,
                                          int32_t *b, int32_t *w)
{
    clear_error();
    if (n < 1 || n > INT32_MAX || !a || !b || !w) return fail(DSSSP_ERR_ARG, "bad argument");
    if (m < n - 1 || (double)m > (double)n * (double)(n - 1) / 2.0)
        return fail(DSSSP_ERR_ARG, "m=%lld impossible for n=%lld", (long long)m, (long long)n);
    if (int s = check_weights(wlo, whi)) return s;
    KeySet seen((size_t)m);
    for (int64_t i = 1; i < n; ++i) {
        const uint32_t j = (uint32_t)bounded(rng_at(seed, 1, (uint64_t)i), (uint64_t)i);
        a[i - 1] = (int32_t)j;
        b[i - 1] = (int32_t)i;
        seen.insert(pair_key(j, (uint32_t)i) + 1);
    }
    int64_t e = n - 1;
    for (uint64_t d = 0; e < m; ++d) {
        const uint64_t r = rng_at(seed, 2, d);
        const uint32_t x = (uint32_t)bounded(r, (uint64_t)n);
        const uint32_t y = (uint32_t)bounded(mix64(r), (uint64_t)n);
        if (x == y) continue;
        if (!seen.insert(pair_key(x, y) + 1)) continue;
        a[e] = (int32_t)x;
        b[e] = (int32_t)y;
        ++e;
    }
    for (int64_t i = 0; i < m; ++i) w[i] = draw_weight(rng_at(seed, 3, (uint64_t)i), wlo, whi);
    return DSSSP_OK;
}

extern "C" int dsssp_gen_rmat(int32_t scale, int32_t edge_factor, double pa, double pb, double pc, int32_t wlo,
                              int32_t whi, uint64_t seed, int64_t *m, int32_t **a, int32_t **b, int32_t **w)
{
    clear_error();
    if (scale < 1 || scale > 30 || edge_factor < 1 || !m || !a || !b || !w) return fail(DSSSP_ERR_ARG, "bad argument");
    if (pa < 0 || pb < 0 || pc < 0 || pa + pb + pc > 1.0) return fail(DSSSP_ERR_ARG, "bad R-MAT probabilities");
    if (int s = check_weights(wlo, whi)) return s;
    const uint64_t draws = (uint64_t)edge_factor << scale;
    // 16-bit thresholds of the quadrant choice
    const uint32_t ta = (uint32_t)llround(pa * 65536.0);
    const uint32_t tb = (uint32_t)llround((pa + pb) * 65536.0);
    const uint32_t tc = (uint32_t)llround((pa + pb + pc) * 65536.0);
    std::vector<uint64_t> keys((size_t)draws);
    const uint64_t NONE = ~0ull;
    parallel_ranges((int64_t)draws, [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; ++i) {
            uint64_t u = 0, v = 0, r = 0;
            for (int l = 0; l < scale; ++l) {
                if ((l & 3) == 0) r = rng_at(seed, 10 + (uint64_t)(l >> 2), (uint64_t)i);
                const uint32_t q = (uint32_t)(r & 0xFFFF);
                r >>= 16;
                const uint64_t bu = q >= tb, bv = (q >= ta && q < tb) || q >= tc;
                u = (u << 1) | bu;
                v = (v << 1) | bv;
            }
            if (u == v) {
                keys[i] = NONE;
            } else {
                const uint64_t lo2 = u < v ? u : v, hi2 = u < v ? v : u;
                keys[i] = (lo2 << scale) | hi2;
            }
        }
    });
    // Real keys have lo < hi < 2^scale, so their low 2*scale bits are never all
    // ones and NONE (self-loops) sorts last; it is dropped explicitly anyway.
    radix_sort_u64(keys.data(), (int64_t)draws, 2 * scale);
    int64_t cnt = 0;
    for (uint64_t i = 0; i < draws; ++i) {
        const uint64_t k = keys[i];
        if (k == NONE) continue;
        if (cnt > 0 && keys[cnt - 1] == k) continue;
        keys[cnt++] = k;
    }
    int32_t *A = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cnt ? cnt : 1));
    int32_t *B = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cnt ? cnt : 1));
    int32_t *W = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cnt ? cnt : 1));
    if (!A || !B || !W) {
        free(A), free(B), free(W);
        return fail(DSSSP_ERR_NOMEM, "out of host memory");
    }
    const uint64_t vmask = (1ull << scale) - 1;
    parallel_ranges(cnt, [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; ++i) {
            const uint64_t k = keys[i];
            A[i] = (int32_t)(k >> scale);
            B[i] = (int32_t)(k & vmask);
            W[i] = draw_weight(mix64(seed ^ 0x5EEDull) ^ mix64(k), wlo, whi);
        }
    });
    *a = A;
    *b = B;
    *w = W;
    *m = cnt;
    return DSSSP_OK;
}

extern "C" int dsssp_gen_grid(int32_t side, double drop, int32_t wlo, int32_t whi, uint64_t seed, int64_t *m,
                              int32_t **a, int32_t **b, int32_t **w)
{
    clear_error();
    if (side < 2 || side > 46340 || drop < 0 || drop >= 1 || !m || !a || !b || !w)
        return fail(DSSSP_ERR_ARG, "bad argument");
    if (int s = check_weights(wlo, whi)) return s;
    const int64_t S = side, E = 2 * S * (S - 1);
    const int64_t n = S * S;
    // edge e < S*(S-1): horizontal (r, c)-(r, c+1) with e = r*(S-1)+c; else vertical
    auto endpoints = [S](int64_t e, int32_t &x, int32_t &y) {
        const int64_t H = S * (S - 1);
        if (e < H) {
            const int64_t r = e / (S - 1), c = e % (S - 1);
            x = (int32_t)(r * S + c);
            y = x + 1;
        } else {
            const int64_t f = e - H, r = f / S, c = f % S;
            x = (int32_t)(r * S + c);
            y = x + (int32_t)S;
        }
    };
    for (uint64_t attempt = 0; attempt < 64; ++attempt) {
        const uint64_t sd = seed + attempt * 0x9E3779B97F4A7C15ull;
        // remove exactly round(drop*E) edges: those with the smallest random keys
        const int64_t nrem = llround(drop * (double)E);
        std::vector<uint64_t> key((size_t)E);
        parallel_ranges(E, [&](int64_t lo, int64_t hi, int) {
            for (int64_t e = lo; e < hi; ++e) key[e] = rng_at(sd, 20, (uint64_t)e);
        });
        std::vector<uint8_t> keep((size_t)E, 1);
        if (nrem > 0) {
            std::vector<std::pair<uint64_t, int64_t>> kv((size_t)E);
            for (int64_t e = 0; e < E; ++e) kv[e] = {key[e], e};
            std::nth_element(kv.begin(), kv.begin() + (nrem - 1), kv.end());
            for (int64_t i = 0; i < nrem; ++i) keep[kv[i].second] = 0;
        }
        DSU dsu(n);
        for (int64_t e = 0; e < E; ++e) {
            if (!keep[e]) continue;
            int32_t x, y;
            endpoints(e, x, y);
            dsu.unite(x, y);
        }
        std::vector<int64_t> size((size_t)n, 0);
        for (int64_t v = 0; v < n; ++v) size[dsu.find((int32_t)v)]++;
        int64_t best = 0;
        for (int64_t v = 0; v < n; ++v)
            if (size[v] > size[best]) best = v;
        if (dsu.find(0) != best) continue;  // vertex (0,0) outside the largest component: redraw
        std::vector<int32_t> A, B, W;
        for (int64_t e = 0; e < E; ++e) {
            if (!keep[e]) continue;
            int32_t x, y;
            endpoints(e, x, y);
            if (dsu.find(x) != best) continue;
            A.push_back(x);
            B.push_back(y);
            W.push_back(draw_weight(rng_at(sd, 21, (uint64_t)e), wlo, whi));
        }
        const size_t cnt = A.size();
        *a = (int32_t *)malloc(sizeof(int32_t) * (cnt ? cnt : 1));
        *b = (int32_t *)malloc(sizeof(int32_t) * (cnt ? cnt : 1));
        *w = (int32_t *)malloc(sizeof(int32_t) * (cnt ? cnt : 1));
        if (!*a || !*b || !*w) {
            free(*a), free(*b), free(*w);
            return fail(DSSSP_ERR_NOMEM, "out of host memory");
        }
        memcpy(*a, A.data(), sizeof(int32_t) * cnt);
        memcpy(*b, B.data(), sizeof(int32_t) * cnt);
        memcpy(*w, W.data(), sizeof(int32_t) * cnt);
        *m = (int64_t)cnt;
        return DSSSP_OK;
    }
    return fail(DSSSP_ERR_ARG, "vertex 0 never fell in the largest component");
}

extern "C" int dsssp_gen_changes(int64_t n, int64_t m, const int32_t *a, const int32_t *b, const int32_t *w, int64_t k,
                                 double ins_frac, int32_t wlo, int32_t whi, int64_t first_deletable, uint64_t seed,
                                 dsssp_change *out)
{
    clear_error();
    (void)w;
    if (n < 2 || m < 0 || k < 0 || (m > 0 && (!a || !b)) || (k > 0 && !out) || ins_frac < 0 || ins_frac > 1 ||
        first_deletable < 0 || first_deletable > m)
        return fail(DSSSP_ERR_ARG, "bad argument");
    if (int s = check_weights(wlo, whi)) return s;
    const int64_t ni = llround((double)k * ins_frac);
    const int64_t nd = k - ni;
    if (nd > m - first_deletable)
        return fail(DSSSP_ERR_ARG, "%lld deletions requested but only %lld deletable edges", (long long)nd,
                    (long long)(m - first_deletable));
    const double pairs = (double)n * (double)(n - 1) / 2.0;
    if ((double)ni > 0.5 * (pairs - (double)m))
        return fail(DSSSP_ERR_ARG, "too many insertions for the absent pairs left");
    // deletions: distinct indices of [first_deletable, m)
    const int64_t span = m - first_deletable;
    std::vector<uint64_t> taken((size_t)((span + 63) / 64), 0);
    int64_t o = 0;
    for (uint64_t d = 0; o < nd; ++d) {
        const uint64_t r = rng_at(seed, 30, d);
        const int64_t e = (int64_t)bounded(r, (uint64_t)span);
        if (taken[e >> 6] >> (e & 63) & 1) continue;
        taken[e >> 6] |= 1ull << (e & 63);
        const int64_t idx = first_deletable + e;
        const bool swap = (r >> 7) & 1;
        out[o].a = swap ? b[idx] : a[idx];
        out[o].b = swap ? a[idx] : b[idx];
        out[o].w = w ? w[idx] : 0;
        out[o].op = 0;
        ++o;
    }
    // insertions: distinct pairs absent from the graph
    if (ni > 0) {
        std::vector<uint64_t> ek((size_t)m);
        parallel_ranges(m, [&](int64_t lo, int64_t hi, int) {
            for (int64_t i = lo; i < hi; ++i) ek[i] = pair_key((uint32_t)a[i], (uint32_t)b[i]);
        });
        radix_sort_u64(ek.data(), m, 32 + bits_for((uint64_t)n));
        KeySet mine((size_t)ni);
        const int64_t block = std::max<int64_t>(1 << 16, ni / 4);
        std::vector<uint64_t> cand;
        std::vector<uint8_t> absent;
        int64_t got = 0;
        for (uint64_t base = 0; got < ni; base += (uint64_t)block) {
            cand.assign((size_t)block, 0);
            absent.assign((size_t)block, 0);
            parallel_ranges(block, [&](int64_t lo, int64_t hi, int) {
                for (int64_t i = lo; i < hi; ++i) {
                    const uint64_t r = rng_at(seed, 31, base + (uint64_t)i);
                    const uint32_t x = (uint32_t)bounded(r, (uint64_t)n);
                    const uint32_t y = (uint32_t)bounded(mix64(r), (uint64_t)n);
                    if (x == y) continue;
                    const uint64_t key = ((uint64_t)x << 32) | y;  // orientation kept
                    cand[i] = key;
                    absent[i] = !std::binary_search(ek.begin(), ek.end(), pair_key(x, y));
                }
            });
            for (int64_t i = 0; i < block && got < ni; ++i) {
                if (!absent[i]) continue;
                const uint32_t x = (uint32_t)(cand[i] >> 32), y = (uint32_t)cand[i];
                if (!mine.insert(pair_key(x, y) + 1)) continue;
                out[o].a = (int32_t)x;
                out[o].b = (int32_t)y;
                out[o].w = draw_weight(rng_at(seed, 32, (uint64_t)got), wlo, whi);
                out[o].op = 1;
                ++o;
                ++got;
            }
        }
    }
    // interleave: Fisher-Yates shuffle
    for (int64_t i = k - 1; i > 0; --i) {
        const int64_t j = (int64_t)bounded(rng_at(seed, 33, (uint64_t)i), (uint64_t)i + 1);
        std::swap(out[i], out[j]);
    }
    return DSSSP_OK;
}

extern "C" int dsssp_gen_changes_grid(int32_t side, int64_t m, const int32_t *a, const int32_t *b, const int32_t *w,
                                      const int32_t *parent, int64_t k, double del_frac, double tree_frac,
                                      int32_t max_manhattan, int32_t wlo, int32_t whi, uint64_t seed,
                                      dsssp_change *out)
{
    clear_error();
    if (side < 2 || m < 1 || !a || !b || !w || !parent || k < 0 || (k > 0 && !out) || del_frac < 0 ||
        del_frac > 1 || tree_frac < 0 || tree_frac > 1 || max_manhattan < 1)
        return fail(DSSSP_ERR_ARG, "bad argument");
    if (int s = check_weights(wlo, whi)) return s;
    const int64_t n = (int64_t)side * side;
    const int64_t nd = llround((double)k * del_frac);
    const int64_t ntree = llround((double)nd * tree_frac);
    const int64_t ni = k - nd;
    // edges of the current tree: (v, parent v) for v with a parent other than itself
    std::vector<int64_t> tree_e, other_e;
    {
        std::vector<uint64_t> tk;
        for (int64_t v = 0; v < n; ++v)
            if (parent[v] >= 0 && parent[v] != v) tk.push_back(pair_key((uint32_t)v, (uint32_t)parent[v]));
        std::sort(tk.begin(), tk.end());
        for (int64_t i = 0; i < m; ++i) {
            if (std::binary_search(tk.begin(), tk.end(), pair_key((uint32_t)a[i], (uint32_t)b[i])))
                tree_e.push_back(i);
            else
                other_e.push_back(i);
        }
    }
    if (ntree > (int64_t)tree_e.size() || nd - ntree > (int64_t)other_e.size())
        return fail(DSSSP_ERR_ARG, "not enough edges to delete");
    int64_t o = 0;
    auto pick = [&](std::vector<int64_t> &pool, int64_t cnt, uint64_t stream) {
        // partial Fisher-Yates: cnt distinct picks
        for (int64_t i = 0; i < cnt; ++i) {
            const int64_t j = i + (int64_t)bounded(rng_at(seed, stream, (uint64_t)i), (uint64_t)(pool.size() - i));
            std::swap(pool[i], pool[j]);
            const int64_t e = pool[i];
            out[o++] = dsssp_change{a[e], b[e], w[e], 0};
        }
    };
    pick(tree_e, ntree, 40);
    pick(other_e, nd - ntree, 41);
    // insertions: distinct absent pairs at Manhattan distance 1..max_manhattan
    std::vector<uint64_t> ek((size_t)m);
    for (int64_t i = 0; i < m; ++i) ek[i] = pair_key((uint32_t)a[i], (uint32_t)b[i]);
    std::sort(ek.begin(), ek.end());
    KeySet mine((size_t)std::max<int64_t>(ni, 1));
    int64_t got = 0;
    for (uint64_t d = 0; got < ni; ++d) {
        if (d > (uint64_t)ni * 1000 + 100000) return fail(DSSSP_ERR_ARG, "could not find enough absent pairs");
        const uint64_t r = rng_at(seed, 42, d);
        const int64_t v = (int64_t)bounded(r, (uint64_t)n);
        const int64_t r0 = v / side, c0 = v % side;
        const int32_t L = max_manhattan;
        const uint64_t r2 = mix64(r);
        const int64_t dr = (int64_t)bounded(r2, (uint64_t)(2 * L + 1)) - L;
        const int64_t rem = L - (dr < 0 ? -dr : dr);
        const int64_t dc = (int64_t)bounded(mix64(r2), (uint64_t)(2 * rem + 1)) - rem;
        if (dr == 0 && dc == 0) continue;
        const int64_t r1 = r0 + dr, c1 = c0 + dc;
        if (r1 < 0 || r1 >= side || c1 < 0 || c1 >= side) continue;
        const uint32_t x = (uint32_t)v, y = (uint32_t)(r1 * side + c1);
        const uint64_t key = pair_key(x, y);
        if (std::binary_search(ek.begin(), ek.end(), key)) continue;
        if (!mine.insert(key + 1)) continue;
        out[o++] = dsssp_change{(int32_t)x, (int32_t)y, draw_weight(rng_at(seed, 43, (uint64_t)got), wlo, whi), 1};
        ++got;
    }
    for (int64_t i = k - 1; i > 0; --i) {
        const int64_t j = (int64_t)bounded(rng_at(seed, 44, (uint64_t)i), (uint64_t)i + 1);
        std::swap(out[i], out[j]);
    }
    return DSSSP_OK;
}
