// host.h -- host-side runtime pieces: errors, thread pool loops, CSR builder,
// parallel radix sort, counter-based RNG.
#pragma once
#include <algorithm>
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dsssp.h"

namespace dsssp {

// Served form: applies nbuf requests -- pinned host record buffers, already in
// g's vertex ids and validated, with ni[i] insertions, nd[i] deletions and
// zw[i] set when an insertion has weight 0 -- to (g, t) as one batch
// (engine.hip).  With check set, the merged records are first checked on the
// device: only the requests before the first one that inserts a key an earlier
// request deletes are applied (one batch would then leave another graph than
// the requests one by one); *applied = the requests applied, always >= 1.  The
// buffers may be reused once it returns.
int update_pinned(dsssp_graph *g, dsssp_tree *t, const dsssp_change *const *buf, const int64_t *len,
                  const int64_t *ni, const int64_t *nd, const uint8_t *zw, int32_t nbuf, bool check,
                  int32_t *applied, dsssp_stats *stats);
// Page-locked host memory (hipHostMalloc / hipHostFree).
int pinned_alloc(void **p, size_t bytes);
void pinned_free(void *p);

// Records the calling thread's last error and returns code.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
const char *last_error();
void clear_error();

int host_threads();

// Runs f(begin, end, tid) over [0, n) split into contiguous ranges, one per thread.
template <class F>
void parallel_ranges(int64_t n, F f, int threads = 0)
{
    if (threads <= 0) threads = host_threads();
    if (n < 65536 || threads <= 1) {
        if (n > 0) f((int64_t)0, n, 0);
        return;
    }
    threads = (int)std::min<int64_t>(threads, n);
    std::vector<std::thread> pool;
    pool.reserve(threads);
    for (int t = 0; t < threads; ++t) {
        const int64_t b = n * t / threads, e = n * (t + 1) / threads;
        pool.emplace_back([=, &f] { f(b, e, t); });
    }
    for (auto &th : pool) th.join();
}

// Runs f(begin, end) over [0, n) in chunks handed out dynamically (uneven work).
template <class F>
void parallel_chunks(int64_t n, int64_t chunk, F f, int threads = 0)
{
    if (threads <= 0) threads = host_threads();
    if (n <= chunk || threads <= 1) {
        if (n > 0) f((int64_t)0, n);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; ++t)
        pool.emplace_back([&] {
            for (;;) {
                const int64_t b = next.fetch_add(chunk);
                if (b >= n) break;
                f(b, std::min(n, b + chunk));
            }
        });
    for (auto &th : pool) th.join();
}

// splitmix64 finaliser: a counter-based generator, identical on every machine.
// Slots of the edge map for `entries` directed entries: load factor <= 0.7.
inline uint64_t emap_slots_for(uint64_t entries)
{
    uint64_t s = 1024;
    while (s * 7 < entries * 10) s <<= 1;
    return s;
}

inline uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
inline uint64_t rng_at(uint64_t seed, uint64_t stream, uint64_t i)
{
    return mix64(mix64(seed ^ (stream * 0xD6E8FEB86659FD93ull)) + i);
}
// Uniform integer in [0, range) from a 64-bit draw (multiply-high, bias < 2^-32).
inline uint64_t bounded(uint64_t r, uint64_t range)
{
    return (uint64_t)(((__uint128_t)r * range) >> 64);
}

// Parallel LSD radix sort of keys whose significant bits are [0, bits).
void radix_sort_u64(uint64_t *keys, int64_t n, int bits);

// Host CSR of an undirected edge list: row v holds (neighbour << 32 | weight)
// for every edge incident to v (a self-loop twice), rows sorted ascending.
// Validates ids (0 <= id < n) and weights (>= 0).  Returns DSSSP_OK or an error.
int build_csr(int64_t n, int64_t m, const int32_t *a, const int32_t *b, const int32_t *w,
              std::vector<uint64_t> &rp, std::vector<uint64_t> &adj);

}  // namespace dsssp
