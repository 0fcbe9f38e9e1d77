// host.cpp -- errors, host thread count, parallel radix sort, CSR builder.
#include "host.h"

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/dsssp.h"

namespace dsssp {

static thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

const char *last_error() { return g_err.c_str(); }
void clear_error() { g_err.clear(); }

int host_threads()
{
    static int n = [] {
        const char *e = getenv("DSSSP_HOST_THREADS");
        int v = e ? atoi(e) : 0;
        if (v <= 0) v = (int)std::thread::hardware_concurrency();
        return std::max(1, std::min(v, 16));
    }();
    return n;
}

void radix_sort_u64(uint64_t *keys, int64_t n, int bits)
{
    if (n <= 1) return;
    if (n < (1 << 16)) {
        std::sort(keys, keys + n);
        return;
    }
    const int T = host_threads();
    std::vector<uint64_t> tmp((size_t)n);
    uint64_t *src = keys, *dst = tmp.data();
    constexpr int DB = 11, R = 1 << DB;
    std::vector<int64_t> hist((size_t)T * R);
    for (int shift = 0; shift < bits; shift += DB) {
        std::fill(hist.begin(), hist.end(), 0);
        parallel_ranges(n, [&](int64_t b, int64_t e, int t) {
            int64_t *h = &hist[(size_t)t * R];
            for (int64_t i = b; i < e; ++i) h[(src[i] >> shift) & (R - 1)]++;
        }, T);
        // exclusive offsets in (digit, thread) order keep the sort stable
        int64_t run = 0;
        for (int d = 0; d < R; ++d)
            for (int t = 0; t < T; ++t) {
                const int64_t c = hist[(size_t)t * R + d];
                hist[(size_t)t * R + d] = run;
                run += c;
            }
        parallel_ranges(n, [&](int64_t b, int64_t e, int t) {
            int64_t *h = &hist[(size_t)t * R];
            for (int64_t i = b; i < e; ++i) dst[h[(src[i] >> shift) & (R - 1)]++] = src[i];
        }, T);
        std::swap(src, dst);
    }
    if (src != keys) std::memcpy(keys, src, sizeof(uint64_t) * (size_t)n);
}

int build_csr(int64_t n, int64_t m, const int32_t *a, const int32_t *b, const int32_t *w,
              std::vector<uint64_t> &rp, std::vector<uint64_t> &adj)
{
    if (n < 0 || m < 0) return fail(DSSSP_ERR_ARG, "negative n or m");
    if (n > 0x7FFFFFFF) return fail(DSSSP_ERR_ARG, "n must be below 2^31");
    if (m > 0 && (!a || !b || !w)) return fail(DSSSP_ERR_ARG, "null edge arrays");
    std::atomic<int64_t> bad_id{-1}, bad_w{-1};
    std::vector<std::atomic<uint64_t>> deg((size_t)n);
    parallel_ranges(n, [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; ++i) deg[i].store(0, std::memory_order_relaxed);
    });
    parallel_ranges(m, [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; ++i) {
            if (a[i] < 0 || a[i] >= n || b[i] < 0 || b[i] >= n) {
                bad_id.store(i);
                return;
            }
            if (w[i] < 0) {
                bad_w.store(i);
                return;
            }
            deg[a[i]].fetch_add(1, std::memory_order_relaxed);
            deg[b[i]].fetch_add(1, std::memory_order_relaxed);
        }
    });
    if (bad_id.load() >= 0) {
        const int64_t i = bad_id.load();
        return fail(DSSSP_ERR_RANGE, "edge %lld (%d %d) has an endpoint outside [0,%lld)", (long long)i, a[i], b[i],
                    (long long)n);
    }
    if (bad_w.load() >= 0) {
        const int64_t i = bad_w.load();
        return fail(DSSSP_ERR_WEIGHT, "edge %lld (%d %d) has negative weight %d", (long long)i, a[i], b[i], w[i]);
    }
    rp.assign((size_t)n + 1, 0);
    for (int64_t v = 0; v < n; ++v) rp[v + 1] = rp[v] + deg[v].load(std::memory_order_relaxed);
    adj.assign((size_t)rp[n], 0);
    parallel_ranges(n, [&](int64_t lo, int64_t hi, int) {
        for (int64_t v = lo; v < hi; ++v) deg[v].store(rp[v], std::memory_order_relaxed);
    });
    parallel_ranges(m, [&](int64_t lo, int64_t hi, int) {
        for (int64_t i = lo; i < hi; ++i) {
            const uint64_t ww = (uint32_t)w[i];
            adj[deg[a[i]].fetch_add(1, std::memory_order_relaxed)] = ((uint64_t)(uint32_t)b[i] << 32) | ww;
            adj[deg[b[i]].fetch_add(1, std::memory_order_relaxed)] = ((uint64_t)(uint32_t)a[i] << 32) | ww;
        }
    });
    // Sort rows; chunks are handed out dynamically because hub rows cluster.
    parallel_chunks(n, 4096, [&](int64_t lo, int64_t hi) {
        for (int64_t v = lo; v < hi; ++v) std::sort(adj.begin() + (int64_t)rp[v], adj.begin() + (int64_t)rp[v + 1]);
    });
    return DSSSP_OK;
}

}  // namespace dsssp
