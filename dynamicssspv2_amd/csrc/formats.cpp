// formats.cpp -- the reference's three text formats (README.md:22-37), read in
// parallel and written with exact integers.
//
//   graph   "a b w"          read_graphEdges   all_structure_undir.cuh:178-198
//   changes "a b w inst"     readin_changes    all_structure_undir.cuh:89-126
//   SSSP    "v parent dist"  read_SSSP         all_structure_undir.cuh:133-171
//
// Fields are separated by any run of blanks; blank lines are skipped; a final
// line without newline is accepted (the reference's own graph.txt ends that
// way).  Extra trailing fields are ignored, as sscanf ignores them.
#include <cerrno>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dsssp.h"
#include "host.h"

namespace dsssp {
namespace {

bool read_file(const char *path, std::vector<char> &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    if (fseek(f, 0, SEEK_END) != 0) {
        fclose(f);
        return false;
    }
    const long sz = ftell(f);
    if (sz < 0) {
        fclose(f);
        return false;
    }
    rewind(f);
    buf.resize((size_t)sz + 1);
    const size_t got = sz > 0 ? fread(buf.data(), 1, (size_t)sz, f) : 0;
    fclose(f);
    if (got != (size_t)sz) return false;
    buf[(size_t)sz] = '\n';
    return true;
}

inline bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// Parses one signed decimal integer token; returns false if the token is not one.
inline bool parse_int(const char *&p, const char *end, int64_t &out)
{
    while (p < end && is_blank(*p)) ++p;
    if (p >= end || *p == '\n') return false;
    bool neg = false;
    if (*p == '-' || *p == '+') {
        neg = *p == '-';
        ++p;
    }
    if (p >= end || *p < '0' || *p > '9') return false;
    int64_t v = 0;
    while (p < end && *p >= '0' && *p <= '9') {
        if (v > (INT64_MAX - 9) / 10) return false;
        v = v * 10 + (*p - '0');
        ++p;
    }
    if (p < end && !is_blank(*p) && *p != '\n') return false;  // e.g. "1.5e+06"
    out = neg ? -v : v;
    return true;
}

// Skips one token of any content.
inline bool skip_token(const char *&p, const char *end)
{
    while (p < end && is_blank(*p)) ++p;
    if (p >= end || *p == '\n') return false;
    while (p < end && !is_blank(*p) && *p != '\n') ++p;
    return true;
}

inline const char *next_line(const char *p, const char *end)
{
    const char *q = (const char *)memchr(p, '\n', (size_t)(end - p));
    return q ? q + 1 : end;
}

inline bool line_is_blank(const char *p, const char *end)
{
    while (p < end && is_blank(*p)) ++p;
    return p >= end || *p == '\n';
}

// Splits buf into per-thread ranges on line boundaries and parses each range
// with fn(begin, end, out_vector, first_line_number_unknown).  Records are
// int64 tuples of width W.
template <int W, class Fn>
int parse_parallel(const std::vector<char> &buf, std::vector<int64_t> &rec, Fn fn)
{
    const char *base = buf.data();
    const int64_t len = (int64_t)buf.size();
    const int T = len > (1 << 22) ? host_threads() : 1;
    std::vector<int64_t> cut(T + 1);
    cut[0] = 0;
    cut[T] = len;
    for (int t = 1; t < T; ++t) {
        int64_t c = len * t / T;
        if (c < cut[t - 1]) c = cut[t - 1];
        const char *q = (const char *)memchr(base + c, '\n', (size_t)(len - c));
        cut[t] = q ? (int64_t)(q - base) + 1 : len;
    }
    std::vector<std::vector<int64_t>> part(T);
    std::vector<int> status(T, DSSSP_OK);
    std::vector<int64_t> bad_off(T, -1);
    std::vector<std::string> msg(T);  // fail() messages are thread-local
    std::vector<std::thread> pool;
    for (int t = 0; t < T; ++t)
        pool.emplace_back([&, t] {
            const char *p = base + cut[t], *end = base + cut[t + 1];
            while (p < end) {
                const char *nl = next_line(p, end);
                if (!line_is_blank(p, nl)) {
                    int64_t v[W];
                    const char *q = p;
                    const int st = fn(q, nl, v);
                    if (st != DSSSP_OK) {
                        status[t] = st;
                        bad_off[t] = p - base;
                        msg[t] = last_error();
                        return;
                    }
                    part[t].insert(part[t].end(), v, v + W);
                }
                p = nl;
            }
        });
    for (auto &th : pool) th.join();
    for (int t = 0; t < T; ++t) {
        if (status[t] != DSSSP_OK) {
            int64_t line = 1;
            for (int64_t i = 0; i < bad_off[t]; ++i) line += base[i] == '\n';
            return fail(status[t], "line %" PRId64 ": %s", line, msg[t].c_str());
        }
    }
    size_t total = 0;
    for (auto &v : part) total += v.size();
    rec.clear();
    rec.reserve(total);
    for (auto &v : part) rec.insert(rec.end(), v.begin(), v.end());
    return DSSSP_OK;
}

int check_id(int64_t v, int64_t n)
{
    if (v < 0 || v > INT32_MAX || (n > 0 && v >= n))
        return fail(DSSSP_ERR_RANGE, "vertex %" PRId64 " outside [0,%" PRId64 ")", v, n);
    return DSSSP_OK;
}

template <class T>
T *dup_array(const std::vector<T> &v)
{
    T *p = (T *)malloc(sizeof(T) * (v.size() ? v.size() : 1));
    if (p && !v.empty()) memcpy(p, v.data(), sizeof(T) * v.size());
    return p;
}

}  // namespace
}  // namespace dsssp

using namespace dsssp;

extern "C" int dsssp_read_graph(const char *path, int64_t n, int64_t *m, int32_t **a, int32_t **b, int32_t **w)
{
    clear_error();
    if (!path || !m || !a || !b || !w) return fail(DSSSP_ERR_ARG, "null argument");
    std::vector<char> buf;
    if (!read_file(path, buf)) return fail(DSSSP_ERR_IO, "cannot read %s", path);
    std::vector<int64_t> rec;
    int st = parse_parallel<3>(buf, rec, [n](const char *&p, const char *e, int64_t *v) {
        for (int i = 0; i < 3; ++i)
            if (!parse_int(p, e, v[i])) return fail(DSSSP_ERR_FORMAT, "expected 'a b w' integers");
        int s = check_id(v[0], n);
        if (s == DSSSP_OK) s = check_id(v[1], n);
        if (s != DSSSP_OK) return s;
        if (v[2] < 0 || v[2] > INT32_MAX) return fail(DSSSP_ERR_WEIGHT, "weight %" PRId64 " outside [0, 2^31)", v[2]);
        return DSSSP_OK;
    });
    if (st != DSSSP_OK) return st;
    const size_t cnt = rec.size() / 3;
    std::vector<int32_t> va(cnt), vb(cnt), vw(cnt);
    for (size_t i = 0; i < cnt; ++i) {
        va[i] = (int32_t)rec[3 * i];
        vb[i] = (int32_t)rec[3 * i + 1];
        vw[i] = (int32_t)rec[3 * i + 2];
    }
    *a = dup_array(va);
    *b = dup_array(vb);
    *w = dup_array(vw);
    if (!*a || !*b || !*w) {
        free(*a), free(*b), free(*w);
        *a = *b = *w = nullptr;
        return fail(DSSSP_ERR_NOMEM, "out of host memory");
    }
    *m = (int64_t)cnt;
    return DSSSP_OK;
}

extern "C" int dsssp_read_changes(const char *path, int64_t n, int64_t *k, dsssp_change **chg)
{
    clear_error();
    if (!path || !k || !chg) return fail(DSSSP_ERR_ARG, "null argument");
    std::vector<char> buf;
    if (!read_file(path, buf)) return fail(DSSSP_ERR_IO, "cannot read %s", path);
    std::vector<int64_t> rec;
    int st = parse_parallel<4>(buf, rec, [n](const char *&p, const char *e, int64_t *v) {
        for (int i = 0; i < 4; ++i)
            if (!parse_int(p, e, v[i])) return fail(DSSSP_ERR_FORMAT, "expected 'a b w inst' integers");
        if (v[3] != 0 && v[3] != 1) return DSSSP_OK;  // ignored by readin_changes (:106,119)
        int s = check_id(v[0], n);
        if (s == DSSSP_OK) s = check_id(v[1], n);
        if (s != DSSSP_OK) return s;
        if (v[2] < 0 || v[2] > INT32_MAX) return fail(DSSSP_ERR_WEIGHT, "weight %" PRId64 " outside [0, 2^31)", v[2]);
        if (v[3] < INT32_MIN || v[3] > INT32_MAX) return fail(DSSSP_ERR_FORMAT, "inst out of range");
        return DSSSP_OK;
    });
    if (st != DSSSP_OK) return st;
    const size_t cnt = rec.size() / 4;
    dsssp_change *out = (dsssp_change *)malloc(sizeof(dsssp_change) * (cnt ? cnt : 1));
    if (!out) return fail(DSSSP_ERR_NOMEM, "out of host memory");
    for (size_t i = 0; i < cnt; ++i) {
        out[i].a = (int32_t)rec[4 * i];
        out[i].b = (int32_t)rec[4 * i + 1];
        out[i].w = (int32_t)rec[4 * i + 2];
        out[i].op = (int32_t)rec[4 * i + 3];
    }
    *chg = out;
    *k = (int64_t)cnt;
    return DSSSP_OK;
}

extern "C" int dsssp_read_sssp(const char *path, int64_t n, int32_t *parent, int64_t *dist)
{
    clear_error();
    if (!path || !parent || !dist || n < 0) return fail(DSSSP_ERR_ARG, "bad argument");
    std::vector<char> buf;
    if (!read_file(path, buf)) return fail(DSSSP_ERR_IO, "cannot read %s", path);
    std::vector<int64_t> rec;
    int st = parse_parallel<3>(buf, rec, [n](const char *&p, const char *e, int64_t *v) {
        if (!parse_int(p, e, v[0]) || !parse_int(p, e, v[1]))
            return fail(DSSSP_ERR_FORMAT, "expected 'v parent dist'");
        int s = check_id(v[0], n);
        if (s != DSSSP_OK) return s;
        if (v[1] == -1) {  // unreachable: any distance token, even a lossy double
            if (!skip_token(p, e)) return fail(DSSSP_ERR_FORMAT, "missing distance");
            v[2] = -1;
            return DSSSP_OK;
        }
        s = check_id(v[1], n);
        if (s != DSSSP_OK) return s;
        if (!parse_int(p, e, v[2]))
            return fail(DSSSP_ERR_FORMAT, "distance is not an exact integer (lossy output of the reference Dijkstra?)");
        if (v[2] < 0) return fail(DSSSP_ERR_FORMAT, "negative distance");
        return DSSSP_OK;
    });
    if (st != DSSSP_OK) return st;
    for (int64_t v = 0; v < n; ++v) {
        parent[v] = -1;
        dist[v] = DSSSP_INF;
    }
    for (size_t i = 0; i + 2 < rec.size(); i += 3) {
        const int64_t v = rec[i];
        if (rec[i + 1] == -1) {
            parent[v] = -1;
            dist[v] = DSSSP_INF;
        } else {
            parent[v] = (int32_t)rec[i + 1];
            dist[v] = rec[i + 2];
        }
    }
    return DSSSP_OK;
}

extern "C" int dsssp_write_sssp(const char *path, int64_t n, const int32_t *parent, const int64_t *dist)
{
    clear_error();
    if (!path || !parent || !dist || n < 0) return fail(DSSSP_ERR_ARG, "bad argument");
    FILE *f = fopen(path, "wb");
    if (!f) return fail(DSSSP_ERR_IO, "cannot write %s", path);
    std::vector<char> out;
    out.reserve(1 << 20);
    char line[64];
    bool ok = true;
    for (int64_t v = 0; v < n && ok; ++v) {
        int len;
        if (parent[v] < 0 || dist[v] == DSSSP_INF)
            len = snprintf(line, sizeof line, "%" PRId64 " -1 -1\n", v);
        else
            len = snprintf(line, sizeof line, "%" PRId64 " %d %" PRId64 "\n", v, parent[v], dist[v]);
        out.insert(out.end(), line, line + len);
        if (out.size() > (1u << 20)) {
            ok = fwrite(out.data(), 1, out.size(), f) == out.size();
            out.clear();
        }
    }
    if (ok && !out.empty()) ok = fwrite(out.data(), 1, out.size(), f) == out.size();
    if (fclose(f) != 0) ok = false;
    if (!ok) return fail(DSSSP_ERR_IO, "write to %s failed", path);
    return DSSSP_OK;
}

extern "C" void dsssp_free(void *p) { free(p); }
