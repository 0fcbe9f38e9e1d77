// serve.cpp -- served form: resident instances and continuous-batching
// request loops on one GPU.
//
// The admitted instances are spread over lanes (largest first, each to the
// lane with the fewest bytes).  A lane is ONE device graph holding its
// instances' vertex ranges end to end, with ONE tree object holding one root
// per instance, its own stream and its own loop thread.  A lane's loop takes,
// in arrival order, the oldest pending request of every one of its instances
// that has one (FIFO per instance), up to max_merge requests / max_records
// records, concatenates them with vertex offsets, uploads and applies them as
// one update, then wakes the waiters.  Requests that arrive while a merged
// update is on the GPU are picked up by the next one; the lanes' updates run
// concurrently on their streams, so small updates of different lanes overlap on
// the device.  Components never share vertices or edges, so each instance's
// result is exactly what it would be alone.
// The reference has no counterpart (one graph and one batch per process,
// SSSPmainForUndirected.cu:27-335).
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/dsssp.h"
#include "host.h"

using namespace dsssp;
using sclock = std::chrono::steady_clock;

namespace {

struct Instance {
    int64_t n = 0, m = 0;
    int32_t lane = 0;
    int64_t offset = 0;  // first vertex in the lane's graph
    std::vector<int32_t> a, b, w;
    int32_t src = 0;
    uint64_t bytes = 0;
};

struct Request {
    uint64_t ticket = 0;
    int32_t inst = 0;
    // The records, staged at submission: pinned host memory, in the lane
    // graph's vertex ids, counted.
    dsssp_change *buf = nullptr;
    int64_t cap = 0;   // > 0: a buffer of its own; 0: a slice of the arena
    int64_t k = 0, ni = 0, nd = 0;
    bool zw = false;
    sclock::time_point submit, start, finish;
    int status = DSSSP_OK;
    std::string error;
    int32_t merged = 0;
    double update_ms = 0;
    bool done = false;
};

// Device bytes of an instance: CSR, delta descriptors/capacities/counters and
// bitmaps, the initial delta pool (E/2 + 12 n + 2^16 entries), the edge map, and the
// tree (state, bitmaps, chunk lists).
// What an instance adds to its lane graph's delta pool (half its entries plus
// the initial delta blocks), which 32-bit offsets address: LANE_ROWS at most.
uint64_t lane_rows(int64_t n, int64_t m) { return (uint64_t)m + 12 * (uint64_t)n; }
constexpr uint64_t LANE_ROWS = 0x7FFFFFFFull - (1u << 20);

uint64_t instance_bytes(int64_t n, int64_t m)
{
    const uint64_t E = 2 * (uint64_t)m, N = (uint64_t)n;
    return 8 * (N + 1) + 8 * E + 16 * N + 3 * N / 8 + 8 * (E / 2 + 12 * N + (1u << 16)) + 8 * emap_slots_for(E + E / 8) +
           8 * N + 3 * N / 8 + 16 * (E / 128 + 1024);
}

}  // namespace

// One lane: a device graph holding some instances end to end, one tree with a
// root per instance, its own stream (the graph's) and its own loop thread.
struct Lane {
    std::vector<int32_t> members;  // instance ids
    int64_t N = 0;                 // vertices of the lane's graph
    dsssp_graph *g = nullptr;
    dsssp_tree *t = nullptr;
    std::deque<Request *> pending;
    std::condition_variable cv_work;
    std::thread loop;
    std::mutex use;  // held while the lane's graph and tree are in use (one caller at a time)
};

struct dsssp_server {
    dsssp_server_config cfg{};
    std::vector<Instance> inst;
    std::vector<Lane> lanes;
    uint64_t used = 0;
    int device = 0;
    bool started = false, stopping = false;
    std::mutex mu;
    std::condition_variable cv_done;
    std::map<uint64_t, Request *> all;
    uint64_t next_ticket = 1;
    // Pinned staging of submitted records: slices of one arena, allocated in
    // submission order and released as requests finish (the oldest live slice
    // bounds the free space); a request that does not fit gets a buffer of its
    // own, kept by capacity for reuse.
    std::mutex pool_mu;
    dsssp_change *arena = nullptr;
    int64_t arena_cap = 0, head = 0;
    struct Slice {
        int64_t start, len;
        bool live;
    };
    std::deque<Slice> slices;
    std::multimap<int64_t, dsssp_change *> pool;
    // totals
    int64_t launches = 0, records = 0, requests_done = 0;
    double busy_ms = 0;
};

namespace {

constexpr int64_t ARENA_RECORDS = int64_t(1) << 22;  // 64 MiB of pinned staging

// Pinned room for k records: an arena slice (*cap = 0) when one fits, else a
// buffer of its own (from the pool, or a new one of a power-of-two capacity).
int take_buffer(dsssp_server *s, int64_t k, dsssp_change **buf, int64_t *cap)
{
    {
        std::lock_guard<std::mutex> lk(s->pool_mu);
        if (s->arena && k > 0) {
            if (s->slices.empty()) s->head = 0;
            const int64_t tail = s->slices.empty() ? 0 : s->slices.front().start;
            int64_t at = -1;
            if (s->slices.empty() || s->head > tail) {  // free: [head, cap) and [0, tail)
                if (s->head + k <= s->arena_cap)
                    at = s->head;
                else if (k <= tail)
                    at = 0;
            } else if (s->head + k <= tail) {  // wrapped: free [head, tail)
                at = s->head;
            }
            if (at >= 0) {
                s->slices.push_back({at, k, true});
                s->head = at + k;
                *buf = s->arena + at;
                *cap = 0;
                return DSSSP_OK;
            }
        }
        auto it = s->pool.lower_bound(k);
        if (it != s->pool.end()) {
            *cap = it->first;
            *buf = it->second;
            s->pool.erase(it);
            return DSSSP_OK;
        }
    }
    int64_t c = 1024;
    while (c < k) c *= 2;
    void *p = nullptr;
    if (int rc = pinned_alloc(&p, sizeof(dsssp_change) * (size_t)c)) return rc;
    *buf = (dsssp_change *)p;
    *cap = c;
    return DSSSP_OK;
}

void give_buffer(dsssp_server *s, dsssp_change *buf, int64_t cap)
{
    if (!buf) return;
    std::lock_guard<std::mutex> lk(s->pool_mu);
    if (cap > 0) {
        s->pool.emplace(cap, buf);
        return;
    }
    const int64_t at = buf - s->arena;
    for (auto &sl : s->slices)
        if (sl.start == at && sl.live) {
            sl.live = false;
            break;
        }
    while (!s->slices.empty() && !s->slices.front().live) s->slices.pop_front();
}

void serve_loop(dsssp_server *s, Lane *L)
{
    (void)dsssp_set_device(s->device);
    std::vector<const dsssp_change *> ptrs;  // reused per merged update
    std::vector<int64_t> lens;
    for (;;) {
        std::vector<Request *> pick;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            L->cv_work.wait(lk, [&] { return s->stopping || !L->pending.empty(); });
            if (L->pending.empty() && s->stopping) return;
            std::vector<char> taken(s->inst.size(), 0);
            int64_t recs = 0;
            for (auto it = L->pending.begin(); it != L->pending.end();) {
                Request *r = *it;
                const bool room = (s->cfg.max_merge <= 0 || (int32_t)pick.size() < s->cfg.max_merge) &&
                                  (pick.empty() || s->cfg.max_records <= 0 || recs + r->k <= s->cfg.max_records);
                if (!taken[r->inst] && room) {
                    taken[r->inst] = 1;
                    recs += r->k;
                    pick.push_back(r);
                    it = L->pending.erase(it);
                } else {
                    taken[r->inst] = 1;  // keeps per-instance FIFO order
                    ++it;
                }
            }
        }
        const auto t0 = sclock::now();
        // The picked requests' staged records are uploaded one after another.
        ptrs.clear();
        lens.clear();
        int64_t nrec = 0, ni = 0, nd = 0;
        bool zw = false;
        for (Request *r : pick) {
            r->start = t0;
            ptrs.push_back(r->buf);
            lens.push_back(r->k);
            nrec += r->k;
            ni += r->ni;
            nd += r->nd;
            zw = zw || r->zw;
        }
        dsssp_stats st{};
        int rc;
        std::string err;
        {
            std::lock_guard<std::mutex> use(L->use);
            rc = update_pinned(L->g, L->t, ptrs.data(), lens.data(), (int32_t)pick.size(), ni, nd, zw, &st);
            if (rc) err = dsssp_last_error();
        }
        for (Request *r : pick) {  // the uploads are complete when the update returns
            give_buffer(s, r->buf, r->cap);
            r->buf = nullptr;
        }
        const auto t1 = sclock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            for (Request *r : pick) {
                r->finish = t1;
                r->status = rc;
                r->error = err;
                r->merged = (int32_t)pick.size();
                r->update_ms = st.t_update_ms;
                r->done = true;
            }
            s->launches += 1;
            s->records += nrec;
            s->requests_done += (int64_t)pick.size();
            s->busy_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
        }
        s->cv_done.notify_all();
    }
}

// Stops the loops and frees the lanes' device objects.
void stop_lanes(dsssp_server *s)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    for (Lane &L : s->lanes) L.cv_work.notify_all();
    for (Lane &L : s->lanes)
        if (L.loop.joinable()) L.loop.join();
    for (Lane &L : s->lanes) {
        if (L.t) dsssp_tree_destroy(L.t);
        if (L.g) dsssp_graph_destroy(L.g);
        L.t = nullptr;
        L.g = nullptr;
    }
}

}  // namespace

extern "C" int dsssp_server_create(const dsssp_server_config *cfg, dsssp_server **out)
{
    clear_error();
    if (!cfg || !out) return fail(DSSSP_ERR_ARG, "null argument");
    *out = nullptr;
    if (cfg->mem_budget_bytes == 0) return fail(DSSSP_ERR_ARG, "memory budget must be positive");
    dsssp_server *s = new dsssp_server();
    s->cfg = *cfg;
    *out = s;
    return DSSSP_OK;
}

extern "C" int dsssp_server_add_instance(dsssp_server *s, int64_t n, int64_t m, const int32_t *a, const int32_t *b,
                                         const int32_t *w, int32_t src, int32_t *id)
{
    clear_error();
    if (!s || !id || n < 1 || m < 0 || (m > 0 && (!a || !b || !w))) return fail(DSSSP_ERR_ARG, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->started) return fail(DSSSP_ERR_STATE, "instances are admitted before dsssp_server_start");
    if (src < 0 || src >= n) return fail(DSSSP_ERR_RANGE, "source %d outside [0,%lld)", src, (long long)n);
    const uint64_t bytes = instance_bytes(n, m);
    if (lane_rows(n, m) > LANE_ROWS)
        return fail(DSSSP_ERR_BUDGET, "instance too large for one device graph (n=%lld, m=%lld)", (long long)n,
                    (long long)m);
    if (s->used + bytes > s->cfg.mem_budget_bytes)
        return fail(DSSSP_ERR_BUDGET, "instance needs %llu bytes, %llu of %llu left", (unsigned long long)bytes,
                    (unsigned long long)(s->cfg.mem_budget_bytes - s->used),
                    (unsigned long long)s->cfg.mem_budget_bytes);
    int64_t total = 0;
    for (auto &x : s->inst) total += x.n;
    if (total + n > 0x7FFFFFFF) return fail(DSSSP_ERR_BUDGET, "more than 2^31-1 vertices in total");
    Instance I;
    I.n = n;
    I.m = m;
    I.a.assign(a, a + m);
    I.b.assign(b, b + m);
    I.w.assign(w, w + m);
    I.src = src;
    I.bytes = bytes;
    s->used += bytes;
    s->inst.push_back(std::move(I));
    *id = (int32_t)s->inst.size() - 1;
    return DSSSP_OK;
}

extern "C" int dsssp_server_start(dsssp_server *s)
{
    clear_error();
    if (!s) return fail(DSSSP_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->started) return fail(DSSSP_ERR_STATE, "already started");
    if (s->inst.empty()) return fail(DSSSP_ERR_STATE, "no instance admitted");
    const int ninst = (int)s->inst.size();
    const int want = std::max(1, std::min(ninst, s->cfg.lanes > 0 ? (int)s->cfg.lanes : 1));
    // Largest instances first, each to the lane with the fewest bytes so far
    // among those its rows still fit in (a device graph's delta pool is
    // addressed with 32 bits); a new lane when none does.
    std::vector<int32_t> order((size_t)ninst);
    for (int i = 0; i < ninst; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int32_t x, int32_t y) { return s->inst[x].bytes > s->inst[y].bytes; });
    std::vector<uint64_t> load, rows;
    std::vector<std::vector<int32_t>> members;
    for (int32_t i : order) {
        const uint64_t r = lane_rows(s->inst[i].n, s->inst[i].m);
        int best = -1;
        if ((int)load.size() >= want)
            for (int l = 0; l < (int)load.size(); ++l)
                if (rows[l] + r <= LANE_ROWS && (best < 0 || load[l] < load[best])) best = l;
        if (best < 0) {
            best = (int)load.size();
            load.push_back(0);
            rows.push_back(0);
            members.emplace_back();
        }
        load[best] += s->inst[i].bytes;
        rows[best] += r;
        members[best].push_back(i);
    }
    const int nl = (int)members.size();
    s->lanes = std::vector<Lane>((size_t)nl);
    int rc = DSSSP_OK;
    for (int l = 0; l < nl && !rc; ++l) {
        Lane &L = s->lanes[l];
        std::sort(members[l].begin(), members[l].end());
        L.members = members[l];
        int64_t N = 0, M = 0;
        for (int32_t i : L.members) {
            s->inst[i].lane = l;
            s->inst[i].offset = N;
            N += s->inst[i].n;
            M += s->inst[i].m;
        }
        L.N = N;
        std::vector<int32_t> A((size_t)M), B((size_t)M), W((size_t)M), roots;
        int64_t e = 0;
        for (int32_t i : L.members) {
            const Instance &x = s->inst[i];
            const int32_t off = (int32_t)x.offset;
            for (int64_t j = 0; j < x.m; ++j, ++e) {
                A[e] = x.a[j] + off;
                B[e] = x.b[j] + off;
                W[e] = x.w[j];
            }
            roots.push_back(x.src + off);
        }
        rc = dsssp_graph_create(N, M, A.data(), B.data(), W.data(), DSSSP_GRAPH_DEFAULT, &L.g);
        if (!rc) rc = dsssp_tree_create_roots(L.g, (int32_t)roots.size(), roots.data(), DSSSP_TREE_DEFAULT, &L.t);
    }
    if (rc) {
        const std::string keep = dsssp_last_error();
        for (Lane &L : s->lanes) {
            if (L.t) dsssp_tree_destroy(L.t);
            if (L.g) dsssp_graph_destroy(L.g);
        }
        s->lanes.clear();
        return fail(rc, "%s", keep.c_str());
    }
    for (auto &x : s->inst) {
        std::vector<int32_t>().swap(x.a);
        std::vector<int32_t>().swap(x.b);
        std::vector<int32_t>().swap(x.w);
    }
    {
        void *p = nullptr;
        if (pinned_alloc(&p, sizeof(dsssp_change) * (size_t)ARENA_RECORDS) == DSSSP_OK) {
            s->arena = (dsssp_change *)p;
            s->arena_cap = ARENA_RECORDS;
            // first touch now rather than in the first submissions
            memset(s->arena, 0, sizeof(dsssp_change) * (size_t)ARENA_RECORDS);
        }
        clear_error();  // without the arena every request gets a buffer of its own
    }
    s->started = true;
    for (Lane &L : s->lanes) L.loop = std::thread(serve_loop, s, &L);
    return DSSSP_OK;
}

extern "C" int dsssp_server_submit(dsssp_server *s, int32_t id, const dsssp_change *chg, int64_t k, uint64_t *ticket)
{
    clear_error();
    if (!s || !ticket || k < 0 || (k > 0 && !chg)) return fail(DSSSP_ERR_ARG, "bad argument");
    int64_t n = 0, off = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->started || s->stopping) return fail(DSSSP_ERR_STATE, "server not running");
        if (id < 0 || id >= (int32_t)s->inst.size()) return fail(DSSSP_ERR_ARG, "unknown instance %d", id);
        n = s->inst[id].n;  // fixed once started
        off = s->inst[id].offset;
    }
    // Validation and staging happen here, in the submitting thread, outside the
    // lock: the records are checked, shifted to the lane graph's ids and counted
    // straight into a pinned buffer, so a loop only enqueues their upload.
    Request *r = new Request();
    r->inst = id;
    r->k = k;
    if (k > 0) {
        if (int rc = take_buffer(s, k, &r->buf, &r->cap)) {
            delete r;
            return rc;
        }
    }
    for (int64_t i = 0; i < k; ++i) {
        dsssp_change c = chg[i];
        if (c.op == 0 || c.op == 1) {
            if (c.a < 0 || c.a >= n || c.b < 0 || c.b >= n || c.w < 0) {
                give_buffer(s, r->buf, r->cap);
                delete r;
                if (c.w < 0 && c.a >= 0 && c.a < n && c.b >= 0 && c.b < n)
                    return fail(DSSSP_ERR_WEIGHT, "change %lld: negative weight", (long long)i);
                return fail(DSSSP_ERR_RANGE, "change %lld: vertex outside [0,%lld) of instance %d", (long long)i,
                            (long long)n, id);
            }
            c.a += (int32_t)off;
            c.b += (int32_t)off;
            if (c.op == 1) {
                ++r->ni;
                if (c.w == 0) r->zw = true;
            } else {
                ++r->nd;
            }
        }
        r->buf[i] = c;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->stopping) {
        lk.unlock();
        give_buffer(s, r->buf, r->cap);
        delete r;
        return fail(DSSSP_ERR_STATE, "server not running");
    }
    r->ticket = s->next_ticket++;
    r->submit = sclock::now();
    Lane &L = s->lanes[s->inst[id].lane];
    L.pending.push_back(r);
    s->all[r->ticket] = r;
    *ticket = r->ticket;
    lk.unlock();
    L.cv_work.notify_one();
    return DSSSP_OK;
}

extern "C" int dsssp_server_wait(dsssp_server *s, uint64_t ticket, dsssp_request_stats *st)
{
    clear_error();
    if (!s) return fail(DSSSP_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->all.find(ticket);
    if (it == s->all.end()) return fail(DSSSP_ERR_ARG, "unknown ticket %llu", (unsigned long long)ticket);
    Request *r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    const int rc = r->status;
    const std::string err = r->error;
    if (st) {
        st->instance = r->inst;
        st->records = r->k;
        st->merged = r->merged;
        st->queue_ms = std::chrono::duration<double, std::milli>(r->start - r->submit).count();
        st->latency_ms = std::chrono::duration<double, std::milli>(r->finish - r->submit).count();
        st->update_ms = r->update_ms;
    }
    s->all.erase(it);
    delete r;
    lk.unlock();
    if (rc) return fail(rc, "%s", err.c_str());
    return DSSSP_OK;
}

extern "C" int dsssp_server_read(dsssp_server *s, int32_t id, int32_t *parent, int64_t *dist)
{
    clear_error();
    if (!s || !parent || !dist) return fail(DSSSP_ERR_ARG, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (!s->started) return fail(DSSSP_ERR_STATE, "server not started");
    if (id < 0 || id >= (int32_t)s->inst.size()) return fail(DSSSP_ERR_ARG, "unknown instance %d", id);
    // wait until no request of this instance is pending or running
    s->cv_done.wait(lk, [&] {
        for (auto &kv : s->all)
            if (kv.second->inst == id && !kv.second->done) return false;
        return true;
    });
    const Instance &I = s->inst[id];
    Lane &L = s->lanes[I.lane];
    lk.unlock();  // the lane's loop may go on with other instances meanwhile
    std::vector<int32_t> par((size_t)L.N);
    std::vector<int64_t> d((size_t)L.N);
    int rc;
    {
        std::lock_guard<std::mutex> use(L.use);
        rc = dsssp_tree_read(L.t, par.data(), d.data());
    }
    if (rc) return rc;
    for (int64_t v = 0; v < I.n; ++v) {
        const int32_t p = par[I.offset + v];
        parent[v] = p < 0 ? -1 : p - (int32_t)I.offset;
        dist[v] = d[I.offset + v];
    }
    return DSSSP_OK;
}

extern "C" int dsssp_server_info(dsssp_server *s, dsssp_server_totals *tot)
{
    clear_error();
    if (!s || !tot) return fail(DSSSP_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    tot->instances = (int32_t)s->inst.size();
    tot->budget_bytes = s->cfg.mem_budget_bytes;
    tot->used_bytes = s->used;
    tot->launches = s->launches;
    tot->records = s->records;
    tot->requests = s->requests_done;
    tot->busy_ms = s->busy_ms;
    int64_t pend = 0;
    for (const Lane &L : s->lanes) pend += (int64_t)L.pending.size();
    tot->pending = pend;
    return DSSSP_OK;
}

extern "C" int dsssp_server_destroy(dsssp_server *s)
{
    clear_error();
    if (!s) return DSSSP_OK;
    stop_lanes(s);
    for (auto &kv : s->all) {
        if (kv.second->cap > 0) pinned_free(kv.second->buf);
        delete kv.second;
    }
    for (auto &kv : s->pool) pinned_free(kv.second);
    pinned_free(s->arena);
    delete s;
    return DSSSP_OK;
}
