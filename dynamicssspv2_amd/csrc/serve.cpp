// serve.cpp -- served form: resident instances and continuous-batching
// request loops on one GPU.
//
// The instances admitted before dsssp_server_start are spread over lanes
// (largest first, each to the lane with the fewest bytes); an instance admitted
// while the server runs gets a lane of its own.  A lane is ONE device graph
// holding its instances' vertex ranges end to end, with ONE tree object holding
// one root per instance, its own stream and its own loop thread.  A lane's loop
// takes the pending requests in arrival order -- for each instance the oldest
// ones, in order, as long as merging them keeps the result (below) -- up to
// max_merge requests / max_records records, concatenates them with vertex
// offsets, uploads and applies them as one update, then wakes the waiters.
// Requests that arrive while a merged update is on the GPU are picked up by the
// next one; the lanes' updates run concurrently on their streams, so updates of
// different lanes overlap on the device.  Components never share vertices or
// edges, so each instance's result is exactly what it would be alone.
//
// Merging consecutive requests of one instance.  A batch applies its
// insertions, then removes every copy of each deleted (a, b, w).  Applying
// requests R1 .. Rj as one batch leaves the same graph as applying them one
// after another unless some key deleted by an earlier request is inserted by a
// later one (one after another it is present at the end, merged it is not).
// When a merge holds several requests of one instance, the merged records are
// checked on the device before anything is applied (update_pinned): only the
// requests before the first such one are applied, the others go back to the
// front of the queue in order.  Equal graphs give equal trees (the tree is the
// exact, canonical one of the graph), so a merged update is byte-identical to
// the requests applied alone.
//
// An instance is removed (dsssp_server_remove_instance) after its requests
// have been applied; its lane, and the lane's share of the memory budget, is
// released when every instance of the lane has been removed.
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
    int32_t lane = -1;
    int64_t offset = 0;  // first vertex in the lane's graph
    std::vector<int32_t> a, b, w;  // until its lane is built
    int32_t src = 0;
    uint64_t bytes = 0;
    bool removed = false;
    int64_t inflight = 0;  // submitted, not yet applied
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

// What an instance adds to its lane graph's delta pool (half its entries plus
// the initial delta blocks), which 32-bit offsets address: LANE_ROWS at most.
uint64_t lane_rows(int64_t n, int64_t m) { return (uint64_t)m + 12 * (uint64_t)n; }
constexpr uint64_t LANE_ROWS = 0x7FFFFFFFull - (1u << 20);

// Device bytes of an instance: CSR, delta descriptors/capacities/counters and
// bitmaps, the initial delta pool (E/2 + 12 n + 2^16 entries), the edge map, the
// tree (state, bitmaps, chunk lists), and the buffers the tree allocates on
// first need (parent picks and levels, rescue candidates).
uint64_t instance_bytes(int64_t n, int64_t m)
{
    const uint64_t E = 2 * (uint64_t)m, N = (uint64_t)n;
    return 8 * (N + 1) + 8 * E + 16 * N + 3 * N / 8 + 8 * (E / 2 + 12 * N + (1u << 16)) + 8 * emap_slots_for(E + E / 8) +
           8 * N + 3 * N / 8 + 16 * (E / 128 + 1024) + 8 * N + N / 8;
}

}  // namespace

// One lane: a device graph holding some instances end to end, one tree with a
// root per instance, its own stream (the graph's) and its own loop thread.
struct Lane {
    std::vector<int32_t> members;  // instance ids
    int64_t N = 0;                 // vertices of the lane's graph
    uint64_t bytes = 0;            // budget its instances hold
    dsssp_graph *g = nullptr;
    dsssp_tree *t = nullptr;
    std::deque<Request *> pending;
    std::condition_variable cv_work;
    std::thread loop;
    std::mutex use;     // held while the lane's graph and tree are in use (one caller at a time)
    bool stop = false;  // the loop ends once its pending requests are applied
    bool alive = false;
};

struct dsssp_server {
    dsssp_server_config cfg{};
    std::deque<Instance> inst;  // deques: elements stay put while others are added
    std::deque<Lane> lanes;
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

// The requests a loop applies next, taken from L->pending in arrival order
// (caller holds s->mu); an instance whose request does not fit stops there, so
// every instance's order is kept.  Returns whether some instance has several
// requests in the merge (the merge is then checked on the device).
bool pick_requests(dsssp_server *s, Lane *L, std::vector<Request *> &pick)
{
    std::map<int32_t, int> seen;  // instance -> requests taken (-1: blocked)
    bool several = false;
    int64_t recs = 0;
    for (auto it = L->pending.begin(); it != L->pending.end();) {
        Request *r = *it;
        const bool room = (s->cfg.max_merge <= 0 || (int32_t)pick.size() < s->cfg.max_merge) &&
                          (pick.empty() || s->cfg.max_records <= 0 || recs + r->k <= s->cfg.max_records);
        int &c = seen[r->inst];
        if (c < 0 || !room) {
            c = -1;  // keeps the instance's order
            ++it;
            continue;
        }
        several = several || c > 0;
        ++c;
        recs += r->k;
        pick.push_back(r);
        it = L->pending.erase(it);
    }
    return several;
}

void serve_loop(dsssp_server *s, Lane *L)
{
    (void)dsssp_set_device(s->device);
    std::vector<const dsssp_change *> ptrs;  // reused per merged update
    std::vector<int64_t> lens, nis, nds;
    std::vector<uint8_t> zws;
    for (;;) {
        std::vector<Request *> pick;
        bool check;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            L->cv_work.wait(lk, [&] { return s->stopping || L->stop || !L->pending.empty(); });
            if (L->pending.empty() && (s->stopping || L->stop)) return;
            check = pick_requests(s, L, pick);
        }
        const auto t0 = sclock::now();
        // The picked requests' staged records are uploaded one after another.
        ptrs.clear();
        lens.clear();
        nis.clear();
        nds.clear();
        zws.clear();
        for (Request *r : pick) {
            if (r->start == sclock::time_point()) r->start = t0;
            ptrs.push_back(r->buf);
            lens.push_back(r->k);
            nis.push_back(r->ni);
            nds.push_back(r->nd);
            zws.push_back(r->zw ? 1 : 0);
        }
        dsssp_stats st{};
        int rc;
        int32_t took = 0;
        std::string err;
        {
            std::lock_guard<std::mutex> use(L->use);
            rc = update_pinned(L->g, L->t, ptrs.data(), lens.data(), nis.data(), nds.data(), zws.data(),
                               (int32_t)pick.size(), check, &took, &st);
            if (rc) err = dsssp_last_error();
        }
        if (rc) took = (int32_t)pick.size();  // the whole merge fails together
        // Requests past the merge check's cut go back to the front, in order.
        std::vector<Request *> later(pick.begin() + took, pick.end());
        pick.resize((size_t)took);
        int64_t nrec = 0;
        for (Request *r : pick) {  // the uploads are complete when the update returns
            nrec += r->k;
            give_buffer(s, r->buf, r->cap);
            r->buf = nullptr;
        }
        const auto t1 = sclock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            L->pending.insert(L->pending.begin(), later.begin(), later.end());
            for (Request *r : pick) {
                r->finish = t1;
                r->status = rc;
                r->error = err;
                r->merged = (int32_t)pick.size();
                r->update_ms = st.t_update_ms;
                r->done = true;
                s->inst[r->inst].inflight -= 1;
            }
            s->launches += 1;
            s->records += nrec;
            s->requests_done += (int64_t)pick.size();
            s->busy_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
        }
        s->cv_done.notify_all();
    }
}

// Builds the device graph and tree of instances `members` (vertex ranges end to
// end, offsets recorded in the instances), releasing their host edge copies.
int build_lane(dsssp_server *s, const std::vector<int32_t> &members, dsssp_graph **g, dsssp_tree **t, int64_t *N_out)
{
    *g = nullptr;
    *t = nullptr;
    int64_t N = 0, M = 0;
    for (int32_t i : members) {
        s->inst[i].offset = N;
        N += s->inst[i].n;
        M += s->inst[i].m;
    }
    std::vector<int32_t> A((size_t)M), B((size_t)M), W((size_t)M), roots;
    int64_t e = 0;
    for (int32_t i : members) {
        const Instance &x = s->inst[i];
        const int32_t off = (int32_t)x.offset;
        for (int64_t j = 0; j < x.m; ++j, ++e) {
            A[e] = x.a[j] + off;
            B[e] = x.b[j] + off;
            W[e] = x.w[j];
        }
        roots.push_back(x.src + off);
    }
    int rc = dsssp_graph_create(N, M, A.data(), B.data(), W.data(), DSSSP_GRAPH_DEFAULT, g);
    if (!rc) rc = dsssp_tree_create_roots(*g, (int32_t)roots.size(), roots.data(), DSSSP_TREE_DEFAULT, t);
    if (rc) {
        const std::string keep = dsssp_last_error();
        if (*t) dsssp_tree_destroy(*t);
        if (*g) dsssp_graph_destroy(*g);
        *g = nullptr;
        *t = nullptr;
        return fail(rc, "%s", keep.c_str());
    }
    for (int32_t i : members) {
        Instance &x = s->inst[i];
        std::vector<int32_t>().swap(x.a);
        std::vector<int32_t>().swap(x.b);
        std::vector<int32_t>().swap(x.w);
    }
    *N_out = N;
    return DSSSP_OK;
}

// Ends one lane's loop (after its pending requests) and frees its device objects.
void stop_lane(dsssp_server *s, Lane &L)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        L.stop = true;
    }
    L.cv_work.notify_all();
    if (L.loop.joinable()) L.loop.join();
    std::lock_guard<std::mutex> use(L.use);  // waits for a reader of the lane's tree
    if (L.t) dsssp_tree_destroy(L.t);
    if (L.g) dsssp_graph_destroy(L.g);
    L.t = nullptr;
    L.g = nullptr;
}

// Stops the loops and frees the lanes' device objects.
void stop_lanes(dsssp_server *s)
{
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    for (Lane &L : s->lanes) L.cv_work.notify_all();
    for (Lane &L : s->lanes) stop_lane(s, L);
}

// Budget and size checks of a new instance (caller holds s->mu).
int admit(dsssp_server *s, int64_t n, int64_t m, int32_t src, uint64_t *bytes)
{
    if (src < 0 || src >= n) return fail(DSSSP_ERR_RANGE, "source %d outside [0,%lld)", src, (long long)n);
    *bytes = instance_bytes(n, m);
    if (lane_rows(n, m) > LANE_ROWS)
        return fail(DSSSP_ERR_BUDGET, "instance too large for one device graph (n=%lld, m=%lld)", (long long)n,
                    (long long)m);
    if (s->used + *bytes > s->cfg.mem_budget_bytes)
        return fail(DSSSP_ERR_BUDGET, "instance needs %llu bytes, %llu of %llu left", (unsigned long long)*bytes,
                    (unsigned long long)(s->cfg.mem_budget_bytes - s->used),
                    (unsigned long long)s->cfg.mem_budget_bytes);
    if (n > 0x7FFFFFFF) return fail(DSSSP_ERR_BUDGET, "more than 2^31-1 vertices in one instance");
    return DSSSP_OK;
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
    for (int64_t j = 0; j < m; ++j)
        if (a[j] < 0 || a[j] >= n || b[j] < 0 || b[j] >= n || w[j] < 0)
            return fail(w[j] < 0 ? DSSSP_ERR_WEIGHT : DSSSP_ERR_RANGE, "edge %lld (%d %d %d) invalid for n=%lld",
                        (long long)j, a[j], b[j], w[j], (long long)n);
    uint64_t bytes = 0;
    bool running;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stopping) return fail(DSSSP_ERR_STATE, "server is being destroyed");
        if (int rc = admit(s, n, m, src, &bytes)) return rc;
        running = s->started;
        s->used += bytes;  // held from here on; given back if the lane cannot be built
        s->inst.emplace_back();
        Instance &I = s->inst.back();
        I.n = n;
        I.m = m;
        I.src = src;
        I.bytes = bytes;
        I.removed = running;  // not open to submissions until its lane runs
        I.a.assign(a, a + m);
        I.b.assign(b, b + m);
        I.w.assign(w, w + m);
        *id = (int32_t)s->inst.size() - 1;
    }
    if (!running) return DSSSP_OK;
    // Admitted while the server runs: a lane of its own, built in this thread
    // while the other lanes go on serving.
    const int32_t me = *id;
    dsssp_graph *g = nullptr;
    dsssp_tree *t = nullptr;
    int64_t N = 0;
    const int rc = build_lane(s, {me}, &g, &t, &N);
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc) {
        s->used -= bytes;
        s->inst[me].lane = -1;  // never served; the id stays unused
        return rc;
    }
    if (s->stopping) {  // destroyed meanwhile
        dsssp_tree_destroy(t);
        dsssp_graph_destroy(g);
        return fail(DSSSP_ERR_STATE, "server is being destroyed");
    }
    s->lanes.emplace_back();
    Lane &L = s->lanes.back();
    L.members = {me};
    L.N = N;
    L.bytes = bytes;
    L.g = g;
    L.t = t;
    L.alive = true;
    s->inst[me].lane = (int32_t)s->lanes.size() - 1;
    s->inst[me].removed = false;
    L.loop = std::thread(serve_loop, s, &L);
    return DSSSP_OK;
}

extern "C" int dsssp_server_remove_instance(dsssp_server *s, int32_t id)
{
    clear_error();
    if (!s) return fail(DSSSP_ERR_ARG, "null argument");
    Lane *L = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        if (id < 0 || id >= (int32_t)s->inst.size()) return fail(DSSSP_ERR_ARG, "unknown instance %d", id);
        Instance &I = s->inst[id];
        if (I.removed || (s->started && I.lane < 0)) return fail(DSSSP_ERR_STATE, "instance %d is not served", id);
        I.removed = true;  // no new submissions
        if (!s->started) {  // admitted but never placed: only the budget to give back
            s->used -= I.bytes;
            std::vector<int32_t>().swap(I.a);
            std::vector<int32_t>().swap(I.b);
            std::vector<int32_t>().swap(I.w);
            return DSSSP_OK;
        }
        s->cv_done.wait(lk, [&] { return I.inflight == 0; });
        Lane &Ln = s->lanes[I.lane];
        bool all = true;
        for (int32_t j : Ln.members) all = all && s->inst[j].removed;
        if (all && Ln.alive) {
            Ln.alive = false;
            L = &Ln;
        }
    }
    if (L) {
        stop_lane(s, *L);
        std::lock_guard<std::mutex> lk(s->mu);
        s->used -= L->bytes;
    }
    return DSSSP_OK;
}

extern "C" int dsssp_server_start(dsssp_server *s)
{
    clear_error();
    if (!s) return fail(DSSSP_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->started) return fail(DSSSP_ERR_STATE, "already started");
    std::vector<int32_t> live;
    for (int32_t i = 0; i < (int32_t)s->inst.size(); ++i)
        if (!s->inst[i].removed) live.push_back(i);
    if (live.empty()) return fail(DSSSP_ERR_STATE, "no instance admitted");
    const int ninst = (int)live.size();
    const int want = std::max(1, std::min(ninst, s->cfg.lanes > 0 ? (int)s->cfg.lanes : 1));
    // Largest instances first, each to the lane with the fewest bytes so far
    // among those its rows still fit in (a device graph's delta pool is
    // addressed with 32 bits); a new lane when none does.
    std::vector<int32_t> order = live;
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
    int rc = DSSSP_OK;
    for (int l = 0; l < nl; ++l) s->lanes.emplace_back();
    for (int l = 0; l < nl && !rc; ++l) {
        Lane &L = s->lanes[l];
        std::sort(members[l].begin(), members[l].end());
        L.members = members[l];
        for (int32_t i : L.members) {
            s->inst[i].lane = l;
            L.bytes += s->inst[i].bytes;
        }
        rc = build_lane(s, L.members, &L.g, &L.t, &L.N);
    }
    if (rc) {
        const std::string keep = dsssp_last_error();
        for (Lane &L : s->lanes) {
            if (L.t) dsssp_tree_destroy(L.t);
            if (L.g) dsssp_graph_destroy(L.g);
        }
        s->lanes.clear();
        for (int32_t i : live) s->inst[i].lane = -1;
        return fail(rc, "%s", keep.c_str());
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
    for (Lane &L : s->lanes) {
        L.alive = true;
        L.loop = std::thread(serve_loop, s, &L);
    }
    return DSSSP_OK;
}

extern "C" int dsssp_server_submit(dsssp_server *s, int32_t id, const dsssp_change *chg, int64_t k, uint64_t *ticket)
{
    clear_error();
    if (!s || !ticket || k < 0 || (k > 0 && !chg)) return fail(DSSSP_ERR_ARG, "bad argument");
    int64_t n = 0, off = 0;
    Instance *I = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->started || s->stopping) return fail(DSSSP_ERR_STATE, "server not running");
        if (id < 0 || id >= (int32_t)s->inst.size()) return fail(DSSSP_ERR_ARG, "unknown instance %d", id);
        I = &s->inst[id];
        if (I->removed || I->lane < 0) return fail(DSSSP_ERR_STATE, "instance %d is not served", id);
        n = I->n;  // fixed once placed
        off = I->offset;
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
    if (s->stopping || I->removed) {
        lk.unlock();
        give_buffer(s, r->buf, r->cap);
        delete r;
        return fail(DSSSP_ERR_STATE, "instance %d is not served", id);
    }
    r->ticket = s->next_ticket++;
    r->submit = sclock::now();
    Lane &L = s->lanes[I->lane];
    L.pending.push_back(r);
    I->inflight += 1;
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
    Instance &I = s->inst[id];
    if (I.removed || I.lane < 0) return fail(DSSSP_ERR_STATE, "instance %d is not served", id);
    // wait until no request of this instance is pending or running
    s->cv_done.wait(lk, [&] { return I.inflight == 0; });
    const int64_t n = I.n, off = I.offset;
    Lane &L = s->lanes[I.lane];
    lk.unlock();  // the lane's loop may go on with other instances meanwhile
    std::vector<int32_t> par((size_t)L.N);
    std::vector<int64_t> d((size_t)L.N);
    int rc;
    {
        std::lock_guard<std::mutex> use(L.use);
        if (!L.t) return fail(DSSSP_ERR_STATE, "instance %d is not served", id);  // retired meanwhile
        rc = dsssp_tree_read(L.t, par.data(), d.data());
    }
    if (rc) return rc;
    for (int64_t v = 0; v < n; ++v) {
        const int32_t p = par[off + v];
        parent[v] = p < 0 ? -1 : p - (int32_t)off;
        dist[v] = d[off + v];
    }
    return DSSSP_OK;
}

extern "C" int dsssp_server_info(dsssp_server *s, dsssp_server_totals *tot)
{
    clear_error();
    if (!s || !tot) return fail(DSSSP_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    int32_t live = 0;
    for (const Instance &I : s->inst) live += !I.removed && (I.lane >= 0 || !s->started);
    tot->instances = live;
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
