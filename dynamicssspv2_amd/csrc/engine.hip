// engine.hip -- device objects and the update pipeline behind the C ABI.
//
// Update pipeline (one HIP stream per graph, no host round-trip per round):
//   1. insert   k_ins_place / wordsum / scan / alloc / pending: the batch's
//               insertions become delta-row entries in HBM (the reference
//               appends them to its host adjacency before upload,
//               all_structure_undir.cuh:106-117)
//   2. classify k_classify: tree-edge test, roots of invalidated subtrees
//               marked; k_tomb: every copy of each deleted entry tombstoned
//               (deleteEdge, gpuFunctions_undir.cuh:179-215)
//   3. invalidate rounds of k_frontier<M_INVAL>, descending the pre-update tree
//               (updateNeighbors_del + filterAffectedNodes loop,
//               SSSPmainForUndirected.cu:219-239, gpuFunctions_undir.cuh:125-144,312-349)
//   4. seed     k_frontier<M_SEED> over every invalidated slot + k_ins_offer
//               (insertEdge, gpuFunctions_undir.cuh:217-251)
//   5. relax    rounds of k_frontier<M_RELAX> until no slot is marked
//               (updateNeighbors + filterAllAffected loop, :162-174,370-433,
//               SSSPmainForUndirected.cu:263-281)
//   6. repair   only for graphs that ever held a zero weight: tight-graph BFS
//               levels and parent re-pick so the tree stays acyclic
// Round i of a loop phase reads frontier bitmap i % 2 and flag / chunk-count
// slot i % 3 and writes the next ones, so a group of rounds has fixed arguments
// and can be replayed from a captured hipGraph.  The host reads one flag pair
// per group of rounds (the reference syncs and reads UVM every round).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dsssp.h"
#include "common.h"
#include "host.h"
#include "kernels.h"

using namespace dsssp;

namespace {

dsssp_tuning g_tuning = {FRONTIER_BLOCK, 1024, 256, 256, 4, 0, 0, 4, 1024, 2, DSSSP_DELTA_AUTO, -1, 120000,
                         DSSSP_RESCUE_AUTO, 16, 0};
std::mutex g_tuning_mu;
thread_local int t_device = 0;

dsssp_tuning tuning()
{
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    return g_tuning;
}

#define HIP_TRY(expr)                                                                                    \
    do {                                                                                                 \
        hipError_t e_ = (expr);                                                                          \
        if (e_ != hipSuccess) return fail(DSSSP_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_)); \
    } while (0)

int dev_alloc(void **p, size_t bytes, const char *what)
{
    *p = nullptr;
    if (bytes == 0) bytes = 8;
    hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(DSSSP_ERR_NOMEM, "hipMalloc of %zu bytes for %s failed: %s", bytes, what, hipGetErrorString(e));
    }
    return DSSSP_OK;
}

template <class T>
void dev_free(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

int check_device(int dev)
{
    static std::mutex mu;
    static std::vector<int> ok_devs;
    std::lock_guard<std::mutex> lk(mu);
    if (std::find(ok_devs.begin(), ok_devs.end(), dev) != ok_devs.end()) return DSSSP_OK;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        (void)hipGetLastError();
        return fail(DSSSP_ERR_NODEV, "no HIP device visible");
    }
    if (dev < 0 || dev >= count) return fail(DSSSP_ERR_NODEV, "device %d not present (%d visible)", dev, count);
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(DSSSP_ERR_NODEV, "device %d is %s; this library is built for gfx950 only", dev, prop.gcnArchName);
    ok_devs.push_back(dev);
    return DSSSP_OK;
}

inline uint32_t words_of(uint64_t bits) { return (uint32_t)((bits + 63) / 64); }

constexpr uint32_t INIT_DCAP = 4;  // initial delta block capacity of every row
constexpr int64_t LOCAL_MAX = 32;  // largest continuation budget of a round (tiles)
constexpr int64_t ROUND_MARGIN = 1;  // rounds a loop's first group adds to the last loop's rounds with input
// The full edge pass: rows longer than SWEEP_PIECE entries go to the item
// launch in pieces of that size; both launches have (at least) as many
// workgroups as the device holds at once at 6 waves per SIMD (24 waves per CU).
// Measured on scale-24 R-MAT (pass ms): pieces of 256 3.77, 512 3.71, 768 3.56,
// 1024 3.57, 1280 3.62, 1536 3.73, 2048 4.31; with pieces of 1024, a tile
// budget of 512 entries 3.95, 1024 (the tuning default) 3.68, 2048 4.92.
constexpr uint32_t SWEEP_PIECE = 1024;
constexpr int SWEEP_WAVES_PER_CU = 24;
// A tree built from scratch starts with full edge passes while a pass still
// lowers at least 1/SCRATCH_PASS_SHARE of the slots (low-diameter graphs),
// at most SCRATCH_PASSES, then relaxation rounds finish it.
constexpr int SCRATCH_PASSES = 4;
constexpr uint64_t SCRATCH_PASS_SHARE = 8;
// Automatic invalidation rescue: candidates are rows of at least
// RESCUE_AUTO_DEGREE entries, in batches with at least n / RESCUE_AUTO_SHARE
// deletions.  Measured on the scale-24 R-MAT graph (MI355X): with 10^7
// deletions 15 % faster per update, with 5 * 10^6 6 %, with 10^6 or fewer a
// few percent slower (the stages' launches cost more than the work they save).
constexpr uint32_t RESCUE_AUTO_DEGREE = 1024;
constexpr uint64_t RESCUE_AUTO_SHARE = 8;
// Automatic ordered relaxation: trees whose distances span more than DEEP_STEPS
// steps, or whose ordered construction took more than DEEP_ROUNDS rounds.
constexpr uint32_t DEEP_STEPS = 64;
constexpr int64_t DEEP_ROUNDS = 512;

}  // namespace

// ---------------------------------------------------------------------------
// Objects.
// ---------------------------------------------------------------------------

struct dsssp_graph {
    int device = 0;
    hipStream_t stream = nullptr;
    int64_t n = 0;
    uint32_t nw = 0;  // words of a vertex bitmap
    uint64_t E = 0;   // base entries
    VRec *vr = nullptr;  // per-vertex records
    uint64_t *adj = nullptr, *pool = nullptr;
    uint64_t pool_cap = 0;
    uint64_t pool_top = 0;       // host copy, valid between updates
    uint64_t delta_entries = 0;  // sum of delta row lengths
    uint64_t base_entries = 0;   // sum of base row lengths (base rows shrink at a repack)
    uint64_t dead_entries = 0;   // tombstoned entries still in rows (since the last repack; duplicate
                                 // deletion records in one batch may count a copy twice)
    int64_t repacks = 0;
    bool emap_stale = false;     // a repack moved entries and the map could not be rebuilt yet
    // Merge check of served requests: hash table and request starts.
    unsigned long long *mkeys = nullptr;
    uint32_t *mreq = nullptr;
    uint64_t mslots = 0;
    uint64_t *mfirst = nullptr;
    int32_t mfirst_cap = 0;
    uint64_t *tbm = nullptr;  // rows whose delta block grows in this batch
    uint64_t *emap = nullptr;  // edge map (fingerprint << 32 | position in row)
    uint64_t emap_slots = 0;
    uint64_t emap_used = 0;    // entries entered since the map was built (dead ones too)
    uint32_t *wsum = nullptr;   // per bitmap word: pool entries needed by the batch
    uint64_t *bsum = nullptr;   // per 256 words: sum of wsum
    uint64_t *bbase = nullptr;  // per 256 words: pool offset (only past INS_FOLD_GROUPS groups)
    GraphCtl *gc = nullptr;
    GraphCtl *gc_host = nullptr;  // pinned
    bool zero_w = false;          // some weight was ever 0: parents need the tight-BFS repair
    uint32_t auto_delta = 1;      // step of ordered relaxation: twice the mean input weight
    // ev[0..6]: update phases; ev[7]: upkeep start; ev[8..9]: host batch copy.
    // Timing events only, created without the system-scope fence (no cache
    // writeback and invalidation at each record); nothing waits on them.
    hipEvent_t ev[10] = {};
    std::vector<hipEvent_t> kev;  // per-launch timing events (kernel_timing)
    // Trees and batches attached to this graph.  dsssp_graph_destroy while any
    // remain only marks the graph; the last detach releases it, so objects may
    // be destroyed in any order.
    int attached = 0;
    bool destroy_pending = false;
    std::mutex mu;
    // dsssp_update from host records reuses one device batch and one pinned
    // staging buffer (grown on demand) instead of allocating per call.
    dsssp_batch *upd = nullptr;
    dsssp_change *stage = nullptr;
    int64_t stage_cap = 0;

    GView view() const { return GView{vr, adj, pool, emap, emap_slots - 1}; }
};

struct dsssp_tree {
    dsssp_graph *g = nullptr;
    int32_t src = -1;
    int32_t ntrees = 1;   // trees sharing the graph (a forest of K sources)
    uint32_t kshift = 0;  // slot = vertex << kshift | tree
    uint64_t nslots = 0;  // n << kshift
    // Distances span many steps of the graph's auto_delta (a high-diameter
    // graph): relaxation is distance-ordered when the tuning leaves it to the
    // library.
    bool deep = false;
    int64_t last_rounds[8] = {};  // rounds the last loop of each mode took
    uint32_t nw = 0;      // words of a slot bitmap
    uint64_t *st = nullptr;
    uint64_t *fb[2] = {nullptr, nullptr};  // frontier bitmaps
    uint64_t *ib = nullptr;                // slots invalidated by the current update
    uint64_t *hq[2] = {nullptr, nullptr};
    uint32_t hcap = 0;
    uint32_t *lvl = nullptr, *best = nullptr;
    bool best_clean = false;   // best[] is all ~0 (the rescue stage's precondition)
    uint64_t *cb = nullptr;    // rescue candidates of the current stage
    unsigned long long *bstats = nullptr;
    Counters *c = nullptr;
    Counters *c_host = nullptr;  // pinned
    // Captured groups of rounds (hipGraph), keyed by mode, first round index
    // mod 6, group size, grid and the exact base arguments.
    struct Captured {
        int mode, phase, group, grid;
        FArgs args;
        hipGraphExec_t exec;
    };
    std::vector<Captured> graphs;
};

struct dsssp_batch {
    dsssp_graph *g = nullptr;
    int64_t k = 0, n_ins = 0, n_del = 0;
    bool zero_w = false;
    int4 *rec = nullptr;
    uint32_t *tslot = nullptr;  // per record copy: position in its row's delta block
    int64_t cap = 0;            // records rec / tslot hold
};

namespace {

struct DevGuard {
    explicit DevGuard(int d) { (void)hipSetDevice(d); }
};

FArgs base_args(dsssp_tree *t)
{
    const dsssp_tuning tu = tuning();
    FArgs a;
    memset(&a, 0, sizeof a);  // padding included: captured graphs are keyed by these bytes
    a.g = t->g->view();
    a.st = t->st;
    a.lvl = t->lvl;
    a.best = t->best;
    a.inv_bm = t->ib;
    a.c = t->c;
    a.bstats = t->bstats;
    a.nslots = (uint32_t)t->nslots;
    a.nwords = t->nw;
    a.hcap = t->hcap;
    a.tiles = 1;
    a.consume = 1;
    a.heavy = (uint32_t)tu.heavy_degree;
    a.chunk = (uint32_t)tu.chunk;
    a.kshift = t->kshift;
    a.ept = (uint32_t)tu.edges_per_thread;
    a.budget = (uint32_t)tu.tile_entries;
    a.local = tu.local_tiles > 0 ? (uint32_t)tu.local_tiles : 0u;
    a.delta = 0;  // set for the relaxation rounds and the seeding before them
    a.block = (uint32_t)tu.block;
    return a;
}

// Workgroups of a frontier launch over a bitmap of nwords words: the tuning's
// grid, or enough workgroups to read the whole bitmap in one pass (a wave
// reads 64 words per pass), so that a nearly empty frontier of a large graph
// costs one round trip to memory rather than one per pass.
int frontier_grid(uint64_t nwords, uint32_t block)
{
    const uint64_t words_per_wg = 64 * (uint64_t)(block / 64);
    const uint64_t one_pass = (nwords + words_per_wg - 1) / words_per_wg;
    return (int)std::max<uint64_t>((uint64_t)std::max(1, tuning().grid), std::min<uint64_t>(one_pass, 4096));
}

// Step of the distance-ordered relaxation of tree t (0: off): the tuning's,
// or in automatic mode the graph's step for deep trees.
uint32_t relax_delta(const dsssp_tree *t)
{
    const int32_t d = tuning().delta;
    if (d == DSSSP_DELTA_AUTO) return t->deep ? t->g->auto_delta : 0u;
    return d > 0 ? (uint32_t)d : 0u;
}

// Arguments of round i of a loop phase; the producer stage before the loop
// wrote bitmap 0, flag 0 and chunk count 0.
FArgs round_args(dsssp_tree *t, const FArgs &base, int64_t i)
{
    FArgs a = base;
    const int fi = (int)(i & 1), ci = (int)(i % 3);
    a.in_bm = t->fb[fi];
    a.in_act = &t->c->act[ci];
    a.in_hq = t->hq[fi];
    a.in_hcnt = &t->c->hcnt[ci];
    a.out_bm = t->fb[fi ^ 1];
    a.out_act = &t->c->act[(ci + 1) % 3];
    a.out_hq = t->hq[fi ^ 1];
    a.out_hcnt = &t->c->hcnt[(ci + 1) % 3];
    a.zero_act = &t->c->act[(ci + 2) % 3];
    a.zero_hcnt = &t->c->hcnt[(ci + 2) % 3];
    a.ring = (uint32_t)ci;
    a.count_round = 1;
    // Deep updates (high-diameter graphs) run many rounds of small frontiers:
    // the continuation budget doubles every 16 rounds, up to LOCAL_MAX tiles.
    if (a.local) a.local = (uint32_t)std::min<int64_t>(LOCAL_MAX, (int64_t)a.local << std::min<int64_t>(i / 16, 8));
    return a;
}

// Clears the round flags, chunk counts and threshold ring (act .. aux_hcnt)
// before a producer.
void zero_ring(dsssp_tree *t)
{
    static_assert(offsetof(Counters, err) == offsetof(Counters, aux_hcnt) + sizeof(uint32_t), "aux_hcnt before err");
    launch_fill_u32(&t->c->act[0], 0u, offsetof(Counters, err) / sizeof(uint32_t), t->g->stream);
}

// One pass of a single-pass mode (seed, parent pick) into bitmap 0 / flag 0.
// Rows longer than the heavy threshold are queued as chunks on list 1 (idle
// here) by the first launch and processed by a second, chunk-only launch.
// ring_clear: the caller has just cleared the ring (zero_ring), chunk count
// included.
void single_pass(dsssp_tree *t, int mode, FArgs a, int min_grid = 0, uint64_t *out_bm = nullptr,
                 bool ring_clear = false)
{
    hipStream_t s = t->g->stream;
    const int grid = std::max(frontier_grid(t->nw, a.block), min_grid);
    if (!ring_clear) launch_fill_u32(&t->c->aux_hcnt, 0u, 1, s);
    a.ring = 2;  // output ring slot 0, read by round 0 of the loop that follows
    a.out_bm = out_bm ? out_bm : t->fb[0];
    a.out_act = &t->c->act[0];
    a.out_hq = t->hq[1];
    a.out_hcnt = &t->c->aux_hcnt;
    a.in_hq = nullptr;
    a.in_hcnt = nullptr;
    launch_frontier(mode, a, grid, s);
    FArgs b = a;
    b.tiles = 0;
    b.in_hq = t->hq[1];
    b.in_hcnt = &t->c->aux_hcnt;
    b.out_hq = nullptr;
    b.out_hcnt = nullptr;
    launch_frontier(mode, b, grid, s);
}

// Arguments of the full edge pass (M_SWEEP) and its least grid.
FArgs sweep_args(dsssp_tree *t, int &grid)
{
    FArgs a = base_args(t);
    a.identity = 1;
    a.consume = 0;
    a.local = 0;
    a.delta = 0;
    a.heavy = a.chunk = SWEEP_PIECE;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, t->g->device) != hipSuccess) cus = 0;
    grid = cus * SWEEP_WAVES_PER_CU / (int)(a.block / 64);
    return a;
}

// Full edge passes at the start of a tree built from scratch (the roots are
// marked in bitmap 0), while a pass lowers at least 1/SCRATCH_PASS_SHARE of
// the slots.  A pass marks the slots it lowers in bitmap 1, which is then
// counted (distinct slots, not improvements) and merged into bitmap 0, so
// relaxation rounds started from bitmap 0 reach the same fixed point as
// without the passes.
int scratch_passes(dsssp_tree *t, int &passes)
{
    passes = 0;
    if (t->kshift >= 6) return DSSSP_OK;
    hipStream_t s = t->g->stream;
    int grid = 0;
    const FArgs a = sweep_args(t, grid);
    for (int k = 0; k < SCRATCH_PASSES; ++k) {
        single_pass(t, M_SWEEP, a, grid, t->fb[1]);
        ++passes;
        launch_merge_count(t->fb[0], t->fb[1], t->nw, &t->c->popc, s);
        HIP_TRY(hipMemcpyAsync(&t->c_host->popc, &t->c->popc, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (t->c_host->popc * SCRATCH_PASS_SHARE < t->nslots) break;
    }
    return DSSSP_OK;
}

// Instantiated hipGraph of rounds first .. first+group-1 of `mode`.  Round
// arguments depend on the round index mod 6 and, through the continuation
// budget, on the index itself below round 128.
int captured_rounds(dsssp_tree *t, int mode, int64_t first, int group, int grid, const FArgs &base,
                    hipGraphExec_t *out)
{
    const int phase = first < 128 ? (int)first + 6 : (int)(first % 6);
    for (auto &e : t->graphs)
        if (e.mode == mode && e.phase == phase && e.group == group && e.grid == grid &&
            !memcmp(&e.args, &base, sizeof base)) {
            *out = e.exec;
            return DSSSP_OK;
        }
    if (t->graphs.size() >= 32) {
        (void)hipGraphExecDestroy(t->graphs.front().exec);
        t->graphs.erase(t->graphs.begin());
    }
    hipStream_t s = t->g->stream;
    hipGraph_t gr = nullptr;
    HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < group; ++i) launch_frontier(mode, round_args(t, base, first + i), grid, s);
    HIP_TRY(hipStreamEndCapture(s, &gr));
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) return fail(DSSSP_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    t->graphs.push_back({mode, phase, group, grid, base, exec});
    *out = exec;
    return DSSSP_OK;
}

void rescue_stage(dsssp_tree *t, int64_t j);

// Runs rounds of `mode` until a round marks nothing and queues no chunk.
// Rounds run back to back on the stream (replayed from captured hipGraphs when
// use_graph is set); the host reads one flag pair per group of rounds.  A round
// after convergence reads an empty bitmap and is a no-op.  With rescue_levels
// (invalidation only), rounds 0 .. rescue_levels - 2 leave rescue candidates
// and each is followed by its rescue stage.
// A tail, when given, is enqueued after every group of rounds in place of the
// flag copy and must copy at least the leading act / hcnt / worked words of
// the counters to the host (the caller's end-of-update copies, so that the
// last convergence check and the update's final synchronisation are one).
int run_rounds(dsssp_tree *t, int mode, int64_t &rounds, double *kernel_ms, bool full_first = false,
               bool no_alloc = false, int rescue_levels = 0, uint32_t rescue_deg = 0,
               const std::function<int()> *tail = nullptr)
{
    dsssp_graph *g = t->g;
    hipStream_t s = g->stream;
    const dsssp_tuning tu = tuning();
    const int max_group = 64;
    // The first group covers as many rounds as the last loop of this mode had
    // rounds with input, plus ROUND_MARGIN: a loop whose last round with input
    // marks nothing is over after that round, so a loop as deep as the last
    // one needs one convergence check and few empty rounds (each ~4 us).
    // Later groups start again from round_batch.
    const int rb = std::max(1, tu.round_batch);
    const int first = t->last_rounds[mode] > 0
                          ? (int)std::min<int64_t>(t->last_rounds[mode] + ROUND_MARGIN, max_group)
                          : rb;
    int group = first;
    FArgs base = base_args(t);
    const int grid = frontier_grid(t->nw, base.block);
    // (no per-launch events or captured graphs for the allocation-free update)
    const bool timed = tu.kernel_timing && kernel_ms && !no_alloc;
    const bool graphs = tu.use_graph && !no_alloc;
    if (mode == M_RELAX) base.delta = relax_delta(t);
    // Push-vs-full-sweep (not with distance ordering, whose rounds visit only
    // slots below a threshold, nor on the tree-vector path of 64-tree forests).
    if (mode == M_RELAX && !base.delta && t->kshift < 6 && tu.sweep > 0)
        base.sweep = (uint32_t)std::max<uint64_t>(1, t->nslots * (uint64_t)tu.sweep / 1000);
    // Round i's arguments; with full_first, round 0 relaxes every slot.
    auto args_of = [&](int64_t r) {
        FArgs a = round_args(t, base, r);
        if (full_first && r == 0) {
            a.identity = 1;
            a.in_act = nullptr;
        }
        return a;
    };
    // Round r, then its rescue stage when it leaves candidates.
    if (rescue_levels <= 0) rescue_deg = 0;
    auto launch_round = [&](int64_t r, hipEvent_t e0, hipEvent_t e1) {
        FArgs a = args_of(r);
        const bool stage = rescue_deg && r + 1 < rescue_levels;
        if (stage) {
            a.rescue_deg = rescue_deg;
            a.cand_bm = t->cb;
            a.cand_flag = &t->c->cand[(r + 1) & 1];
        }
        launch_frontier(stage ? M_INVAL_RESCUE : mode, a, grid, s, e0, e1);
        if (stage) rescue_stage(t, r + 1);
    };
    // Every round settles at least one more hop of every shortest path (of
    // every tree at once), so a loop needs at most n rounds; each threshold move
    // of ordered relaxation follows a round that visited nothing and settles at
    // least the smallest marked slot, so it needs at most about twice that.
    const int64_t max_rounds = (base.delta ? 3 : 1) * g->n + 64;
    // A loop that runs longer than the give-up time is abandoned (every kernel
    // has returned by the time the host looks); the caller scrubs the tree.
    const auto t_start = std::chrono::steady_clock::now();
    const double give_up_ms = tu.give_up_ms > 0 ? (double)tu.give_up_ms : 120000.0;
    int64_t i = 0;
    for (;;) {
        if (timed) {
            if (g->kev.size() < (size_t)(2 * group)) {
                const size_t old = g->kev.size();
                g->kev.resize((size_t)(2 * group));
                for (size_t j = old; j < g->kev.size(); ++j)
                    HIP_TRY(hipEventCreateWithFlags(&g->kev[j], hipEventDisableSystemFence));
            }
            for (int j = 0; j < group; ++j) launch_round(i + j, g->kev[2 * j], g->kev[2 * j + 1]);
        } else if (graphs && !(full_first && i == 0) && !(rescue_deg && i + 1 < rescue_levels)) {
            hipGraphExec_t exec = nullptr;
            if (int e = captured_rounds(t, mode, i, group, grid, base, &exec)) return e;
            HIP_TRY(hipGraphLaunch(exec, s));
        } else {
            for (int j = 0; j < group; ++j) launch_round(i + j, nullptr, nullptr);
        }
        i += group;
        rounds += group;
        HIP_TRY(hipGetLastError());
        const int ci = (int)(i % 3);  // flag / chunk slot written by the last round
        // act, hcnt and worked in one copy (they lead Counters)
        static_assert(offsetof(Counters, hcnt) == 3 * sizeof(uint32_t) &&
                          offsetof(Counters, worked) == 6 * sizeof(uint32_t), "act, hcnt, worked lead Counters");
        if (tail) {
            if (int e = (*tail)()) return e;
        } else {
            HIP_TRY(hipMemcpyAsync(t->c_host, t->c, 7 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        }
        HIP_TRY(hipStreamSynchronize(s));
        if (timed) {
            for (int j = 0; j < group; ++j) {
                float ms = 0;
                HIP_TRY(hipEventElapsedTime(&ms, g->kev[2 * j], g->kev[2 * j + 1]));
                *kernel_ms += ms;
            }
        }
        if (t->c_host->act[ci] == 0 && t->c_host->hcnt[ci] == 0) break;
        // More rounds than max_rounds can only come from an internal error.
        if (i > max_rounds) return fail(DSSSP_ERR_STATE, "frontier did not empty after %lld rounds", (long long)i);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
        if (ms > give_up_ms)
            return fail(DSSSP_ERR_STATE, "frontier did not empty within %.0f ms (%lld rounds); gave up", give_up_ms,
                        (long long)i);
        group = i == first && first < max_group ? rb : std::min(group * 2, max_group);
    }
    t->last_rounds[mode] = t->c_host->worked;
    return DSSSP_OK;
}

int ensure_pick_buffers(dsssp_tree *t)
{
    if (!t->lvl) {
        if (int s = dev_alloc((void **)&t->lvl, sizeof(uint32_t) * (size_t)t->nslots, "levels")) return s;
    }
    if (!t->best) {
        if (int s = dev_alloc((void **)&t->best, sizeof(uint32_t) * (size_t)t->nslots, "parent picks")) return s;
    }
    return DSSSP_OK;
}

// Recomputes every parent: smallest-id tight neighbour (use_lvl = 0), or the
// smallest-id tight neighbour one BFS level closer to a root in the tight graph
// (use_lvl = 1), which stays acyclic with zero weights.
int repick_parents(dsssp_tree *t, bool use_lvl, int64_t &rounds, bool no_alloc = false)
{
    dsssp_graph *g = t->g;
    if (int s = ensure_pick_buffers(t)) return s;
    const uint32_t ns = (uint32_t)t->nslots;
    if (use_lvl) {
        zero_ring(t);
        launch_lvl_init(ns, t->kshift, t->st, t->lvl, t->best, t->fb[0], &t->c->act[0], g->stream);
        if (int s = run_rounds(t, M_TBFS, rounds, nullptr, false, no_alloc)) return s;
    } else {
        launch_fill_u32(t->best, 0xFFFFFFFFu, ns, g->stream);
    }
    FArgs a = base_args(t);
    a.identity = 1;
    a.consume = 0;
    a.use_lvl = use_lvl ? 1u : 0u;
    zero_ring(t);
    single_pass(t, M_PICK, a, 0, nullptr, true);
    launch_apply_best(ns, t->kshift, t->st, t->best, t->c, g->stream);
    t->best_clean = false;
    HIP_TRY(hipGetLastError());
    return DSSSP_OK;
}

// Buffers of the invalidation rescue: the candidate bitmap and the parent
// picks, all ~0 between stages (the commit resets what it used).  Without
// them (an allocation-free update before any other update) rescue stays off.
bool rescue_ready(dsssp_tree *t, bool no_alloc)
{
    hipStream_t s = t->g->stream;
    // (an allocation that fails only leaves the rescue off: its message is dropped)
    if (!t->cb) {
        if (no_alloc) return false;
        if (dev_alloc((void **)&t->cb, sizeof(uint64_t) * t->nw, "rescue candidates")) {
            clear_error();
            return false;
        }
        launch_fill_u64(t->cb, 0ull, t->nw, s);
    }
    if (!t->best) {
        if (no_alloc) return false;
        if (dev_alloc((void **)&t->best, sizeof(uint32_t) * (size_t)t->nslots, "parent picks")) {
            clear_error();
            return false;
        }
        t->best_clean = false;
    }
    if (!t->best_clean) {
        launch_fill_u32(t->best, 0xFFFFFFFFu, t->nslots, s);
        t->best_clean = true;
    }
    return true;
}

// Rescue stage j of an invalidation loop (stage 0 follows the classification,
// stage j > 0 follows round j - 1, whose candidates raised c->cand[j & 1]).
// Every candidate in t->cb still holds its distance; the parent-pick pass
// (M_PICK without levels: the smallest-id live neighbour y with d[y] + w equal
// to that distance, y not reset) writes best, and k_rescue_commit keeps each
// candidate under its pick or resets it into round j's input bitmap and flag.
// All launches return at once when the flag is clear.  The pick pass's split
// rows use the item list round j - 1 read, free until round j writes it.
void rescue_stage(dsssp_tree *t, int64_t j)
{
    hipStream_t s = t->g->stream;
    FArgs a = base_args(t);
    a.in_bm = t->cb;
    a.consume = 0;
    a.use_lvl = 0;
    a.in_act = &t->c->cand[j & 1];
    a.out_hq = t->hq[(j & 1) ^ 1];
    a.out_hcnt = &t->c->aux_hcnt;
    const int grid = frontier_grid(t->nw, a.block);
    launch_frontier(M_PICK, a, grid, s);
    FArgs b = a;
    b.tiles = 0;
    b.in_hq = a.out_hq;
    b.in_hcnt = &t->c->aux_hcnt;
    b.out_hq = nullptr;
    b.out_hcnt = nullptr;
    launch_frontier(M_PICK, b, grid, s);
    launch_rescue_commit(t->cb, t->nw, t->st, t->best, t->ib, t->fb[j & 1], &t->c->act[j % 3], &t->c->cand[j & 1],
                         &t->c->cand[(j + 1) & 1], &t->c->aux_hcnt, t->bstats, s);
}

// Item lists of the frontier kernel.  A launch leaves ceil(deg / chunk) items
// per deferred slot (every slot is in a launch's input at most once): the slots
// with deg > heavy entries (at most entries / heavy of them per tree) and, per
// tile that passes its budget of entries, at most 63 others (so at most
// 64 * entries / budget per tree, and at most n).
int ensure_hq(dsssp_tree *t, uint64_t total_entries)
{
    const dsssp_tuning tu = tuning();
    const uint64_t n = (uint64_t)t->g->n;
    const uint64_t per_tree = total_entries / (uint64_t)std::max(64, tu.chunk) +
                              total_entries / (uint64_t)std::max(64, tu.heavy_degree) +
                              std::min<uint64_t>(n, 64 * total_entries / (uint64_t)std::max(64, tu.tile_entries)) + 1;
    const uint64_t need = (per_tree << t->kshift) + 1024;
    if (need > 0xFFFFFFFFull) return fail(DSSSP_ERR_ARG, "item list would exceed 2^32 entries; raise chunk");
    if (t->hcap >= need) return DSSSP_OK;
    dev_free(t->hq[0]);
    dev_free(t->hq[1]);
    t->hcap = 0;
    const uint64_t cap = std::min<uint64_t>(need + need / 4, 0xFFFFFFFFull);
    if (int s = dev_alloc((void **)&t->hq[0], 16 * cap, "item list")) return s;
    if (int s = dev_alloc((void **)&t->hq[1], 16 * cap, "item list")) return s;
    t->hcap = (uint32_t)cap;
    return DSSSP_OK;
}

int build_emap(dsssp_graph *g, uint64_t entries, uint64_t extra);
int ensure_emap(dsssp_graph *g, uint64_t add);

// Repack of the resident CSR on the device, in the graph's stream: every row's
// tombstoned entries are dropped (live base entries moved to the front of the
// row, in order; live delta entries copied to a fresh block of a new pool) and
// the edge map is rebuilt.  Room is left in the new pool for n_ins more
// insertions under the growth bound of the insertion kernels (a batch allocates
// at most 2 x live capacity + 10 x n_ins entries).
int repack(dsssp_graph *g, uint64_t n_ins)
{
    hipStream_t s = g->stream;
    const uint64_t n = (uint64_t)g->n;
    const uint64_t add = 10 * n_ins;
    // every new block holds max(INIT_DCAP, live delta entries of its row)
    const uint64_t ub = INIT_DCAP * n + g->delta_entries;
    const uint64_t ncap = std::min<uint64_t>(std::max<uint64_t>(g->pool_cap, 3 * ub + add + (1u << 16)), 0xFFFFFFFFull);
    const uint64_t nb = (n + 255) / 256;
    uint32_t *cap = nullptr;
    uint64_t *bsum = nullptr, *bbase = nullptr, *npool = nullptr;
    int st = dev_alloc((void **)&npool, sizeof(uint64_t) * ncap, "delta pool");
    if (!st) st = dev_alloc((void **)&cap, sizeof(uint32_t) * n, "repack capacities");
    if (!st) st = dev_alloc((void **)&bsum, sizeof(uint64_t) * nb, "repack sums");
    if (!st) st = dev_alloc((void **)&bbase, sizeof(uint64_t) * nb, "repack offsets");
    if (st) {
        dev_free(npool);
        dev_free(cap);
        dev_free(bsum);
        dev_free(bbase);
        return st;
    }
    // From here on entries move: nothing below can leave the graph half done
    // except the edge map, which is then rebuilt before the next use.
    g->emap_stale = true;
    launch_fill_u64((uint64_t *)&g->gc->live[0], 0ull, 3, s);  // live[0], live[1], total
    launch_repack_rows(g->vr, g->adj, g->pool, n, INIT_DCAP, cap, g->gc, s);
    launch_repack_scan(cap, n, bsum, bbase, &g->gc->total, s);
    launch_repack_delta(g->vr, g->pool, npool, n, cap, bbase, s);
    const hipError_t e1 = hipMemcpyAsync(g->gc_host, g->gc, sizeof(GraphCtl), hipMemcpyDeviceToHost, s);
    const hipError_t e2 = hipStreamSynchronize(s);
    dev_free(cap);
    dev_free(bsum);
    dev_free(bbase);
    if (e1 != hipSuccess || e2 != hipSuccess) {
        dev_free(npool);
        return fail(DSSSP_ERR_HIP, "repack failed: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    }
    const uint64_t total = g->gc_host->total;
    dev_free(g->pool);
    g->pool = npool;
    g->pool_cap = ncap;
    g->pool_top = total;
    g->gc_host->pool_top = total;
    launch_fill_u64((uint64_t *)&g->gc->pool_top, total, 1, s);
    g->base_entries = g->gc_host->live[0];
    g->delta_entries = g->gc_host->live[1];
    g->dead_entries = 0;
    ++g->repacks;
    if (3 * total + add > ncap) return fail(DSSSP_ERR_NOMEM, "delta pool would exceed 2^32 entries");
    const uint64_t live = g->base_entries + g->delta_entries;
    if (int e = build_emap(g, live, live / 2 + 2 * n_ins)) return e;
    g->emap_stale = false;
    return DSSSP_OK;
}

// Upkeep before a batch with n_ins insertions: a repack when the delta pool's
// growth bound for the batch does not fit, or when tombstones are more than
// 1/8 of the entries that row scans read; then room in the edge map.
int ensure_graph(dsssp_graph *g, uint64_t n_ins)
{
    const bool pool = n_ins > 0 && 3 * g->pool_top + 10 * n_ins > g->pool_cap;
    const bool dead = g->dead_entries * 8 > g->base_entries + g->delta_entries;
    if (pool || dead || g->emap_stale) {
        if (g->emap_stale && !pool && !dead) {
            const uint64_t live = g->base_entries + g->delta_entries;
            if (int e = build_emap(g, live, live / 2 + 2 * n_ins)) return e;
            g->emap_stale = false;
        } else if (int e = repack(g, n_ins)) {
            return e;
        }
    }
    return n_ins > 0 ? ensure_emap(g, 2 * n_ins) : DSSSP_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return ms;
}

int finish_errors(dsssp_tree *t, uint32_t gerr)
{
    const uint32_t err = t->c_host->err | gerr;
    if (err & ERR_DIST_OVERFLOW) return fail(DSSSP_ERR_OVERFLOW, "a distance reached 2^32-1; the tree is not exact");
    if (err & ERR_POOL) return fail(DSSSP_ERR_STATE, "delta pool overflow (internal)");
    if (err & ERR_QUEUE) return fail(DSSSP_ERR_STATE, "item list overflow (internal)");
    if (err & ERR_NO_TIGHT) return fail(DSSSP_ERR_TREE, "a reachable vertex has no tight in-edge: distances not exact");
    return DSSSP_OK;
}

// Decides a suspected overflow (a candidate of 2^32-1 or more offered to an
// unreachable slot): it is one iff some reachable slot still has a live edge to
// an unreachable slot.  Runs only when the suspicion was raised; refreshes the
// host copy of the counters.
int confirm_overflow(dsssp_tree *t)
{
    const uint32_t err = t->c_host->err;
    if (!(err & ERR_DIST_SUSPECT) || (err & ERR_DIST_OVERFLOW)) return DSSSP_OK;
    hipStream_t s = t->g->stream;
    FArgs a = base_args(t);
    a.identity = 1;
    a.consume = 0;
    zero_ring(t);
    single_pass(t, M_OFCHECK, a, 0, nullptr, true);
    launch_reduce_stats(t->c, t->bstats, s);
    HIP_TRY(hipMemcpyAsync(t->c_host, t->c, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return DSSSP_OK;
}

// After a failed update or setup: every bitmap, flag and insertion counter is
// cleared so later calls start from a consistent state.
void scrub(dsssp_tree *t)
{
    hipStream_t s = t->g->stream;
    launch_fill_u64(t->fb[0], 0ull, t->nw, s);
    launch_fill_u64(t->fb[1], 0ull, t->nw, s);
    launch_fill_u64(t->ib, 0ull, t->nw, s);
    if (t->cb) launch_fill_u64(t->cb, 0ull, t->nw, s);
    t->best_clean = false;  // an interrupted rescue stage may have left picks
    launch_fill_u64(t->g->tbm, 0ull, t->g->nw, s);
    zero_ring(t);
    (void)hipStreamSynchronize(s);
}

// (Re)builds the edge map for `entries` directed entries plus headroom for
// `extra` more: every live entry of every row is entered by an identity pass of
// the frontier kernel, whose long rows and over-budget tiles go to a temporary
// item list that a second launch spreads over all waves.
int build_emap(dsssp_graph *g, uint64_t entries, uint64_t extra)
{
    const uint64_t slots = emap_slots_for(entries + extra);
    if (slots != g->emap_slots) {
        dev_free(g->emap);
        g->emap_slots = 0;
        if (int s = dev_alloc((void **)&g->emap, 8 * slots, "edge map")) return s;
        g->emap_slots = slots;
    }
    launch_fill_u64(g->emap, ~0ull, slots, g->stream);
    const uint32_t heavy = 1024, chunk = 1024, budget = 4096;
    const uint64_t E = g->E + g->delta_entries;
    const uint64_t items = 2 * E / chunk + E / heavy + 63 * E / budget + 1024;
    if (items > 0xFFFFFFFFull) return fail(DSSSP_ERR_ARG, "edge map build: item list would exceed 2^32 entries");
    uint64_t *list = nullptr;
    Counters *c = nullptr;
    if (int s = dev_alloc((void **)&list, 16 * items, "edge map items")) return s;
    if (int s = dev_alloc((void **)&c, sizeof(Counters), "edge map counters")) {
        dev_free(list);
        return s;
    }
    launch_fill_u32((uint32_t *)c, 0u, sizeof(Counters) / 4, g->stream);
    FArgs a;
    memset(&a, 0, sizeof a);
    a.g = g->view();
    a.c = c;
    a.nslots = (uint32_t)g->n;
    a.nwords = g->nw;
    a.hcap = (uint32_t)items;
    a.identity = 1;
    a.tiles = 1;
    a.heavy = heavy;
    a.chunk = chunk;
    a.budget = budget;
    a.ept = 4;
    a.out_hq = list;
    a.out_hcnt = &c->aux_hcnt;
    a.block = FRONTIER_BLOCK;
    const int grid = frontier_grid(a.nwords, a.block);
    launch_frontier(M_EMAP, a, grid, g->stream);
    FArgs b = a;
    b.tiles = 0;
    b.in_hq = list;
    b.in_hcnt = &c->aux_hcnt;
    b.out_hq = nullptr;
    b.out_hcnt = nullptr;
    launch_frontier(M_EMAP, b, grid, g->stream);
    Counters h;
    const hipError_t e1 = hipMemcpyAsync(&h, c, sizeof(Counters), hipMemcpyDeviceToHost, g->stream);
    const hipError_t e2 = hipStreamSynchronize(g->stream);
    dev_free(list);
    dev_free(c);
    if (e1 != hipSuccess || e2 != hipSuccess)
        return fail(DSSSP_ERR_HIP, "edge map build failed: %s", hipGetErrorString(e1 != hipSuccess ? e1 : e2));
    if (h.err & ERR_QUEUE) return fail(DSSSP_ERR_STATE, "edge map build: item list overflow (internal)");
    g->emap_used = entries;
    return DSSSP_OK;
}

// Keeps the edge map's load factor at most 0.7 with `add` more entries.
int ensure_emap(dsssp_graph *g, uint64_t add)
{
    if ((g->emap_used + add) * 10 <= g->emap_slots * 7) return DSSSP_OK;
    const uint64_t entries = g->base_entries + g->delta_entries;  // live entries are at most these
    return build_emap(g, entries, entries / 2 + add);
}

// Tombstones every copy of every deleted (a, b, w) in both rows.
void tombstone(dsssp_graph *g, dsssp_tree *t, dsssp_batch *b)
{
    launch_tomb(b->rec, b->k, g->view(), t->bstats, g->stream);
}

// Whether the update of a batch with n_ins insertions fits the graph's and
// tree's current buffers (the allocation-free update refuses otherwise).
int fits_without_upkeep(dsssp_graph *g, dsssp_tree *t, uint64_t n_ins, bool zero_w)
{
    if (g->emap_stale)
        return fail(DSSSP_ERR_STATE, "allocation-free update: the edge map must be rebuilt first "
                                     "(apply one update with dsssp_update_batch)");
    if (n_ins > 0 && 3 * g->pool_top + 10 * n_ins > g->pool_cap)
        return fail(DSSSP_ERR_STATE, "allocation-free update: the delta pool would have to grow for %llu insertions "
                                     "(apply one update with dsssp_update_batch first)", (unsigned long long)n_ins);
    if (n_ins > 0 && (g->emap_used + 2 * n_ins) * 10 > g->emap_slots * 7)
        return fail(DSSSP_ERR_STATE, "allocation-free update: the edge map would have to be rebuilt for %llu "
                                     "insertions", (unsigned long long)n_ins);
    const dsssp_tuning tu = tuning();
    const uint64_t total = g->E + g->delta_entries + 2 * n_ins;
    const uint64_t per_tree = total / (uint64_t)std::max(64, tu.chunk) + total / (uint64_t)std::max(64, tu.heavy_degree) +
                              std::min<uint64_t>((uint64_t)g->n, 64 * total / (uint64_t)std::max(64, tu.tile_entries)) + 1;
    if ((per_tree << t->kshift) + 1024 > t->hcap)
        return fail(DSSSP_ERR_STATE, "allocation-free update: the tree's item lists would have to grow");
    if ((g->zero_w || zero_w) && (!t->lvl || !t->best))
        return fail(DSSSP_ERR_STATE, "allocation-free update: zero weights need the parent-repair buffers "
                                     "(apply one update with dsssp_update_batch first)");
    return DSSSP_OK;
}

int update_impl(dsssp_graph *g, dsssp_tree *t, dsssp_batch *b, dsssp_stats *stats, bool h2d_timed,
                bool no_alloc = false)
{
    DevGuard dg(g->device);
    hipStream_t s = g->stream;
    const dsssp_tuning tu = tuning();
    // ev[7] and ev[6] bracket the update; with phase_timing, ev[0] .. ev[5]
    // mark the ends of its phases (each record holds the stream a few us).
    const bool phases = tu.phase_timing != 0;
    auto phase_mark = [&](int k) { return phases ? hipEventRecord(g->ev[k], s) : hipSuccess; };
    // Upkeep of the resident graph (pool room, edge map load, item lists) is
    // part of the update's time: ev[7] opens the window.
    HIP_TRY(hipEventRecord(g->ev[7], s));
    if (no_alloc) {
        if (int e = fits_without_upkeep(g, t, (uint64_t)b->n_ins, b->zero_w)) return e;
    } else {
        if (int e = ensure_graph(g, (uint64_t)b->n_ins)) return e;
        if (int e = ensure_hq(t, g->E + g->delta_entries + 2 * (uint64_t)b->n_ins)) return e;
    }
    int64_t inv_rounds = 0, relax_rounds = 0, repair_rounds = 0;
    double relax_kernel_ms = 0;

    HIP_TRY(phase_mark(0));
    // the tree's counters (ring included) and the graph's insertion error word
    launch_zero2_u32((uint32_t *)t->c, sizeof(Counters) / 4, &g->gc->err, b->n_ins > 0 ? 1 : 0, s);
    // 1. insertions into the resident CSR
    if (b->n_ins > 0) {
        launch_ins_place(b->rec, b->k, g->view(), b->tslot, g->tbm, s);
        launch_ins_wordsum(g->tbm, g->nw, g->vr, g->wsum, g->bsum, s);
        launch_ins_alloc(g->tbm, g->nw, g->wsum, g->bsum, g->bbase, g->vr, g->pool, g->pool_cap, g->gc, s);
        launch_ins_pending(b->rec, b->k, g->view(), b->tslot, g->gc, s);
        g->emap_used += 2 * (uint64_t)b->n_ins;
        if (b->zero_w) g->zero_w = true;
    }
    HIP_TRY(phase_mark(1));
    // 2-3. deletions: classify, then invalidate the subtrees.  With rescue, a
    // vertex with a long row is not reset at once: a rescue stage first looks
    // for another tight parent that still holds (positive weights only, so a
    // kept vertex's parent is strictly closer to the root; the tree-vector
    // layout of 64 trees resets as before).
    uint32_t rdeg = 0;
    if (tu.rescue_degree > 0)
        rdeg = (uint32_t)tu.rescue_degree;
    else if (tu.rescue_degree == DSSSP_RESCUE_AUTO && (uint64_t)b->n_del * RESCUE_AUTO_SHARE >= (uint64_t)g->n)
        rdeg = RESCUE_AUTO_DEGREE;
    const bool rescue = b->n_del > 0 && rdeg && !g->zero_w && t->kshift < 6 && rescue_ready(t, no_alloc);
    if (!rescue) rdeg = 0;
    int code = DSSSP_OK;
    if (b->n_del > 0) {
        // (the ring is still clear from the start of the update: the insertion
        // kernels touch only the graph)
        launch_classify(b->rec, b->k, t->kshift, t->st, t->fb[0], t->ib, &t->c->act[0], g->vr, rdeg, t->cb,
                        &t->c->cand[0], s);
        tombstone(g, t, b);
    }
    HIP_TRY(phase_mark(2));
    if (rescue) rescue_stage(t, 0);
    if (b->n_del > 0)
        code = run_rounds(t, M_INVAL, inv_rounds, nullptr, false, no_alloc, rescue ? tu.rescue_levels : 0, rdeg);
    HIP_TRY(phase_mark(3));
    // 4. seed invalidated slots from valid neighbours, offer inserted edges
    if (!code) {
        zero_ring(t);
        if (b->n_del > 0) {
            FArgs a = base_args(t);
            a.in_bm = t->ib;  // consumed, so clean for the next update
            a.delta = relax_delta(t);
            single_pass(t, M_SEED, a, 0, nullptr, true);
        }
        if (b->n_ins > 0)
            launch_ins_offer(b->rec, b->k, t->kshift, b->tslot, g->view(), t->st, t->fb[0], &t->c->act[0], t->c,
                             relax_delta(t) != 0, s);
    }
    HIP_TRY(phase_mark(4));
    // 5. relaxation.  Without the zero-weight repair after it, the update's
    // closing statistics and copies ride on each convergence check of the
    // relaxation loop, so the last check is also the update's last wait.
    auto closing = [&]() -> int {
        HIP_TRY(hipEventRecord(g->ev[6], s));
        launch_reduce_stats(t->c, t->bstats, s, g->gc);  // the graph's pool top and error word too
        HIP_TRY(hipMemcpyAsync(t->c_host, t->c, sizeof(Counters), hipMemcpyDeviceToHost, s));
        return DSSSP_OK;
    };
    const std::function<int()> relax_tail = [&]() -> int {
        HIP_TRY(phase_mark(5));
        return closing();
    };
    const bool merged_close = !g->zero_w;
    if (!code)
        code = run_rounds(t, M_RELAX, relax_rounds, &relax_kernel_ms, false, no_alloc, 0, 0,
                          merged_close ? &relax_tail : nullptr);
    if (code || !merged_close) {
        HIP_TRY(phase_mark(5));
        // 6. zero-weight parent repair
        if (!code && g->zero_w) code = repick_parents(t, true, repair_rounds, no_alloc);
        if (int e = closing()) return e;
        HIP_TRY(hipStreamSynchronize(s));
    }
    HIP_TRY(hipGetLastError());
    if (b->n_ins > 0) {
        g->pool_top = t->c_host->g_pool_top;
        g->delta_entries += 2 * (uint64_t)b->n_ins;
    }
    g->dead_entries += t->c_host->stats[ST_KILLS];
    if (!code) code = confirm_overflow(t);
    if (!code) code = finish_errors(t, b->n_ins > 0 ? t->c_host->g_err : 0u);
    if (code) {
        const std::string keep = last_error();
        scrub(t);
        return fail(code, "%s", keep.c_str());
    }
    if (stats) {
        const Counters &c = *t->c_host;
        dsssp_stats S{};
        S.n_insert = b->n_ins;
        S.n_delete = b->n_del;
        S.n_invalidated = (int64_t)c.stats[ST_INVALIDATED];
        S.inv_rounds = inv_rounds;
        S.relax_rounds = relax_rounds;
        S.relax_launches = relax_rounds;
        S.frontier_vertices = (int64_t)c.stats[ST_RELAX_FRONT];
        S.relax_edges = (int64_t)c.stats[ST_RELAX_EDGES];
        S.relax_improved = (int64_t)c.stats[ST_RELAX_MARKS];
        S.relax_enqueued = (int64_t)c.stats[ST_RELAX_MARKS];
        S.other_edges = (int64_t)c.stats[ST_OTHER_EDGES];
        S.probes = (int64_t)c.stats[ST_PROBES];
        S.t_h2d_ms = h2d_timed ? elapsed(g->ev[8], g->ev[9]) : 0.0;
        if (phases) {
            S.t_upkeep_ms = elapsed(g->ev[7], g->ev[0]);
            S.t_insert_ms = elapsed(g->ev[0], g->ev[1]);
            S.t_classify_ms = elapsed(g->ev[1], g->ev[2]);
            S.t_invalidate_ms = elapsed(g->ev[2], g->ev[3]);
            S.t_seed_ms = elapsed(g->ev[3], g->ev[4]);
            S.t_relax_ms = elapsed(g->ev[4], g->ev[5]) + elapsed(g->ev[5], g->ev[6]);
        }
        S.t_update_ms = elapsed(g->ev[7], g->ev[6]);
        S.t_relax_kernel_ms = relax_kernel_ms;
        // Algorithmic bytes of relaxation: per frontier slot 16 B row bounds +
        // 8 B delta descriptor + 8 B own state; per scanned entry 8 B entry +
        // 8 B neighbour state; per improvement 8 B state + 8 B bitmap word.
        // Each round reads (and clears) its input bitmap: 8 B per word.
        const double fr = (double)c.stats[ST_RELAX_FRONT], ed = (double)c.stats[ST_RELAX_EDGES];
        const double mk = (double)c.stats[ST_RELAX_MARKS];
        const double relax = 32.0 * fr + 16.0 * ed + 16.0 * mk + 8.0 * t->nw * (double)relax_rounds;
        const double other = 16.0 * c.stats[ST_OTHER_EDGES] + 2 * 32.0 * c.stats[ST_INVALIDATED] +
                             8.0 * c.stats[ST_PROBES] + 32.0 * b->n_del + 72.0 * b->n_ins +
                             8.0 * t->nw * (double)inv_rounds;
        S.alg_bytes_relax = relax;
        S.alg_bytes = relax + other;
        S.relax_sweeps = (int64_t)c.stats[ST_SWEEPS];
        S.n_rescued = (int64_t)c.stats[ST_RESCUED];
        *stats = S;
    }
    return DSSSP_OK;
}

int validate_changes(int64_t n, const dsssp_change *chg, int64_t k, int64_t &ni, int64_t &nd, bool &zero_w)
{
    ni = nd = 0;
    zero_w = false;
    for (int64_t i = 0; i < k; ++i) {
        const dsssp_change &c = chg[i];
        if (c.op != 0 && c.op != 1) continue;  // ignored, as readin_changes does (all_structure_undir.cuh:106,119)
        if (c.a < 0 || c.a >= n || c.b < 0 || c.b >= n)
            return fail(DSSSP_ERR_RANGE, "change %lld (%d %d %d %d): vertex outside [0,%lld)", (long long)i, c.a, c.b,
                        c.w, c.op, (long long)n);
        if (c.w < 0)
            return fail(DSSSP_ERR_WEIGHT, "change %lld (%d %d %d %d): negative weight", (long long)i, c.a, c.b, c.w,
                        c.op);
        if (c.op == 1) {
            ++ni;
            if (c.w == 0) zero_w = true;
        } else {
            ++nd;
        }
    }
    return DSSSP_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI.
// ---------------------------------------------------------------------------

extern "C" const char *dsssp_last_error(void) { return last_error(); }
extern "C" const char *dsssp_version(void) { return "dsssp-mi355x 0.2 (gfx950)"; }

static int scratch_of_device(int dev);

extern "C" int dsssp_set_device(int device)
{
    clear_error();
    if (int s = check_device(device)) return s;
    if (int s = scratch_of_device(device)) return s;  // the allocation-free update's scratch address
    t_device = device;
    return DSSSP_OK;
}

extern "C" int dsssp_get_tuning(dsssp_tuning *t)
{
    clear_error();
    if (!t) return fail(DSSSP_ERR_ARG, "null argument");
    *t = tuning();
    return DSSSP_OK;
}

extern "C" int dsssp_set_tuning(const dsssp_tuning *t)
{
    clear_error();
    if (!t) return fail(DSSSP_ERR_ARG, "null argument");
    std::lock_guard<std::mutex> lk(g_tuning_mu);
    dsssp_tuning nt = g_tuning;
    if (t->block) nt.block = t->block;
    if (t->grid) nt.grid = t->grid;
    if (t->heavy_degree) nt.heavy_degree = t->heavy_degree;
    if (t->chunk) nt.chunk = t->chunk;
    if (t->round_batch) nt.round_batch = t->round_batch;
    if (t->edges_per_thread) nt.edges_per_thread = t->edges_per_thread;
    if (t->tile_entries) nt.tile_entries = t->tile_entries;
    if (t->local_tiles) nt.local_tiles = t->local_tiles;
    if (t->delta) nt.delta = t->delta;
    if (t->sweep) nt.sweep = t->sweep;
    if (t->give_up_ms) nt.give_up_ms = t->give_up_ms;
    if (t->rescue_degree) nt.rescue_degree = t->rescue_degree;
    if (t->rescue_levels) nt.rescue_levels = t->rescue_levels;
    nt.kernel_timing = t->kernel_timing;
    nt.use_graph = t->use_graph;
    nt.phase_timing = t->phase_timing;
    if (nt.grid < 1 || nt.grid > 65536 || nt.heavy_degree < 64 || nt.chunk < 64 || nt.chunk % 64 ||
        nt.round_batch < 1 || nt.round_batch > 64 || nt.tile_entries < 64 || nt.local_tiles < -1 ||
        nt.local_tiles > 1024 || nt.delta < DSSSP_DELTA_AUTO || nt.sweep < -1 || nt.sweep > 1000 || nt.give_up_ms < 1 ||
        nt.rescue_degree < DSSSP_RESCUE_AUTO || nt.rescue_levels < 1 || nt.rescue_levels > 1024 ||
        (nt.phase_timing != 0 && nt.phase_timing != 1) ||
        (nt.block != 128 && nt.block != 256 && nt.block != 512) ||
        (nt.edges_per_thread != 1 && nt.edges_per_thread != 2 && nt.edges_per_thread != 4))
        return fail(DSSSP_ERR_ARG, "tuning out of range");
    g_tuning = nt;
    return DSSSP_OK;
}

static void graph_release(dsssp_graph *g);

static void graph_attach(dsssp_graph *g)
{
    std::lock_guard<std::mutex> lk(g->mu);
    ++g->attached;
}

static void graph_detach(dsssp_graph *g)
{
    bool release;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        release = --g->attached == 0 && g->destroy_pending;
    }
    if (release) graph_release(g);
}

static void graph_release(dsssp_graph *g)
{
    if (!g) return;
    DevGuard dg(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    dev_free(g->vr);
    dev_free(g->adj);
    dev_free(g->pool);
    dev_free(g->tbm);
    dev_free(g->emap);
    if (g->upd) {
        dev_free(g->upd->rec);
        dev_free(g->upd->tslot);
        delete g->upd;
    }
    if (g->stage) (void)hipHostFree(g->stage);
    dev_free(g->wsum);
    dev_free(g->bsum);
    dev_free(g->bbase);
    dev_free(g->gc);
    dev_free(g->mkeys);
    dev_free(g->mreq);
    dev_free(g->mfirst);
    if (g->gc_host) (void)hipHostFree(g->gc_host);
    for (auto &e : g->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : g->kev) (void)hipEventDestroy(e);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
}

static int graph_alloc(dsssp_graph *g, uint64_t pool_cap)
{
    DevGuard dg(g->device);
    HIP_TRY(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
    for (auto &e : g->ev) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    const size_t n = (size_t)g->n;
    g->nw = words_of(n);
    if (int s = dev_alloc((void **)&g->vr, sizeof(VRec) * n, "vertex records")) return s;
    if (int s = dev_alloc((void **)&g->adj, sizeof(uint64_t) * g->E, "adjacency")) return s;
    if (int s = dev_alloc((void **)&g->tbm, sizeof(uint64_t) * g->nw, "touched bitmap")) return s;
    if (int s = dev_alloc((void **)&g->wsum, sizeof(uint32_t) * g->nw, "word sizes")) return s;
    const size_t nb = (g->nw + 255) / 256;
    if (int s = dev_alloc((void **)&g->bsum, sizeof(uint64_t) * nb, "group sizes")) return s;
    if (nb > INS_FOLD_GROUPS)
        if (int s = dev_alloc((void **)&g->bbase, sizeof(uint64_t) * nb, "group bases")) return s;
    if (int s = dev_alloc((void **)&g->gc, sizeof(GraphCtl), "graph control")) return s;
    g->pool_cap = pool_cap;
    if (int s = dev_alloc((void **)&g->pool, sizeof(uint64_t) * pool_cap, "delta pool")) return s;
    HIP_TRY(hipHostMalloc((void **)&g->gc_host, sizeof(GraphCtl), hipHostMallocDefault));
    memset(g->gc_host, 0, sizeof(GraphCtl));
    HIP_TRY(hipMemset(g->tbm, 0, sizeof(uint64_t) * g->nw));
    HIP_TRY(hipMemset(g->gc, 0, sizeof(GraphCtl)));
    return DSSSP_OK;
}

extern "C" int dsssp_graph_create(int64_t n, int64_t m, const int32_t *a, const int32_t *b, const int32_t *w,
                                  uint32_t flags, dsssp_graph **out)
{
    clear_error();
    if (!out) return fail(DSSSP_ERR_ARG, "null output");
    *out = nullptr;
    if (flags != 0) return fail(DSSSP_ERR_ARG, "unknown flags");
    if (n < 1) return fail(DSSSP_ERR_ARG, "n must be positive");
    if (int s = check_device(t_device)) return s;
    std::vector<uint64_t> rp, adj;
    if (int s = build_csr(n, m, a, b, w, rp, adj)) return s;
    if ((uint64_t)n * INIT_DCAP * 3 + adj.size() / 2 > 0xFFFFFFFFull / 2)
        return fail(DSSSP_ERR_ARG, "graph too large for the delta pool (n=%lld, entries=%zu)", (long long)n,
                    adj.size());
    dsssp_graph *g = new dsssp_graph();
    g->device = t_device;
    g->n = n;
    g->E = adj.size();
    {
        double sum = 0;
        for (int64_t i = 0; i < m; ++i) {
            sum += w[i];
            if (w[i] == 0) g->zero_w = true;
        }
        g->auto_delta = (uint32_t)std::min(std::max(2.0 * sum / (double)std::max<int64_t>(m, 1), 1.0), 1e9);
    }
    // Every row starts with a delta block of INIT_DCAP entries, so most
    // insertions land in place; the pool leaves room for the growth bound of
    // ensure_pool.
    const uint64_t pool_cap = std::min<uint64_t>(g->E / 2 + 3 * INIT_DCAP * (uint64_t)n + (1u << 16),
                                                 0xFFFFFFFFull / 2);
    if (int s = graph_alloc(g, pool_cap)) {
        graph_release(g);
        return s;
    }
    DevGuard dg(g->device);
    std::vector<VRec> vr((size_t)n);
    parallel_ranges(n, [&](int64_t lo, int64_t hi, int) {
        for (int64_t v = lo; v < hi; ++v) {
            const uint64_t *row = adj.data() + rp[v], *end = adj.data() + rp[v + 1];
            const uint32_t blo = (uint32_t)(std::partition_point(row, end, [v](uint64_t e) {
                                                return e_col(e) < (uint32_t)v;
                                            }) - row);
            vr[v] = VRec{rp[v], (uint32_t)(rp[v + 1] - rp[v]), INIT_DCAP, ((uint64_t)v * INIT_DCAP) << 32, blo, 0u};
        }
    });
    if (hipMemcpy(g->vr, vr.data(), sizeof(VRec) * (size_t)n, hipMemcpyHostToDevice) != hipSuccess ||
        (g->E && hipMemcpy(g->adj, adj.data(), sizeof(uint64_t) * g->E, hipMemcpyHostToDevice) != hipSuccess)) {
        const int s = fail(DSSSP_ERR_HIP, "graph upload failed: %s", hipGetErrorString(hipGetLastError()));
        graph_release(g);
        return s;
    }
    g->pool_top = INIT_DCAP * (uint64_t)n;
    g->base_entries = g->E;
    g->gc_host->pool_top = g->pool_top;
    int s = hipMemcpy(&g->gc->pool_top, &g->pool_top, sizeof(uint64_t), hipMemcpyHostToDevice) == hipSuccess
                ? DSSSP_OK
                : fail(DSSSP_ERR_HIP, "graph upload failed");
    if (!s) s = build_emap(g, g->E, g->E / 8);
    if (!s && hipStreamSynchronize(g->stream) != hipSuccess)
        s = fail(DSSSP_ERR_HIP, "edge map build failed: %s", hipGetErrorString(hipGetLastError()));
    if (s) {
        const std::string keep = last_error();
        graph_release(g);
        return fail(s, "%s", keep.c_str());
    }
    *out = g;
    return DSSSP_OK;
}

extern "C" int dsssp_graph_destroy(dsssp_graph *g)
{
    clear_error();
    if (!g) return DSSSP_OK;
    bool release;
    {
        std::lock_guard<std::mutex> lk(g->mu);
        g->destroy_pending = true;
        release = g->attached == 0;
    }
    if (release) graph_release(g);
    return DSSSP_OK;
}

extern "C" int dsssp_graph_info(const dsssp_graph *g, int64_t *n, int64_t *m_live, uint64_t *device_bytes)
{
    clear_error();
    if (!g) return fail(DSSSP_ERR_ARG, "null graph");
    if (n) *n = g->n;
    if (m_live) {
        int64_t m = 0;
        int s = dsssp_graph_export(const_cast<dsssp_graph *>(g), &m, nullptr, nullptr, nullptr);
        if (s) return s;
        *m_live = m;
    }
    if (device_bytes)
        *device_bytes = 32 * (uint64_t)g->n + 8 * g->E + 12 * (uint64_t)g->nw +
                        16 * (uint64_t)((g->nw + 255) / 256) + 8 * g->pool_cap + 8 * g->emap_slots;
    return DSSSP_OK;
}

extern "C" int dsssp_graph_stats(const dsssp_graph *g, dsssp_graph_upkeep *out)
{
    clear_error();
    if (!g || !out) return fail(DSSSP_ERR_ARG, "null argument");
    out->n = g->n;
    out->base_entries = (int64_t)g->base_entries;
    out->delta_entries = (int64_t)g->delta_entries;
    out->dead_entries = (int64_t)g->dead_entries;
    out->pool_top = g->pool_top;
    out->pool_cap = g->pool_cap;
    out->emap_slots = g->emap_slots;
    out->repacks = g->repacks;
    return DSSSP_OK;
}

extern "C" int dsssp_graph_repack(dsssp_graph *g)
{
    clear_error();
    if (!g) return fail(DSSSP_ERR_ARG, "null argument");
    DevGuard dg(g->device);
    HIP_TRY(hipStreamSynchronize(g->stream));
    return repack(g, 0);
}

extern "C" int dsssp_graph_export(dsssp_graph *g, int64_t *m, int32_t *a, int32_t *b, int32_t *w)
{
    clear_error();
    if (!g || !m) return fail(DSSSP_ERR_ARG, "null argument");
    DevGuard dg(g->device);
    const size_t n = (size_t)g->n;
    std::vector<VRec> vr(n);
    std::vector<uint64_t> adj(g->E), pool(g->pool_top);
    HIP_TRY(hipStreamSynchronize(g->stream));
    HIP_TRY(hipMemcpy(vr.data(), g->vr, sizeof(VRec) * n, hipMemcpyDeviceToHost));
    if (g->E) HIP_TRY(hipMemcpy(adj.data(), g->adj, sizeof(uint64_t) * g->E, hipMemcpyDeviceToHost));
    if (g->pool_top) HIP_TRY(hipMemcpy(pool.data(), g->pool, sizeof(uint64_t) * g->pool_top, hipMemcpyDeviceToHost));
    // Each undirected edge is stored once per endpoint (a self-loop twice in its
    // row): keep the copy in the row of the smaller endpoint.
    int64_t cnt = 0;
    auto take = [&](uint32_t v, uint64_t e, uint64_t &selfs) {
        if (e & E_DEAD) return;
        const uint32_t c = e_col(e);
        if (c < v) return;
        if (c == v && (selfs++ & 1)) return;
        if (a) {
            a[cnt] = (int32_t)v;
            b[cnt] = (int32_t)c;
            w[cnt] = (int32_t)e_w(e);
        }
        ++cnt;
    };
    for (size_t v = 0; v < n; ++v) {
        uint64_t selfs = 0;
        for (uint64_t j = vr[v].boff; j < vr[v].boff + vr[v].blen; ++j) take((uint32_t)v, adj[j], selfs);
        const uint64_t beg = vr[v].dbl >> 32, len = (uint32_t)vr[v].dbl;
        for (uint64_t j = 0; j < len; ++j) take((uint32_t)v, pool[beg + j], selfs);
    }
    *m = cnt;
    return DSSSP_OK;
}

extern "C" int dsssp_graph_clone(dsssp_graph *src, dsssp_graph **out)
{
    clear_error();
    if (!src || !out) return fail(DSSSP_ERR_ARG, "null argument");
    *out = nullptr;
    dsssp_graph *g = new dsssp_graph();
    g->device = src->device;
    g->n = src->n;
    g->E = src->E;
    if (int s = graph_alloc(g, src->pool_cap)) {
        graph_release(g);
        return s;
    }
    if (int s = dsssp_graph_copy(g, src)) {
        graph_release(g);
        return s;
    }
    *out = g;
    return DSSSP_OK;
}

extern "C" int dsssp_graph_copy(dsssp_graph *dst, dsssp_graph *src)
{
    clear_error();
    if (!dst || !src) return fail(DSSSP_ERR_ARG, "null argument");
    if (dst->n != src->n || dst->E != src->E) return fail(DSSSP_ERR_ARG, "graphs differ in shape");
    DevGuard dg(dst->device);
    HIP_TRY(hipStreamSynchronize(src->stream));
    if (dst->pool_cap < src->pool_top) {
        dev_free(dst->pool);
        if (int s = dev_alloc((void **)&dst->pool, sizeof(uint64_t) * src->pool_cap, "delta pool")) return s;
        dst->pool_cap = src->pool_cap;
    }
    const size_t n = (size_t)src->n;
    hipStream_t s = dst->stream;
    HIP_TRY(hipMemcpyAsync(dst->vr, src->vr, sizeof(VRec) * n, hipMemcpyDeviceToDevice, s));
    if (src->E) HIP_TRY(hipMemcpyAsync(dst->adj, src->adj, sizeof(uint64_t) * src->E, hipMemcpyDeviceToDevice, s));
    if (src->pool_top)
        HIP_TRY(hipMemcpyAsync(dst->pool, src->pool, sizeof(uint64_t) * src->pool_top, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemcpyAsync(dst->gc, src->gc, sizeof(GraphCtl), hipMemcpyDeviceToDevice, s));
    if (dst->emap_slots != src->emap_slots) {
        dev_free(dst->emap);
        dst->emap_slots = 0;
        if (int e = dev_alloc((void **)&dst->emap, 8 * src->emap_slots, "edge map")) return e;
        dst->emap_slots = src->emap_slots;
    }
    HIP_TRY(hipMemcpyAsync(dst->emap, src->emap, 8 * src->emap_slots, hipMemcpyDeviceToDevice, s));
    HIP_TRY(hipMemsetAsync(dst->tbm, 0, sizeof(uint64_t) * dst->nw, s));
    HIP_TRY(hipStreamSynchronize(s));
    dst->pool_top = src->pool_top;
    dst->delta_entries = src->delta_entries;
    dst->base_entries = src->base_entries;
    dst->dead_entries = src->dead_entries;
    dst->repacks = src->repacks;
    dst->emap_stale = src->emap_stale;
    dst->emap_used = src->emap_used;
    dst->zero_w = src->zero_w;
    dst->auto_delta = src->auto_delta;
    return DSSSP_OK;
}

static void tree_release(dsssp_tree *t)
{
    if (!t) return;
    dsssp_graph *g = t->g;
    {
        DevGuard dg(g->device);
        (void)hipStreamSynchronize(g->stream);
        dev_free(t->st);
        dev_free(t->fb[0]);
        dev_free(t->fb[1]);
        dev_free(t->ib);
        dev_free(t->hq[0]);
        dev_free(t->hq[1]);
        dev_free(t->lvl);
        dev_free(t->best);
        dev_free(t->cb);
        dev_free(t->bstats);
        dev_free(t->c);
        if (t->c_host) (void)hipHostFree(t->c_host);
        for (auto &e : t->graphs) (void)hipGraphExecDestroy(e.exec);
        delete t;
    }
    graph_detach(g);
}

static int tree_alloc(dsssp_tree *t)
{
    dsssp_graph *g = t->g;
    DevGuard dg(g->device);
    const size_t ns = (size_t)t->nslots;
    t->nw = words_of(ns);
    if (int s = dev_alloc((void **)&t->st, sizeof(uint64_t) * ns, "tree state")) return s;
    if (int s = dev_alloc((void **)&t->fb[0], sizeof(uint64_t) * t->nw, "frontier bitmap")) return s;
    if (int s = dev_alloc((void **)&t->fb[1], sizeof(uint64_t) * t->nw, "frontier bitmap")) return s;
    if (int s = dev_alloc((void **)&t->ib, sizeof(uint64_t) * t->nw, "invalidated bitmap")) return s;
    if (int s = dev_alloc((void **)&t->bstats, sizeof(unsigned long long) * STAT_SLOTS * NSTAT, "statistics"))
        return s;
    if (int s = dev_alloc((void **)&t->c, sizeof(Counters), "counters")) return s;
    HIP_TRY(hipHostMalloc((void **)&t->c_host, sizeof(Counters), hipHostMallocDefault));
    memset(t->c_host, 0, sizeof(Counters));
    HIP_TRY(hipMemsetAsync(t->fb[0], 0, sizeof(uint64_t) * t->nw, g->stream));
    HIP_TRY(hipMemsetAsync(t->fb[1], 0, sizeof(uint64_t) * t->nw, g->stream));
    HIP_TRY(hipMemsetAsync(t->ib, 0, sizeof(uint64_t) * t->nw, g->stream));
    HIP_TRY(hipMemsetAsync(t->bstats, 0, sizeof(unsigned long long) * STAT_SLOTS * NSTAT, g->stream));
    HIP_TRY(hipMemsetAsync(t->c, 0, sizeof(Counters), g->stream));
    if (int s = ensure_hq(t, g->E + g->delta_entries)) return s;
    return DSSSP_OK;
}

// Creates a tree object of `ntrees` trees (slot stride 2^kshift) and fills it:
// from host arrays (one tree), or from scratch on the GPU from `roots`
// (per_tree: root j belongs to tree j; otherwise all roots share tree 0).
static int tree_make(dsssp_graph *g, int32_t ntrees, const std::vector<uint32_t> &roots, bool per_tree,
                     const int32_t *parent, const int64_t *dist, uint32_t flags, dsssp_tree **out)
{
    *out = nullptr;
    if (flags & ~DSSSP_TREE_CANONICAL) return fail(DSSSP_ERR_ARG, "unknown flags");
    const int64_t n = g->n;
    uint32_t kshift = 0;
    while ((1 << kshift) < ntrees) ++kshift;
    // Slot ids, bitmap bit indices and chunk-list slot fields are 32-bit.
    if (((uint64_t)n << kshift) > 0xFFFFFF00ull)
        return fail(DSSSP_ERR_ARG, "n * trees must stay below 2^32 - 256 (n=%lld, trees=%d)", (long long)n, ntrees);
    std::vector<uint64_t> packed;
    uint64_t max_dist = 0;
    if (parent) {
        if (!dist) return fail(DSSSP_ERR_ARG, "parent given without dist");
        packed.resize((size_t)n);
        for (int64_t v = 0; v < n; ++v) {
            const int32_t p = parent[v];
            if (p == -1 || dist[v] == DSSSP_INF) {
                packed[v] = ST_INF;
                continue;
            }
            if (p < 0 || p >= n) return fail(DSSSP_ERR_RANGE, "vertex %lld: parent %d outside [0,n)", (long long)v, p);
            if (dist[v] < 0 || dist[v] >= (int64_t)DIST_INF)
                return fail(DSSSP_ERR_OVERFLOW, "vertex %lld: distance %lld outside [0, 2^32-1)", (long long)v,
                            (long long)dist[v]);
            if (p == v && dist[v] != 0)
                return fail(DSSSP_ERR_TREE, "root %lld has distance %lld, not 0", (long long)v, (long long)dist[v]);
            packed[v] = pack_state((uint32_t)dist[v], (uint32_t)p);
            max_dist = std::max<uint64_t>(max_dist, (uint64_t)dist[v]);
        }
    } else {
        for (uint32_t r : roots)
            if ((int64_t)r >= n) return fail(DSSSP_ERR_RANGE, "source %u outside [0,%lld)", r, (long long)n);
    }
    dsssp_tree *t = new dsssp_tree();
    t->g = g;
    graph_attach(g);
    t->src = roots.empty() ? -1 : (int32_t)roots[0];
    t->ntrees = ntrees;
    t->kshift = kshift;
    t->nslots = (uint64_t)n << kshift;
    // A given tree is deep when its distances span DEEP_STEPS automatic steps; a
    // tree built from scratch is built ordered and is deep when that took more
    // than DEEP_ROUNDS rounds.
    t->deep = parent ? max_dist > (uint64_t)DEEP_STEPS * g->auto_delta : true;
    if (int s = tree_alloc(t)) {
        tree_release(t);
        return s;
    }
    DevGuard dg(g->device);
    int64_t rounds = 0;
    int s = DSSSP_OK;
    uint32_t *d_roots = nullptr;
    launch_fill_u32((uint32_t *)t->c, 0u, sizeof(Counters) / 4, g->stream);
    if (parent) {
        if (hipMemcpyAsync(t->st, packed.data(), sizeof(uint64_t) * (size_t)n, hipMemcpyHostToDevice, g->stream) !=
            hipSuccess)
            s = fail(DSSSP_ERR_HIP, "tree upload failed");
        if (!s && (flags & DSSSP_TREE_CANONICAL)) s = repick_parents(t, g->zero_w, rounds);
    } else {
        // From scratch: every slot unreachable but the roots, then relax.
        launch_fill_u64(t->st, ST_INF, t->nslots, g->stream);
        s = dev_alloc((void **)&d_roots, sizeof(uint32_t) * std::max<size_t>(roots.size(), 1), "roots");
        if (!s && !roots.empty() &&
            hipMemcpyAsync(d_roots, roots.data(), sizeof(uint32_t) * roots.size(), hipMemcpyHostToDevice,
                           g->stream) != hipSuccess)
            s = fail(DSSSP_ERR_HIP, "root upload failed");
        if (!s) {
            zero_ring(t);
            launch_set_sources(t->st, d_roots, (uint32_t)roots.size(), t->kshift, per_tree, t->fb[0], &t->c->act[0],
                               g->stream);
            launch_fill_u32(&t->c->fenc[0][0], ~0u, 1, g->stream);  // the roots were marked at distance 0
            int passes = 0;
            s = scratch_passes(t, passes);
            if (!s) s = run_rounds(t, M_RELAX, rounds, nullptr);
            t->deep = rounds > DEEP_ROUNDS;
            // A build's depth says nothing about an update's: updates start
            // from round_batch and learn their own.
            t->last_rounds[M_RELAX] = 0;
        }
        if (!s && g->zero_w) s = repick_parents(t, true, rounds);
    }
    launch_reduce_stats(t->c, t->bstats, g->stream);
    if (hipMemcpyAsync(t->c_host, t->c, sizeof(Counters), hipMemcpyDeviceToHost, g->stream) != hipSuccess ||
        hipStreamSynchronize(g->stream) != hipSuccess) {
        if (!s) s = fail(DSSSP_ERR_HIP, "tree setup failed: %s", hipGetErrorString(hipGetLastError()));
    } else if (!s) {
        s = confirm_overflow(t);
        if (!s) s = finish_errors(t, 0);
    }
    dev_free(d_roots);
    if (s) {
        const std::string keep = last_error();
        tree_release(t);
        return fail(s, "%s", keep.c_str());
    }
    *out = t;
    return DSSSP_OK;
}

extern "C" int dsssp_tree_create(dsssp_graph *g, int32_t src, const int32_t *parent, const int64_t *dist,
                                 uint32_t flags, dsssp_tree **out)
{
    clear_error();
    if (!g || !out) return fail(DSSSP_ERR_ARG, "null argument");
    if (!parent && (src < 0 || src >= g->n))
        return fail(DSSSP_ERR_RANGE, "source %d outside [0,%lld)", src, (long long)g->n);
    std::vector<uint32_t> roots;
    if (!parent) roots.push_back((uint32_t)src);
    return tree_make(g, 1, roots, false, parent, dist, flags, out);
}

extern "C" int dsssp_tree_create_roots(dsssp_graph *g, int32_t nroots, const int32_t *roots, uint32_t flags,
                                       dsssp_tree **out)
{
    clear_error();
    if (!g || !out || nroots < 1 || !roots) return fail(DSSSP_ERR_ARG, "bad argument");
    std::vector<uint32_t> r((size_t)nroots);
    for (int32_t i = 0; i < nroots; ++i) {
        if (roots[i] < 0 || roots[i] >= g->n) return fail(DSSSP_ERR_RANGE, "root %d outside [0,n)", roots[i]);
        r[i] = (uint32_t)roots[i];
    }
    return tree_make(g, 1, r, false, nullptr, nullptr, flags, out);
}

extern "C" int dsssp_forest_create(dsssp_graph *g, int32_t ntrees, const int32_t *srcs, uint32_t flags,
                                   dsssp_tree **out)
{
    clear_error();
    if (!g || !out || ntrees < 1 || ntrees > 1024 || !srcs) return fail(DSSSP_ERR_ARG, "bad argument");
    std::vector<uint32_t> r((size_t)ntrees);
    for (int32_t i = 0; i < ntrees; ++i) {
        if (srcs[i] < 0 || srcs[i] >= g->n) return fail(DSSSP_ERR_RANGE, "source %d outside [0,n)", srcs[i]);
        r[i] = (uint32_t)srcs[i];
    }
    return tree_make(g, ntrees, r, true, nullptr, nullptr, flags, out);
}

extern "C" int dsssp_tree_count(const dsssp_tree *t, int32_t *ntrees)
{
    clear_error();
    if (!t || !ntrees) return fail(DSSSP_ERR_ARG, "null argument");
    *ntrees = t->ntrees;
    return DSSSP_OK;
}

extern "C" int dsssp_tree_destroy(dsssp_tree *t)
{
    clear_error();
    tree_release(t);
    return DSSSP_OK;
}

extern "C" int dsssp_forest_read(dsssp_tree *t, int32_t tree, int32_t *parent, int64_t *dist)
{
    clear_error();
    if (!t || !parent || !dist) return fail(DSSSP_ERR_ARG, "null argument");
    if (tree < 0 || tree >= t->ntrees) return fail(DSSSP_ERR_ARG, "tree %d outside [0,%d)", tree, t->ntrees);
    DevGuard dg(t->g->device);
    const size_t n = (size_t)t->g->n;
    const size_t stride = (size_t)1 << t->kshift;
    HIP_TRY(hipStreamSynchronize(t->g->stream));
    std::vector<uint64_t> st(n);
    if (stride == 1) {
        HIP_TRY(hipMemcpy(st.data(), t->st, sizeof(uint64_t) * n, hipMemcpyDeviceToHost));
    } else {
        // the tree's states gathered into one contiguous buffer on the device
        uint64_t *col = nullptr;
        if (int s = dev_alloc((void **)&col, sizeof(uint64_t) * n, "forest read")) return s;
        launch_gather_tree(t->st, n, t->kshift, (uint32_t)tree, col, t->g->stream);
        hipError_t e = hipMemcpyAsync(st.data(), col, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, t->g->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(t->g->stream);
        dev_free(col);
        if (e != hipSuccess) return fail(DSSSP_ERR_HIP, "forest read failed: %s", hipGetErrorString(e));
    }
    for (size_t v = 0; v < n; ++v) {
        if (st[v] == ST_INF) {
            parent[v] = -1;
            dist[v] = DSSSP_INF;
        } else {
            parent[v] = (int32_t)st_parent(st[v]);
            dist[v] = (int64_t)st_dist(st[v]);
        }
    }
    return DSSSP_OK;
}

extern "C" int dsssp_tree_read(dsssp_tree *t, int32_t *parent, int64_t *dist)
{
    return dsssp_forest_read(t, 0, parent, dist);
}

extern "C" int dsssp_tree_copy(dsssp_tree *dst, dsssp_tree *src)
{
    clear_error();
    if (!dst || !src) return fail(DSSSP_ERR_ARG, "null argument");
    if (dst->nslots != src->nslots || dst->kshift != src->kshift) return fail(DSSSP_ERR_ARG, "trees differ in shape");
    DevGuard dg(dst->g->device);
    HIP_TRY(hipStreamSynchronize(src->g->stream));
    const size_t ns = (size_t)src->nslots;
    HIP_TRY(hipMemcpyAsync(dst->st, src->st, sizeof(uint64_t) * ns, hipMemcpyDeviceToDevice, dst->g->stream));
    HIP_TRY(hipStreamSynchronize(dst->g->stream));
    dst->src = src->src;
    dst->deep = src->deep;
    return DSSSP_OK;
}

extern "C" int dsssp_tree_relax_full(dsssp_tree *t, dsssp_stats *stats)
{
    clear_error();
    if (!t) return fail(DSSSP_ERR_ARG, "null tree");
    dsssp_graph *g = t->g;
    DevGuard dg(g->device);
    hipStream_t s = g->stream;
    if (int e = ensure_hq(t, g->E + g->delta_entries)) return e;
    int64_t rounds = 0, repair = 0;
    double kernel_ms = 0;
    // Not a measure of update depth: kept, and the first group is round_batch.
    const int64_t keep_rounds = t->last_rounds[M_RELAX];
    t->last_rounds[M_RELAX] = 0;
    HIP_TRY(hipEventRecord(g->ev[0], s));
    launch_fill_u32((uint32_t *)t->c, 0u, sizeof(Counters) / 4, s);
    launch_fill_u64(t->fb[0], 0ull, t->nw, s);
    launch_fill_u64(t->fb[1], 0ull, t->nw, s);
    // The pass's own statistics (front, entries, marks, gathers) and time.
    uint64_t sw[4] = {0, 0, 0, 0};
    double sweep_ms = 0;
    int code = DSSSP_OK;
    if (t->kshift < 6) {
        // Each live edge once, both directions, from its higher-id endpoint's
        // row; the slots it lowers are marked in bitmap 0 for round 0.
        int grid = 0;
        const FArgs a = sweep_args(t, grid);
        HIP_TRY(hipEventRecord(g->ev[3], s));
        single_pass(t, M_SWEEP, a, grid);
        HIP_TRY(hipEventRecord(g->ev[4], s));
        launch_reduce_stats(t->c, t->bstats, s);
        HIP_TRY(hipMemcpyAsync(t->c_host, t->c, sizeof(Counters), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        sweep_ms = elapsed(g->ev[3], g->ev[4]);
        const Counters &c = *t->c_host;
        sw[0] = c.stats[ST_RELAX_FRONT];
        sw[1] = c.stats[ST_RELAX_EDGES];
        sw[2] = c.stats[ST_RELAX_MARKS];
        sw[3] = c.stats[ST_SWEEP_GATHERS];
        // Distance-ordered rounds start from a threshold at distance 0 (an
        // empty ring would mean no threshold at all).
        if (relax_delta(t)) launch_fill_u32(&t->c->fenc[0][0], ~0u, 1, s);
        code = run_rounds(t, M_RELAX, rounds, &kernel_ms);
    } else {
        if (relax_delta(t)) launch_fill_u32(&t->c->fenc[0][0], ~0u, 1, s);
        code = run_rounds(t, M_RELAX, rounds, &kernel_ms, true);  // round 0 pushes from every slot
    }
    t->last_rounds[M_RELAX] = keep_rounds;
    HIP_TRY(hipEventRecord(g->ev[1], s));
    if (!code && g->zero_w) code = repick_parents(t, true, repair);
    HIP_TRY(hipEventRecord(g->ev[2], s));
    launch_reduce_stats(t->c, t->bstats, s);
    HIP_TRY(hipMemcpyAsync(t->c_host, t->c, sizeof(Counters), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipGetLastError());
    if (!code) code = confirm_overflow(t);
    if (!code) code = finish_errors(t, 0);
    if (code) {
        const std::string keep = last_error();
        scrub(t);
        return fail(code, "%s", keep.c_str());
    }
    if (stats) {
        const Counters &c = *t->c_host;
        dsssp_stats S{};
        S.relax_rounds = rounds;
        S.relax_launches = rounds;
        S.frontier_vertices = (int64_t)c.stats[ST_RELAX_FRONT];
        S.relax_edges = (int64_t)c.stats[ST_RELAX_EDGES];
        S.relax_improved = (int64_t)c.stats[ST_RELAX_MARKS];
        S.relax_enqueued = (int64_t)c.stats[ST_RELAX_MARKS];
        S.other_edges = (int64_t)c.stats[ST_OTHER_EDGES];
        S.relax_sweeps = (int64_t)c.stats[ST_SWEEPS] + 1;
        S.t_relax_ms = elapsed(g->ev[0], g->ev[2]);
        S.t_update_ms = S.t_relax_ms;
        // The pass: per slot 32 B (row fields + own state), per scanned entry
        // 8 B, per entry that read its neighbour's state 8 B more, per
        // improvement 16 B (state atomicMin + bitmap word).
        S.t_sweep_ms = sweep_ms;
        S.alg_bytes_sweep = 32.0 * (double)sw[0] + 8.0 * (double)sw[1] + 8.0 * (double)sw[3] + 16.0 * (double)sw[2];
        S.t_relax_kernel_ms = kernel_ms + (tuning().kernel_timing ? sweep_ms : 0.0);
        // The rounds: the same accounting as an update's relaxation (update_impl).
        const double relax = 32.0 * (double)(c.stats[ST_RELAX_FRONT] - sw[0]) +
                             16.0 * (double)(c.stats[ST_RELAX_EDGES] - sw[1]) +
                             16.0 * (double)(c.stats[ST_RELAX_MARKS] - sw[2]) + 8.0 * t->nw * (double)rounds;
        S.alg_bytes_relax = relax + S.alg_bytes_sweep;
        S.alg_bytes = S.alg_bytes_relax + 16.0 * (double)c.stats[ST_OTHER_EDGES];
        *stats = S;
    }
    return DSSSP_OK;
}

extern "C" int dsssp_probe(int32_t kind, uint64_t bytes, uint64_t count, int32_t reps, double *ms)
{
    clear_error();
    if ((kind != 0 && kind != 1) || bytes < 8 || reps < 1 || !ms) return fail(DSSSP_ERR_ARG, "bad argument");
    if (int s = check_device(t_device)) return s;
    DevGuard dg(t_device);
    const uint64_t n = bytes / 8;
    const uint64_t out_words = kind == 0 ? n : 8192ull * 256;
    uint64_t *in = nullptr, *out = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int st = dev_alloc((void **)&in, 8 * n, "probe input");
    if (!st) st = dev_alloc((void **)&out, 8 * out_words, "probe output");
    if (!st && (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                hipEventCreate(&e1) != hipSuccess))
        st = fail(DSSSP_ERR_HIP, "probe setup failed");
    if (!st) {
        launch_fill_u64(in, 1ull, n, s);
        launch_probe(kind, in, n, count, out, s);  // warm-up
        (void)hipEventRecord(e0, s);
        for (int32_t r = 0; r < reps; ++r) launch_probe(kind, in, n, count, out, s);
        (void)hipEventRecord(e1, s);
        if (hipStreamSynchronize(s) != hipSuccess)
            st = fail(DSSSP_ERR_HIP, "probe failed: %s", hipGetErrorString(hipGetLastError()));
        else
            *ms = elapsed(e0, e1) / reps;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    dev_free(in);
    dev_free(out);
    return st;
}

extern "C" int dsssp_batch_upload(dsssp_graph *g, const dsssp_change *chg, int64_t k, dsssp_batch **out)
{
    clear_error();
    if (!g || !out || k < 0 || (k > 0 && !chg)) return fail(DSSSP_ERR_ARG, "bad argument");
    *out = nullptr;
    int64_t ni, nd;
    bool zw;
    if (int s = validate_changes(g->n, chg, k, ni, nd, zw)) return s;
    DevGuard dg(g->device);
    dsssp_batch *b = new dsssp_batch();
    b->g = g;
    graph_attach(g);
    b->k = k;
    b->n_ins = ni;
    b->n_del = nd;
    b->zero_w = zw;
    int s = dev_alloc((void **)&b->rec, sizeof(int4) * (size_t)std::max<int64_t>(k, 1), "batch");
    if (!s) s = dev_alloc((void **)&b->tslot, sizeof(uint32_t) * 2 * (size_t)std::max<int64_t>(k, 1), "batch slots");
    if (!s && k > 0 &&
        hipMemcpyAsync(b->rec, chg, sizeof(int4) * (size_t)k, hipMemcpyHostToDevice, g->stream) != hipSuccess)
        s = fail(DSSSP_ERR_HIP, "batch upload failed");
    if (!s && hipStreamSynchronize(g->stream) != hipSuccess) s = fail(DSSSP_ERR_HIP, "batch upload failed");
    if (s) {
        const std::string keep = last_error();
        dsssp_batch_destroy(b);
        return fail(s, "%s", keep.c_str());
    }
    *out = b;
    return DSSSP_OK;
}

extern "C" int dsssp_batch_destroy(dsssp_batch *b)
{
    if (!b) return DSSSP_OK;
    dsssp_graph *g = b->g;
    {
        DevGuard dg(g->device);
        (void)hipStreamSynchronize(g->stream);
        dev_free(b->rec);
        dev_free(b->tslot);
        delete b;
    }
    graph_detach(g);
    return DSSSP_OK;
}

extern "C" int dsssp_update_batch(dsssp_graph *g, dsssp_tree *t, dsssp_batch *b, dsssp_stats *stats)
{
    clear_error();
    if (!g || !t || !b) return fail(DSSSP_ERR_ARG, "null argument");
    if (t->g != g || b->g != g) return fail(DSSSP_ERR_ARG, "tree or batch belongs to another graph");
    if (b->k == 0) {
        if (stats) *stats = dsssp_stats{};
        return DSSSP_OK;
    }
    return update_impl(g, t, b, stats, false);
}

// The graph's reusable device batch, with room for k records.
static int ensure_batch(dsssp_graph *g, int64_t k)
{
    if (!g->upd) {
        g->upd = new dsssp_batch();
        g->upd->g = g;  // owned by the graph, not attached to it
    }
    dsssp_batch *b = g->upd;
    if (b->cap < k) {
        const int64_t cap = std::max<int64_t>(k, 2 * b->cap);
        HIP_TRY(hipStreamSynchronize(g->stream));
        dev_free(b->rec);
        dev_free(b->tslot);
        b->cap = 0;
        if (int s = dev_alloc((void **)&b->rec, sizeof(int4) * (size_t)cap, "batch")) return s;
        if (int s = dev_alloc((void **)&b->tslot, sizeof(uint32_t) * 2 * (size_t)cap, "batch slots")) return s;
        b->cap = cap;
    }
    return DSSSP_OK;
}

// The graph's pinned staging buffer, with room for k records.
static int ensure_stage(dsssp_graph *g, int64_t k)
{
    if (g->stage_cap < k) {
        const int64_t cap = std::max<int64_t>(k, 2 * g->stage_cap);
        HIP_TRY(hipStreamSynchronize(g->stream));
        if (g->stage) (void)hipHostFree(g->stage);
        g->stage = nullptr;
        g->stage_cap = 0;
        HIP_TRY(hipHostMalloc((void **)&g->stage, sizeof(dsssp_change) * (size_t)cap, hipHostMallocDefault));
        g->stage_cap = cap;
    }
    return DSSSP_OK;
}

// Applies the k records whose upload to the device batch is enqueued between
// ev[8] and ev[9].
static int apply_uploaded(dsssp_graph *g, dsssp_tree *t, int64_t k, int64_t ni, int64_t nd, bool zw,
                          dsssp_stats *stats)
{
    dsssp_batch *b = g->upd;
    b->k = k;
    b->n_ins = ni;
    b->n_del = nd;
    b->zero_w = zw;
    return update_impl(g, t, b, stats, true);
}

extern "C" int dsssp_update(dsssp_graph *g, dsssp_tree *t, const dsssp_change *chg, int64_t k, dsssp_stats *stats)
{
    clear_error();
    if (!g || !t || k < 0 || (k > 0 && !chg)) return fail(DSSSP_ERR_ARG, "bad argument");
    if (t->g != g) return fail(DSSSP_ERR_ARG, "tree belongs to another graph");
    int64_t ni, nd;
    bool zw;
    if (int s = validate_changes(g->n, chg, k, ni, nd, zw)) return s;
    if (k == 0) {
        if (stats) *stats = dsssp_stats{};
        return DSSSP_OK;
    }
    DevGuard dg(g->device);
    if (int s = ensure_batch(g, k)) return s;
    if (int s = ensure_stage(g, k)) return s;
    // One host thread: a check and copy split over the host threads (3x faster
    // for 10^6 records), or fused into one pass, left later upkeep (repacks of
    // graphs updated that way) stalling in the device allocator for up to
    // 0.7 s now and then (MI355X, ROCm 7); the plain check and memcpy did not.
    memcpy(g->stage, chg, sizeof(dsssp_change) * (size_t)k);
    HIP_TRY(hipEventRecord(g->ev[8], g->stream));
    HIP_TRY(hipMemcpyAsync(g->upd->rec, g->stage, sizeof(int4) * (size_t)k, hipMemcpyHostToDevice, g->stream));
    HIP_TRY(hipEventRecord(g->ev[9], g->stream));
    return apply_uploaded(g, t, k, ni, nd, zw, stats);
}

// ---------------------------------------------------------------------------
// Allocation-free update.  Its only device memory beyond the caller's records
// and the graph's and tree's own buffers is `scratch`: per record copy, the
// position it took in its row's delta block, then four validation counters.
// Each device holds its own copy for the life of the process; every call on a
// device shares it and leaves the last call's contents, so calls that overlap
// on one device race.  Host code uses the per-device address obtained from
// hipGetSymbolAddress, never the array's name.
// ---------------------------------------------------------------------------

namespace {

constexpr int64_t NOALLOC_MAX_RECORDS = 1 << 21;
constexpr size_t NOALLOC_SLOT_BYTES = 2 * (size_t)NOALLOC_MAX_RECORDS * sizeof(uint32_t);
constexpr size_t NOALLOC_SCRATCH_BYTES = NOALLOC_SLOT_BYTES + 256;
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
void *g_scratch_at[NOALLOC_MAX_DEVICES] = {};

// The device address of `scratch` on device dev (looked up once per device,
// with that device current).
int scratch_of(int dev, unsigned char **out)
{
    *out = nullptr;
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(DSSSP_ERR_NODEV, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_at[dev]) {
        DevGuard dg(dev);
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(DSSSP_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", dev,
                        hipGetErrorString(e));
        g_scratch_at[dev] = p;
    }
    *out = (unsigned char *)g_scratch_at[dev];
    return DSSSP_OK;
}

}  // namespace

static int scratch_of_device(int dev)
{
    unsigned char *p = nullptr;
    return scratch_of(dev, &p);
}

extern "C" int dsssp_update_noalloc(dsssp_graph *g, dsssp_tree *t, const dsssp_change *d_chg, int64_t k,
                                    dsssp_stats *stats)
{
    clear_error();
    if (!g || !t || k < 0 || (k > 0 && !d_chg)) return fail(DSSSP_ERR_ARG, "bad argument");
    if (t->g != g) return fail(DSSSP_ERR_ARG, "tree belongs to another graph");
    if (k > NOALLOC_MAX_RECORDS)
        return fail(DSSSP_ERR_ARG, "allocation-free update: %lld records exceed its maximum of %lld", (long long)k,
                    (long long)NOALLOC_MAX_RECORDS);
    if (k == 0) {
        if (stats) *stats = dsssp_stats{};
        return DSSSP_OK;
    }
    unsigned char *sc = nullptr;
    if (int s = scratch_of(g->device, &sc)) return s;
    DevGuard dg(g->device);
    hipStream_t s = g->stream;
    unsigned long long *cnt = (unsigned long long *)(sc + NOALLOC_SLOT_BYTES);
    launch_validate((const int4 *)d_chg, k, g->n, cnt, s);
    unsigned long long h[4] = {0, 0, 0, 0};
    HIP_TRY(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (h[3] != ~0ull) {
        dsssp_change c{};
        HIP_TRY(hipMemcpyAsync(&c, d_chg + h[3], sizeof c, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        return fail(c.w < 0 ? DSSSP_ERR_WEIGHT : DSSSP_ERR_RANGE,
                    "change %llu (%d %d %d %d): vertex outside [0,%lld) or negative weight", h[3], c.a, c.b, c.w,
                    c.op, (long long)g->n);
    }
    dsssp_batch b;
    b.g = g;
    b.k = k;
    b.n_ins = (int64_t)h[0];
    b.n_del = (int64_t)h[1];
    b.zero_w = h[2] != 0;
    b.rec = (int4 *)d_chg;
    b.tslot = (uint32_t *)sc;
    b.cap = k;
    return update_impl(g, t, &b, stats, false, true);
}

namespace dsssp {

int pinned_alloc(void **p, size_t bytes)
{
    *p = nullptr;
    if (hipHostMalloc(p, bytes, hipHostMallocPortable) != hipSuccess) {
        *p = nullptr;
        return fail(DSSSP_ERR_NOMEM, "pinned host allocation of %zu bytes failed: %s", bytes,
                    hipGetErrorString(hipGetLastError()));
    }
    return DSSSP_OK;
}

void pinned_free(void *p)
{
    if (p) (void)hipHostFree(p);
}

int update_pinned(dsssp_graph *g, dsssp_tree *t, const dsssp_change *const *buf, const int64_t *len,
                  const int64_t *ni, const int64_t *nd, const uint8_t *zw, int32_t nbuf, bool check,
                  int32_t *applied, dsssp_stats *stats)
{
    clear_error();
    *applied = 0;
    if (!g || !t || nbuf < 1 || !buf || !len) return fail(DSSSP_ERR_ARG, "bad argument");
    if (t->g != g) return fail(DSSSP_ERR_ARG, "tree belongs to another graph");
    std::vector<uint64_t> first((size_t)nbuf + 1, 0);
    uint64_t dels = 0;
    for (int32_t i = 0; i < nbuf; ++i) {
        first[i + 1] = first[i] + (uint64_t)len[i];
        dels += (uint64_t)nd[i];
    }
    const int64_t k = (int64_t)first[nbuf];
    DevGuard dg(g->device);
    if (k > 0) {
        if (int s = ensure_batch(g, k)) return s;
    }
    dsssp_batch *b = g->upd;
    HIP_TRY(hipEventRecord(g->ev[8], g->stream));
    for (int32_t i = 0; i < nbuf; ++i)
        if (len[i])
            HIP_TRY(hipMemcpyAsync(b->rec + first[i], buf[i], sizeof(int4) * (size_t)len[i], hipMemcpyHostToDevice,
                                   g->stream));
    HIP_TRY(hipEventRecord(g->ev[9], g->stream));
    int32_t take = nbuf;
    if (check && nbuf > 1 && dels > 0) {
        uint64_t slots = 1024;
        while (slots < 2 * dels) slots *= 2;
        if (g->mslots < slots) {
            HIP_TRY(hipStreamSynchronize(g->stream));
            dev_free(g->mkeys);
            dev_free(g->mreq);
            g->mslots = 0;
            if (int s = dev_alloc((void **)&g->mkeys, 8 * slots, "merge check keys")) return s;
            if (int s = dev_alloc((void **)&g->mreq, 4 * slots, "merge check requests")) return s;
            g->mslots = slots;
        }
        if (g->mfirst_cap < nbuf) {
            HIP_TRY(hipStreamSynchronize(g->stream));
            dev_free(g->mfirst);
            g->mfirst_cap = 0;
            const int32_t cap = std::max(64, 2 * nbuf);
            if (int s = dev_alloc((void **)&g->mfirst, 8 * (size_t)cap, "merge check starts")) return s;
            g->mfirst_cap = cap;
        }
        HIP_TRY(hipMemcpyAsync(g->mfirst, first.data(), 8 * (size_t)nbuf, hipMemcpyHostToDevice, g->stream));
        launch_merge_check(b->rec, k, g->mfirst, (uint32_t)nbuf, g->mkeys, g->mreq, g->mslots, &g->gc->merge_first,
                           g->stream);
        HIP_TRY(hipMemcpyAsync(&g->gc_host->merge_first, &g->gc->merge_first, sizeof(uint32_t),
                               hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        const uint32_t f = g->gc_host->merge_first;
        if (f < (uint32_t)nbuf) take = std::max<int32_t>(1, (int32_t)f);
    }
    int64_t ki = 0, kd = 0;
    bool kz = false;
    for (int32_t i = 0; i < take; ++i) {
        ki += ni[i];
        kd += nd[i];
        kz = kz || zw[i];
    }
    *applied = take;
    if (first[take] == 0) {
        if (stats) *stats = dsssp_stats{};
        return DSSSP_OK;
    }
    return apply_uploaded(g, t, (int64_t)first[take], ki, kd, kz, stats);
}

}  // namespace dsssp

// ---------------------------------------------------------------------------
// Autotuner: coordinate descent over the launch knobs on sample batches.  The
// graph and tree are snapshotted and restored around every trial, so they end
// as they started; every trial's final tree must digest to the first trial's
// (the knobs never change results).  No reference counterpart.
// ---------------------------------------------------------------------------

namespace {

struct TuneRun {
    dsssp_graph *g;
    dsssp_tree *t;
    dsssp_graph *snap = nullptr;
    uint64_t *st_save = nullptr;
    unsigned long long *digest = nullptr;
    bool deep = false;
    int64_t last_rounds[8] = {};

    int restore()
    {
        if (int s = dsssp_graph_copy(g, snap)) return s;
        DevGuard dg(g->device);
        HIP_TRY(hipMemcpyAsync(t->st, st_save, sizeof(uint64_t) * t->nslots, hipMemcpyDeviceToDevice, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        t->deep = deep;
        memcpy(t->last_rounds, last_rounds, sizeof last_rounds);
        return DSSSP_OK;
    }
    // Applies the batches from the snapshot; *ms = summed device time, *dig =
    // digest of the resulting tree.
    int trial(dsssp_batch *const *batches, int32_t nb, double *ms, unsigned long long *dig)
    {
        if (int s = restore()) return s;
        double sum = 0;
        for (int32_t i = 0; i < nb; ++i) {
            dsssp_stats st{};
            if (int s = dsssp_update_batch(g, t, batches[i], &st)) return s;
            sum += st.t_update_ms;
        }
        DevGuard dg(g->device);
        launch_digest(t->st, t->nslots, digest, g->stream);
        HIP_TRY(hipMemcpyAsync(dig, digest, sizeof *dig, hipMemcpyDeviceToHost, g->stream));
        HIP_TRY(hipStreamSynchronize(g->stream));
        *ms = sum;
        return DSSSP_OK;
    }
};

}  // namespace

extern "C" int dsssp_autotune(dsssp_graph *g, dsssp_tree *t, dsssp_batch *const *batches, int32_t nb,
                              int32_t passes, dsssp_tuning *best, double *best_ms)
{
    clear_error();
    if (!g || !t || !batches || nb < 1 || passes < 1) return fail(DSSSP_ERR_ARG, "bad argument");
    if (t->g != g) return fail(DSSSP_ERR_ARG, "tree belongs to another graph");
    for (int32_t i = 0; i < nb; ++i)
        if (!batches[i] || batches[i]->g != g) return fail(DSSSP_ERR_ARG, "batch %d belongs to another graph", i);
    DevGuard dg(g->device);
    TuneRun R{g, t};
    const dsssp_tuning start = tuning();
    int s = dsssp_graph_clone(g, &R.snap);
    if (!s) s = dev_alloc((void **)&R.st_save, sizeof(uint64_t) * t->nslots, "autotune snapshot");
    if (!s) s = dev_alloc((void **)&R.digest, sizeof(unsigned long long), "autotune digest");
    if (!s && hipMemcpy(R.st_save, t->st, sizeof(uint64_t) * t->nslots, hipMemcpyDeviceToDevice) != hipSuccess)
        s = fail(DSSSP_ERR_HIP, "autotune snapshot failed");
    R.deep = t->deep;
    memcpy(R.last_rounds, t->last_rounds, sizeof R.last_rounds);
    dsssp_tuning cur = start;
    cur.kernel_timing = 0;
    cur.use_graph = 0;
    cur.phase_timing = 0;
    double cur_ms = 0;
    unsigned long long ref = 0;
    if (!s) s = dsssp_set_tuning(&cur);
    if (!s) s = R.trial(batches, nb, &cur_ms, &ref);
    // Candidate values per knob; each pass tries every other value of one knob
    // with the others at their best so far, and keeps a value that is faster by
    // more than 1 %.
    struct Knob {
        int32_t dsssp_tuning::*field;
        std::vector<int32_t> values;
    };
    const std::vector<Knob> knobs = {
        {&dsssp_tuning::block, {128, 256, 512}},
        {&dsssp_tuning::grid, {512, 1024, 2048}},
        {&dsssp_tuning::edges_per_thread, {2, 4}},
        {&dsssp_tuning::heavy_degree, {128, 256, 512}},
        {&dsssp_tuning::tile_entries, {512, 1024, 2048}},
        {&dsssp_tuning::local_tiles, {1, 2, 4}},
        {&dsssp_tuning::round_batch, {4, 8}},
        {&dsssp_tuning::delta, {-1, DSSSP_DELTA_AUTO}},
        {&dsssp_tuning::sweep, {-1, 100, 400}},
        {&dsssp_tuning::rescue_degree, {-1, DSSSP_RESCUE_AUTO, 256, 1024}},
        {&dsssp_tuning::rescue_levels, {4, 16, 32}},
    };
    for (int p = 0; p < passes && !s; ++p)
        for (const Knob &k : knobs) {
            for (int32_t v : k.values) {
                if (s || v == cur.*k.field) continue;
                dsssp_tuning trial = cur;
                trial.*k.field = v;
                double ms = 0;
                unsigned long long dig = 0;
                s = dsssp_set_tuning(&trial);
                if (!s) s = R.trial(batches, nb, &ms, &dig);
                if (!s && dig != ref) s = fail(DSSSP_ERR_STATE, "autotune: a knob setting changed the result");
                if (!s && ms < 0.99 * cur_ms) {
                    cur = trial;
                    cur_ms = ms;
                }
            }
        }
    const std::string keep = s ? last_error() : std::string();
    int r = R.restore();  // the graph and tree as they were
    if (!s) s = r;
    if (R.snap) dsssp_graph_destroy(R.snap);
    dev_free(R.st_save);
    dev_free(R.digest);
    dsssp_tuning fin = s ? start : cur;
    if (!s) {
        fin.kernel_timing = start.kernel_timing;
        fin.use_graph = start.use_graph;
        fin.phase_timing = start.phase_timing;
    }
    (void)dsssp_set_tuning(&fin);
    if (s) return fail(s, "%s", keep.empty() ? last_error() : keep.c_str());
    if (best) *best = fin;
    if (best_ms) *best_ms = cur_ms / nb;
    return DSSSP_OK;
}
