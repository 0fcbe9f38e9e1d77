// kernels.h -- launch interface of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace dsssp {

// Read-mostly view of the device CSR handed to kernels by value.
struct GView {
    const uint64_t *rp;   // n+1 base offsets
    uint64_t *adj;        // base entries
    const uint64_t *dbl;  // delta descriptors
    uint64_t *pool;       // delta entries
};

// Frontier kernel modes.
enum Mode : int {
    M_RELAX = 0,  // push (d[u]+w, u) to every live neighbour with a packed atomicMin
    M_INVAL = 1,  // invalidate children (parent == u) of an invalidated vertex
    M_SEED = 2,   // pull the best live neighbour into an invalidated vertex
    M_TBFS = 3,   // BFS levels over tight edges (zero-weight parent repair)
    M_PICK = 4,   // smallest-id tight neighbour (optionally one level up) as parent
};

struct FArgs {
    GView g;
    uint64_t *st;
    uint32_t *stamp;
    uint32_t *lvl;
    uint32_t *best;
    // input: light queue (in_q == nullptr means the identity list [0, in_n)),
    // its length at *in_cnt (or in_n when in_cnt == nullptr), heavy chunks
    const uint32_t *in_q;
    const uint32_t *in_cnt;
    uint32_t in_n;
    const uint64_t *in_hq;
    const uint32_t *in_hcnt;
    // output queues
    uint32_t *out_q;
    uint32_t *out_cnt;
    uint32_t out_cap;
    uint64_t *out_hq;
    uint32_t *out_hcnt;
    uint32_t out_hcap;
    uint32_t *zero_cnt;  // two counters cleared by workgroup 0 (next round's output)
    uint32_t *invl;      // M_INVAL: list of every invalidated vertex
    uint32_t *inv_total;
    Counters *c;
    uint32_t epoch;   // stamp value of this round's enqueues
    uint32_t heavy;   // rows longer than this are deferred as chunks (loop modes)
    uint32_t chunk;   // entries per chunk, multiple of 64
    uint32_t use_lvl; // M_PICK: require lvl[parent] + 1 == lvl[v]
    uint32_t kshift;  // log2 of the trees per vertex: state slot = vertex << kshift | tree
    // dev_iter = 1: in/out queues, counters and epoch come from Counters'
    // round indexing (identical arguments for every round of a loop phase,
    // so rounds can be replayed from a captured hipGraph)
    uint32_t dev_iter;
    uint32_t *qv[2];
    uint64_t *hqv[2];
};

constexpr int FRONTIER_BLOCK = 256;

void launch_frontier(int mode, const FArgs &a, int grid, hipStream_t s);
void launch_reset_counters(Counters *c, hipStream_t s);
void launch_reset_graphctl(GraphCtl *gc, hipStream_t s);
void launch_ins_count(const int4 *rec, int64_t k, uint32_t *tcnt, uint64_t *slot, uint32_t *touched,
                      GraphCtl *gc, hipStream_t s);
void launch_ins_alloc(const uint32_t *touched, uint32_t max_touched, uint64_t *dbl, uint32_t *tcnt,
                      uint64_t *pool, uint64_t pool_cap, GraphCtl *gc, hipStream_t s);
void launch_ins_scatter(const int4 *rec, int64_t k, const uint64_t *slot, const uint64_t *dbl,
                        const uint32_t *tcnt, uint64_t *pool, uint64_t *pos, hipStream_t s);
void launch_ins_reset(const uint32_t *touched, uint32_t max_touched, uint32_t *tcnt, GraphCtl *gc,
                      hipStream_t s);
void launch_classify(const int4 *rec, int64_t k, uint32_t kshift, GView g, uint64_t *st, uint32_t *out_q,
                     uint32_t *out_cnt, uint32_t *invl, uint32_t *inv_total, uint32_t cap, Counters *c, hipStream_t s);
void launch_ins_offer(const int4 *rec, int64_t k, uint32_t kshift, const uint64_t *pos, const uint64_t *pool,
                      uint64_t *st, uint32_t *stamp, uint32_t epoch, uint32_t *out_q, uint32_t *out_cnt, uint32_t cap,
                      Counters *c, hipStream_t s);
// Roots: for j < count, vertex src[j] becomes a root (distance 0, parent itself)
// of tree (per_tree ? j : 0) and is queued.
void launch_set_sources(uint64_t *st, const uint32_t *src, uint32_t count, uint32_t kshift, bool per_tree,
                        uint32_t *stamp, uint32_t epoch, uint32_t *q, uint32_t *cnt, uint32_t cap, Counters *c,
                        hipStream_t s);
void launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
void launch_lvl_init(uint32_t nslots, uint32_t kshift, const uint64_t *st, uint32_t *lvl, uint32_t *best,
                     uint32_t *stamp, uint32_t epoch, uint32_t *out_q, uint32_t *out_cnt, uint32_t cap, Counters *c,
                     hipStream_t s);
// Starts a loop phase's round indexing (iter = 0).
void launch_phase_init(Counters *c, uint32_t ring_base, uint32_t q_base, uint32_t epoch_base, hipStream_t s);
void launch_apply_best(uint32_t nslots, uint32_t kshift, uint64_t *st, const uint32_t *best, Counters *c,
                       hipStream_t s);

}  // namespace dsssp
