// kernels.h -- launch interface of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "common.h"

namespace dsssp {

// Read-mostly view of the device CSR handed to kernels by value.
struct GView {
    VRec *vr;             // per-vertex records
    uint64_t *adj;        // base entries
    uint64_t *pool;       // delta entries
    uint64_t *emap;       // edge map: slots of fingerprint << 32 | position in row
    uint64_t emask;       // edge map slots - 1
};

// Frontier kernel modes.
enum Mode : int {
    M_RELAX = 0,  // push (d[u]+w, u) to every live neighbour with a packed atomicMin
    M_INVAL = 1,  // invalidate children (parent == u) of an invalidated vertex
    M_SEED = 2,   // pull the best live neighbour into an invalidated vertex
    M_TBFS = 3,   // BFS levels over tight edges (zero-weight parent repair)
    M_PICK = 4,   // smallest-id tight neighbour (optionally one level up) as parent
    M_OFCHECK = 5,  // a reachable slot with a live edge to an unreachable one: overflow
    M_EMAP = 6,     // enter every live entry into the edge map (identity pass)
    // One pass over every live edge, both directions at once, from the row of
    // its higher-id endpoint only (base rows are sorted by neighbour, so that
    // part is a prefix): push (d[v]+w, v) to the neighbour, pull (d[u]+w, u)
    // into the owner.  Identity frontier; not with 64 trees per vertex.
    M_SWEEP = 7,
    // INVAL that leaves rescue candidates (A.rescue_deg, A.cand_bm): a child
    // with a long row is marked in cand_bm instead of being reset.  A mode of
    // its own so that plain invalidation rounds keep their register count.
    M_INVAL_RESCUE = 8,
};
constexpr bool is_inval(int mode) { return mode == M_INVAL || mode == M_INVAL_RESCUE; }

// Arguments of one frontier launch.  The input frontier is a bitmap over
// state slots (or every slot when identity is set, or in a full-sweep round)
// plus a list of heavy-row chunks; marks go to the output bitmap.
struct FArgs {
    GView g;
    uint64_t *st;
    uint32_t *lvl;
    uint32_t *best;
    uint64_t *in_bm;        // input frontier bitmap
    const uint32_t *in_act; // set when the previous launch marked in_bm (null: always scan)
    const uint64_t *in_hq;  // heavy chunks to process
    const uint32_t *in_hcnt;
    uint64_t *out_bm;       // marks of this launch
    uint64_t *inv_bm;       // M_INVAL: every invalidated slot is also marked here
    uint32_t *out_act;      // set to 1 when a mark is made
    uint64_t *out_hq;       // heavy chunks found by this launch
    uint32_t *out_hcnt;
    uint32_t *zero_act;     // cleared by workgroup 0 (the flag two rounds ahead)
    uint32_t *zero_hcnt;
    Counters *c;
    unsigned long long *bstats;  // STAT_SLOTS x NSTAT per-workgroup statistics
    // INVAL with rescue_deg set: a child whose row has at least rescue_deg
    // entries is not reset but marked in cand_bm (and *cand_flag set), to be
    // kept by the rescue stage if another tight neighbour still holds it.
    uint64_t *cand_bm;
    uint32_t *cand_flag;
    uint32_t nslots;        // slots = n << kshift
    uint32_t nwords;        // words of a bitmap = ceil(nslots / 64)
    uint32_t hcap;          // capacity of each heavy-chunk list
    uint32_t identity;      // 1: the input frontier is every slot
    uint32_t consume;       // 1: clear input bitmap words as they are read
    uint32_t tiles;         // 0: process only the heavy chunks
    uint32_t heavy;         // rows longer than this are split into chunks
    uint32_t chunk;         // entries per chunk, multiple of 64
    uint32_t use_lvl;       // M_PICK: require lvl[parent] + 1 == lvl[v]
    uint32_t kshift;        // log2 of the trees per vertex: slot = vertex << kshift | tree
    uint32_t ept;           // edges in flight per lane: 1, 2 or 4
    uint32_t budget;        // entries a fresh tile takes; later slots become items
    uint32_t local;         // push modes: tiles of newly marked slots a wave visits itself
    // RELAX with delta > 0: a round visits only slots below its threshold and
    // leaves the others marked; when the previous round marked nothing below
    // its threshold, the threshold moves to the smallest marked distance plus
    // delta.  SEED then records the distances it marks.  Ring slot `ring` of
    // the Counters' near / thr / fenc is this launch's input, ring + 1 its
    // output, ring + 2 is cleared.
    uint32_t delta;
    uint32_t ring;
    // RELAX with delta off (push-vs-full-sweep): when the previous round marked
    // at least `sweep` slots (ring slot `ring` of Counters::nmark), this round
    // relaxes every slot, its rows read in vertex order, instead of the marked
    // ones; it still consumes the input bitmap.  0: never.  Rounds count their
    // marks into ring slot ring + 1 only when sweep is set.
    uint32_t sweep;
    uint32_t block;  // threads per workgroup: 128, 256 or 512 (the launch's grid is in workgroups of this size)
    uint32_t rescue_deg;
    uint32_t count_round;  // a round of a loop: counts itself in Counters::worked when it has input
    uint32_t pad_;
};

constexpr int FRONTIER_BLOCK = 256;  // default workgroup size of frontier launches

// With start/stop events, they are recorded by the dispatch itself
// (hipExtLaunchKernel), bracketing exactly this kernel.
void launch_frontier(int mode, const FArgs &a, int grid, hipStream_t s, hipEvent_t start = nullptr,
                     hipEvent_t stop = nullptr);
void launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s);
void launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
void launch_zero2_u32(uint32_t *p, uint64_t n, uint32_t *q, uint64_t m, hipStream_t s);
// Sums the per-workgroup statistics into c->stats; with gc, also copies the
// graph's pool_top and err into c->g_pool_top / g_err.
void launch_reduce_stats(Counters *c, unsigned long long *bstats, hipStream_t s, const GraphCtl *gc = nullptr);
// Checks k records in device memory: cnt[0..2] = insertions, deletions,
// zero-weight insertions; cnt[3] = smallest index of an invalid record (~0: none).
void launch_validate(const int4 *rec, int64_t k, int64_t n, unsigned long long *cnt, hipStream_t s);
// Probe launches of known footprint: kind 0 copies in[0, n) to out; kind 1 reads
// count words at hashed positions of in[0, n) (8192 x 256 threads, out holds one
// word per thread).
void launch_probe(int kind, const uint64_t *in, uint64_t n, uint64_t count, uint64_t *out, hipStream_t s);
// dst[i] |= src[i] and src[i] = 0 for i < nwords; *count = set bits src held.
void launch_merge_count(uint64_t *dst, uint64_t *src, uint64_t nwords, unsigned long long *count, hipStream_t s);
// out[v] = st[v << kshift | tree] for v < n (one tree of a forest, contiguous).
void launch_gather_tree(const uint64_t *st, uint64_t n, uint32_t kshift, uint32_t tree, uint64_t *out,
                        hipStream_t s);
// *out = an order-independent digest of p[0, n) (equal arrays, equal digests).
void launch_digest(const uint64_t *p, uint64_t n, unsigned long long *out, hipStream_t s);
// Insertion of a batch into the delta blocks.  Thread 2i+s places copy s of
// record i (tslot[2i+s] = its position in the delta block, bit 31 set while
// pending); rows that overflow are marked in tbm and grown: pool space per
// bitmap word (wsum) and per 256-word group (bsum), the block moves (each
// group's base from the pool top and the sums before it), then the pending
// copies (and the pool top's advance).
constexpr uint32_t SLOT_PENDING = 0x80000000u;
void launch_ins_place(const int4 *rec, int64_t k, GView g, uint32_t *tslot, uint64_t *tbm, hipStream_t s);
void launch_ins_wordsum(const uint64_t *tbm, uint32_t nwords, const VRec *vr, uint32_t *wsum, uint64_t *bsum,
                        hipStream_t s);
// Up to this many 256-word groups, each workgroup of the block moves sums the
// group sums before its own (no scan launch); beyond it a single-workgroup scan
// into bbase (nwords / 256 entries) runs first, so the work stays linear.
constexpr uint32_t INS_FOLD_GROUPS = 2048;
void launch_ins_alloc(uint64_t *tbm, uint32_t nwords, const uint32_t *wsum, const uint64_t *bsum, uint64_t *bbase,
                      VRec *vr, uint64_t *pool, uint64_t pool_cap, GraphCtl *gc, hipStream_t s);
void launch_ins_pending(const int4 *rec, int64_t k, GView g, uint32_t *tslot, GraphCtl *gc, hipStream_t s);
// Deletions: the tree-edge test (marks the roots of invalidated subtrees) and
// the tombstoning of every copy of each deleted entry.
// With rescue_deg set, a victim whose row has at least rescue_deg entries is
// marked in cand_bm (and *cand_flag set) instead of being reset.
void launch_classify(const int4 *rec, int64_t k, uint32_t kshift, uint64_t *st, uint64_t *out_bm, uint64_t *inv_bm,
                     uint32_t *out_act, const VRec *vr, uint32_t rescue_deg, uint64_t *cand_bm, uint32_t *cand_flag,
                     hipStream_t s);
// Rescue commit over the candidate bitmap (consumed): a candidate with a pick in
// best (the smallest-id tight neighbour that was valid) keeps its distance with
// that parent (best reset to ~0); the others are reset and marked in inv_bm and
// out_bm (*out_act set).  Returns at once when *cand_flag is clear; clears
// *clear_flag and *aux_hcnt for the next stage.
void launch_rescue_commit(uint64_t *cand_bm, uint32_t nwords, uint64_t *st, uint32_t *best, uint64_t *inv_bm,
                          uint64_t *out_bm, uint32_t *out_act, const uint32_t *cand_flag, uint32_t *clear_flag,
                          uint32_t *aux_hcnt, unsigned long long *bstats, hipStream_t s);
// Tombstones every copy of each deleted entry, found through the edge map; the
// copies tombstoned are counted in the statistics (ST_KILLS).
void launch_tomb(const int4 *rec, int64_t k, GView g, unsigned long long *bstats, hipStream_t s);
// Merge check of k records made of nreq requests (request q starts at record
// first_rec[q]; first_rec has nreq entries): *first = the smallest q whose
// insertions include a key (a, b, w) that a request before q deletes, ~0 if
// none.  keys/req: a hash table of `slots` entries (a power of two, at least
// twice the deletions), cleared here.
void launch_merge_check(const int4 *rec, int64_t k, const uint64_t *first_rec, uint32_t nreq,
                        unsigned long long *keys, uint32_t *req, uint64_t slots, uint32_t *first, hipStream_t s);
// Repack of the resident CSR: live base entries to the front of each row
// (blen, blo rewritten; cap[v] = capacity of the row's new delta block; live
// totals added to gc->live), the exclusive scan of cap (bbase per RB rows, the
// sum in *total), then the live delta entries copied to npool (dbl, dcap
// rewritten).
void launch_repack_rows(VRec *vr, uint64_t *adj, const uint64_t *pool, uint64_t n, uint32_t min_cap, uint32_t *cap,
                        GraphCtl *gc, hipStream_t s);
void launch_repack_scan(const uint32_t *cap, uint64_t n, uint64_t *bsum, uint64_t *bbase, unsigned long long *total,
                        hipStream_t s);
void launch_repack_delta(VRec *vr, const uint64_t *pool, uint64_t *npool, uint64_t n, const uint32_t *cap,
                         const uint64_t *bbase, hipStream_t s);
// With far set, the distances it marks go to ring slot 0 of c->fenc.
void launch_ins_offer(const int4 *rec, int64_t k, uint32_t kshift, const uint32_t *tslot, GView g, uint64_t *st,
                      uint64_t *out_bm, uint32_t *out_act, Counters *c, bool far, hipStream_t s);
// Roots: for j < count, vertex src[j] becomes a root (distance 0, parent itself)
// of tree (per_tree ? j : 0) and is marked.
void launch_set_sources(uint64_t *st, const uint32_t *src, uint32_t count, uint32_t kshift, bool per_tree,
                        uint64_t *out_bm, uint32_t *out_act, hipStream_t s);
void launch_lvl_init(uint32_t nslots, uint32_t kshift, const uint64_t *st, uint32_t *lvl, uint32_t *best,
                     uint64_t *out_bm, uint32_t *out_act, hipStream_t s);
void launch_apply_best(uint32_t nslots, uint32_t kshift, uint64_t *st, const uint32_t *best, Counters *c,
                       hipStream_t s);

}  // namespace dsssp
