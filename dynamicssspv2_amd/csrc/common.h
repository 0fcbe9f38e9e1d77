// common.h -- layouts and constants shared by the host code and the gfx950 kernels.
//
// HBM layout (all arrays 64-bit aligned, one allocation each):
//   graph  vr   VRec[n]       per-vertex record (32 B): base row offset and
//                             length, delta block capacity, delta block offset
//                             and length -- one cache line holds two
//          adj  uint64[E]     base entries (neighbour << 32 | weight), rows sorted
//                             ascending, i.e. by (neighbour, weight)
//          pool uint64[cap]   delta blocks: entries appended by insertions (rows
//                             unsorted); a block's capacity doubles when outgrown
//          emap uint64[2^k]   edge map: per live directed entry, a 32-bit key
//                             fingerprint and its position in its row
//          tbm  uint64[n/64]  rows whose delta block must grow in this batch
//          wsum/bsum/bbase    pool space per bitmap word / per 256 words, and
//                             (graphs past 2048 x 256 words) the pool offset of
//                             each 256-word group, for one batch
//   tree   st   uint64[S]     packed state (distance << 32 | parent), S = n * K'
//                             slots (K' = trees rounded up to a power of two)
//          fb   2 x uint64[S/64]  frontier bitmaps (ping-pong between rounds)
//          ib   uint64[S/64]  bitmap of slots invalidated by the current update
//          hq   2 x 2 x uint64[h] deferred pieces of rows (items)
// The weight's bit 31 is the tombstone bit (weights are >= 0), playing the role
// of the reference's wt = -1 (gpuFunctions_undir.cuh:203,210,321,384).
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#define DSSSP_HD __host__ __device__ __forceinline__
#else
#define DSSSP_HD inline
#endif

namespace dsssp {

constexpr uint64_t ST_INF = ~0ull;          // unreachable / invalidated: dist 2^32-1, parent -1
constexpr uint32_t DIST_INF = 0xFFFFFFFFu;  // distances must stay below this
constexpr uint32_t W_DEAD = 0x80000000u;    // tombstone bit in an entry's weight
constexpr uint64_t E_DEAD = 0x80000000ull;  // the same bit in a packed entry
constexpr uint32_t LVL_INF = 0xFFFFFFFFu;

DSSSP_HD uint64_t pack_state(uint32_t d, uint32_t p) { return ((uint64_t)d << 32) | p; }
DSSSP_HD uint32_t st_dist(uint64_t s) { return (uint32_t)(s >> 32); }
DSSSP_HD uint32_t st_parent(uint64_t s) { return (uint32_t)s; }
DSSSP_HD uint64_t pack_entry(uint32_t col, uint32_t w) { return ((uint64_t)col << 32) | w; }
DSSSP_HD uint32_t e_col(uint64_t e) { return (uint32_t)(e >> 32); }
DSSSP_HD uint32_t e_w(uint64_t e) { return (uint32_t)e; }
DSSSP_HD uint64_t e_key(uint64_t e) { return e & ~E_DEAD; }

// Per-vertex record.  A row's entries are its base entries adj[boff, boff+blen)
// followed by its delta entries pool[dbl >> 32, (dbl >> 32) + (uint32)dbl); a
// position in a row counts base entries first.
struct VRec {
    uint64_t boff;  // first base entry
    uint32_t blen;  // base entries
    uint32_t dcap;  // delta block capacity
    uint64_t dbl;   // delta block: pool offset << 32 | length
    uint32_t blo;   // base entries whose neighbour id is below the vertex's own
                    // (a prefix of the sorted row; positions never move)
    uint32_t pad;
};
static_assert(sizeof(VRec) == 32, "VRec is 32 bytes");

// Error bits raised by kernels (Counters::err / GraphCtl::err).
constexpr uint32_t ERR_DIST_OVERFLOW = 1u;
constexpr uint32_t ERR_POOL = 2u;
constexpr uint32_t ERR_QUEUE = 4u;
constexpr uint32_t ERR_NO_TIGHT = 8u;
// A candidate of 2^32-1 or more was offered to an unreachable slot.  That is an
// overflow only if the slot is still unreachable at convergence; a check pass
// over the converged tree decides (a finite slot with a live edge to an
// unreachable one proves it).
constexpr uint32_t ERR_DIST_SUSPECT = 16u;

// Statistics, accumulated per workgroup slot (no single-address hot spot) and
// summed into Counters::stats once per update.
enum Stat : int {
    ST_RELAX_FRONT = 0,  // frontier slots processed by relaxation
    ST_RELAX_EDGES,      // adjacency entries scanned by relaxation
    ST_RELAX_MARKS,      // successful improvements (bits set) by relaxation
    ST_OTHER_EDGES,      // entries scanned by invalidation, seeding, repair
    ST_INVALIDATED,      // slots seeded (= invalidated by the update)
    ST_PROBES,           // entries read by deletion searches
    ST_SWEEPS,           // relaxation rounds run as full sweeps (every slot visited)
    ST_SWEEP_GATHERS,    // entries of an M_SWEEP pass that read their neighbour's state
    ST_KILLS,            // entries tombstoned by deletions
    ST_RESCUED,          // invalidation candidates kept through another tight parent
    NSTAT
};
constexpr int STAT_SLOTS = 2048;

// Device-side control block of one tree.
struct Counters {
    uint32_t act[3];    // a round marked its output bitmap; indexed by round % 3
    uint32_t hcnt[3];   // heavy chunks queued for the next round; round % 3
    uint32_t worked;    // rounds of the current loop that had input (bits or chunks)
    // Distance-ordered relaxation (a threshold on the slots a round visits),
    // indexed by round % 3 like act:
    uint32_t near[3];      // a round marked a slot below its threshold
    uint32_t thr[3];       // the threshold a round used
    uint32_t fenc[3][64];  // ~(smallest distance marked at or above it), max-reduced over 64 words
    // Push-vs-full-sweep: slots a relaxation round marked in its output bitmap,
    // spread over 64 words (summed by the next round); round % 3.
    uint32_t nmark[3][64];
    // Rescue of invalidation candidates: stage j of a loop reads cand[j & 1]
    // (set when the step before it left a candidate) and clears the other.
    uint32_t cand[2];
    uint32_t aux_hcnt;  // heavy chunks queued by a single-pass stage
    uint32_t err;
    unsigned long long stats[NSTAT];
    unsigned long long popc;  // set bits counted by the last bitmap merge
    // The graph's words an update's caller reads (GraphCtl::pool_top, err),
    // copied here by the statistics reduction, so that one copy to the host
    // brings the results of both.
    unsigned long long g_pool_top;
    uint32_t g_err;
    uint32_t g_pad;
};

// Device-side control block of one graph (delta pool allocation, upkeep).
struct GraphCtl {
    unsigned long long pool_top;   // next free pool entry
    unsigned long long batch_new;  // pool entries placed by the current batch
    uint32_t err;
    uint32_t pad;
    unsigned long long live[2];    // repack: live base / delta entries counted
    unsigned long long total;      // repack: pool entries of the new delta blocks
    uint32_t merge_first;          // merge check: first request that cannot join the merge (~0: none)
    uint32_t pad2;
};

}  // namespace dsssp
