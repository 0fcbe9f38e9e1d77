// common.h -- layouts and constants shared by the host code and the gfx950 kernels.
//
// HBM layout (all arrays 64-bit aligned, one allocation each):
//   graph  rp   uint64[n+1]   base CSR row offsets
//          adj  uint64[E]     base entries (neighbour << 32 | weight), rows sorted
//                             ascending, i.e. by (neighbour, weight)
//          dbl  uint64[n]     delta row descriptor (pool offset << 32 | length)
//          pool uint64[cap]   delta entries appended by insertions (rows unsorted)
//          tcnt uint32[n]     per-vertex insertion counters, zero between updates
//   tree   st   uint64[n]     packed state (distance << 32 | parent)
//          stamp uint32[n]    epoch of the last enqueue (frontier de-duplication)
//          q    2 x uint32[n] light frontier queues (ping-pong)
//          hq   2 x uint64[h] heavy-row chunk queues (vertex << 32 | chunk)
//          invl uint32[n]     every vertex invalidated by the current update
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

// Error bits raised by kernels (Counters::err).
constexpr uint32_t ERR_DIST_OVERFLOW = 1u;
constexpr uint32_t ERR_POOL = 2u;
constexpr uint32_t ERR_QUEUE = 4u;
constexpr uint32_t ERR_NO_TIGHT = 8u;

// Device-side control block of one tree: ring counters, statistics, errors.
struct Counters {
    uint32_t ring[3][2];  // (light, heavy) queue lengths, indexed by round % 3
    uint32_t inv_total;   // length of invl
    uint32_t err;
    uint32_t aux_hcnt;    // heavy chunks queued by a single-pass stage
    // Device-side round indexing of a loop phase: round i reads ring slot
    // (ring_base + i) % 3 and queue buffer (q_base + i) & 1 and stamps with
    // epoch_base + i; the last workgroup to finish a round advances iter.
    uint32_t iter;
    uint32_t blocks_done;
    uint32_t ring_base;
    uint32_t q_base;
    uint32_t epoch_base;
    unsigned long long relax_frontier, relax_edges, relax_improved, relax_enq;
    unsigned long long other_edges, probes, tombstones;
};

// Device-side control block of one graph (delta pool allocation).
struct GraphCtl {
    unsigned long long pool_top;  // next free pool entry
    uint32_t touched;             // vertices touched by the current batch's insertions
    uint32_t err;
};

}  // namespace dsssp
