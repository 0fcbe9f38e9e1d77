// kernels.hip -- hand-written gfx950 kernels of the dynamic SSSP update.
//
// Work mapping (CDNA4, wave64):
//  * A frontier is a bitmap over state slots.  Marking a slot is one
//    non-returning 64-bit atomicOr at a random address, so there is no shared
//    queue counter and duplicates vanish by construction.  A round's kernel
//    consumes the input bitmap word by word (64 lanes load 64 words, a ballot
//    finds the non-zero ones) and clears each word it reads, so the bitmap is
//    clean again for use two rounds later.
//  * One non-zero word is a tile of up to 64 consecutive slots.  Each lane loads
//    one slot's row bounds, the wave forms an exclusive prefix sum of the
//    degrees with shuffle steps, stages the per-slot fields in LDS, and lane l
//    then handles edge base+l (and base+l+64*j for EPT edges in flight) of the
//    tile's concatenated rows, finding its owner with a 6-step binary search in
//    LDS.  Neighbouring lanes read neighbouring 8-byte entries (coalesced), and
//    consecutive slots share cache lines of the row offsets and states.
//  * Rows longer than `heavy` entries are not scanned inside a tile: they are
//    split into `chunk`-entry pieces on a chunk list processed by the next
//    launch, one chunk per wave, so an R-MAT hub never serialises a wave.
//  * Distance and parent live in one uint64 (dist << 32 | parent) and change
//    together through one 64-bit atomicMin, so the smaller distance wins and,
//    among equal distances, the smaller parent id.  Every atomic in the edge
//    loop is non-returning; the returning ones are one per tile with split
//    rows (item list), the per-row insertion counters and the edge-map claims
//    of delta entries past a row's first few.
//  * A wave keeps the slots it has just marked on an LDS stack and visits a
//    bounded number of tiles of them in the same launch; with a distance step
//    (delta) a round visits only slots below a threshold that every workgroup
//    derives from the previous round's marks.
//  * Statistics go to one slot per workgroup and are summed once per update;
//    no kernel issues device-scope fences.
// The reference does one thread per frontier vertex with unsynchronised
// read-modify-writes (gpuFunctions_undir.cuh:370-433) and one full O(n) filter
// pass over 16-byte records per round (:125-174); none of that is reused here.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace dsssp {

namespace {

constexpr int RB = 256;  // block size of per-record / per-vertex kernels

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

__device__ __forceinline__ void mark(uint64_t *bm, uint32_t s)
{
    atomicOr((unsigned long long *)&bm[s >> 6], 1ull << (s & 63u));
}

__device__ __forceinline__ void add_stat(unsigned long long *bstats, int which, uint64_t v)
{
    if (v) atomicAdd(&bstats[(blockIdx.x % STAT_SLOTS) * NSTAT + which], (unsigned long long)v);
}

// Edge map.  Every live directed entry (row, col, w) has a slot holding a
// 32-bit fingerprint of the key and the entry's position in its row (base
// entries first, then the delta block; positions never change, as blocks move
// in order).  A lookup verifies the entry itself, so it is exact; ~0 = empty.
// The first EMAP_DSKIP delta entries of a row are not in the map: a deletion
// scans them (a row's first delta entries share one or two 64-byte sectors),
// so the common insertion, into a short delta row, claims no map slot.
constexpr uint32_t EMAP_DSKIP = 8;

__device__ __forceinline__ bool emap_keeps(uint32_t pos, uint32_t blen) { return pos < blen || pos - blen >= EMAP_DSKIP; }
__device__ __forceinline__ uint64_t emap_hash(uint32_t row, uint32_t col, uint32_t w)
{
    uint64_t h = ((uint64_t)row << 32 | col) * 0x9E3779B97F4A7C15ull;
    h ^= (uint64_t)w * 0xC2B2AE3D27D4EB4Full;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    return h ^ (h >> 29);
}

__device__ __forceinline__ uint32_t emap_fp(uint64_t h)
{
    const uint32_t f = (uint32_t)(h >> 32);
    return f == 0xFFFFFFFFu ? 0xFFFFFFFEu : f;
}

__device__ __forceinline__ void emap_insert(const GView &g, uint32_t row, uint32_t col, uint32_t w, uint32_t pos)
{
    const uint64_t h = emap_hash(row, col, w);
    const uint64_t v = (uint64_t)emap_fp(h) << 32 | pos;
    for (uint64_t s = h & g.emask;; s = (s + 1) & g.emask)
        if (atomicCAS((unsigned long long *)&g.emap[s], ~0ull, (unsigned long long)v) == ~0ull) return;
}

// Tombstones every live copy of (row, col, w); returns map slots and delta
// entries read, and adds the copies it tombstoned to `kills`.
__device__ __forceinline__ uint32_t emap_kill(const GView &g, uint32_t row, uint32_t col, uint32_t w, uint32_t &kills)
{
    const uint64_t h = emap_hash(row, col, w);
    const uint32_t fp = emap_fp(h);
    const uint64_t key = pack_entry(col, w);
    // The first map slot depends only on the key: its load goes out with the
    // vertex record's.
    uint64_t s = h & g.emask;
    uint64_t v = g.emap[s];
    const uint64_t b0 = g.vr[row].boff;
    const uint32_t blen = g.vr[row].blen;
    const uint64_t d = g.vr[row].dbl;
    const uint32_t dl = (uint32_t)d;
    // The row's first delta entries (never in the map, so no kill of the probe
    // below touches them) are loaded before the probe starts.
    const uint32_t ns = min(dl, EMAP_DSKIP);
    uint64_t *dp = &g.pool[d >> 32];
    uint64_t de[EMAP_DSKIP];
#pragma unroll
    for (uint32_t i = 0; i < EMAP_DSKIP; ++i) de[i] = i < ns ? dp[i] : E_DEAD;
    uint32_t probes = 0;
    for (;;) {
        ++probes;
        if (v == ~0ull) break;
        const uint32_t pos = (uint32_t)v;
        if ((uint32_t)(v >> 32) == fp && pos < blen + dl) {
            uint64_t *ep = pos < blen ? &g.adj[b0 + pos] : &g.pool[(d >> 32) + (pos - blen)];
            const uint64_t e = *ep;
            if (e_key(e) == key && !(e & E_DEAD)) {
                *ep = e | E_DEAD;
                ++kills;  // (duplicate records killing one copy together count it twice)
            }
        }
        s = (s + 1) & g.emask;
        v = g.emap[s];
    }
#pragma unroll
    for (uint32_t i = 0; i < EMAP_DSKIP; ++i)
        if (!(de[i] & E_DEAD) && e_key(de[i]) == key) {  // (entries past ns read as dead)
            dp[i] = de[i] | E_DEAD;
            ++kills;
        }
    return probes + ns;
}

struct WaveStats {
    uint32_t front;  // wave-uniform
    uint64_t edges;  // wave-uniform
    uint32_t marks;  // per lane
    bool any;        // per lane: this lane set an output bit
    // Distance-ordered relaxation: this round's threshold (wave-uniform; 0 in
    // SEED, whose marks all count as far), and per lane whether a mark fell
    // below it and ~(the smallest distance marked at or above it).
    uint32_t thr;
    bool near;
    uint32_t fenc;
    uint32_t outm;  // wave-uniform: slots marked in the output bitmap (RELAX with sweep set)
    uint32_t gath;  // per lane: SWEEP entries that read their neighbour's state
    bool cand;      // per lane: INVAL marked a rescue candidate
};

// Raises *p to v.  A relaxed load first skips the atomic when v is no record
// (a stale read only costs an atomic).
__device__ __forceinline__ void raise_u32(uint32_t *p, uint32_t v)
{
    if (v > __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMax(p, v);
}

// Bookkeeping of a slot marked for the next round with distance d.
__device__ __forceinline__ void note_mark(WaveStats &S, uint32_t d)
{
    if (d < S.thr)
        S.near = true;
    else
        S.fenc = max(S.fenc, ~d);
}

// Each edge is handled in two steps so that a lane can keep several edges in
// flight: gather() issues the loads of the neighbour's state, apply() decides
// and writes.  The owner slot os is vertex os >> kshift of tree os & mask.
struct Pre {
    uint64_t sx;  // neighbour state
    uint32_t lx;  // neighbour BFS level (TBFS / PICK)
    bool skip;
};

template <int MODE>
__device__ __forceinline__ Pre gather(const FArgs &A, uint32_t os, uint64_t ent)
{
    Pre p{ST_INF, LVL_INF, true};
    const uint32_t u = os >> A.kshift;
    const uint32_t x = e_col(ent);
    if (x == u) return p;  // a self-loop never carries a shortest path
    if constexpr (MODE == M_SWEEP) {
        if (x > u) return p;  // relaxed from the row of x, the higher id
    }
    if constexpr (!is_inval(MODE)) {
        if (e_w(ent) & W_DEAD) return p;
    }
    const uint32_t xs = (x << A.kshift) | (os & ((1u << A.kshift) - 1u));
    p.skip = false;
    p.sx = A.st[xs];  // possibly stale within a round: only ever too large
    if constexpr (MODE == M_TBFS) p.lx = A.lvl[xs];
    if constexpr (MODE == M_PICK) {
        if (A.use_lvl) p.lx = A.lvl[xs];
    }
    return p;
}

// Applies one edge.  The push modes (RELAX, INVAL, TBFS) return the slot to
// visit next (~0u: none); a slot found through a stale state read is harmless
// to visit, it changed anyway.  The pull modes (SEED, PICK) only lower `pull`,
// the owner's candidate.
constexpr uint32_t NO_SLOT = 0xFFFFFFFFu;

template <int MODE>
__device__ __forceinline__ uint32_t apply(const FArgs &A, uint32_t os, uint32_t aux0, uint32_t aux1, uint64_t ent,
                                          const Pre &p, WaveStats &S, uint64_t &pull, uint32_t &td)
{
    if (p.skip) return NO_SLOT;
    const uint32_t u = os >> A.kshift;
    const uint32_t x = e_col(ent);
    const uint32_t w = e_w(ent);
    const uint32_t xs = (x << A.kshift) | (os & ((1u << A.kshift) - 1u));
    if constexpr (MODE == M_RELAX) {
        const uint64_t nd = (uint64_t)aux0 + w;
        if (nd >= DIST_INF) {  // cannot beat a finite distance
            if (p.sx == ST_INF) atomicOr(&A.c->err, ERR_DIST_SUSPECT);
            return NO_SLOT;
        }
        const uint64_t cand = pack_state((uint32_t)nd, u);
        // A root (parent itself, distance 0) never changes: with zero weights a
        // candidate (0, smaller id) would otherwise take its place.
        if (cand >= p.sx || st_parent(p.sx) == x) return NO_SLOT;
        atomicMin((unsigned long long *)&A.st[xs], (unsigned long long)cand);
        S.marks++;
        td = (uint32_t)nd;
        return xs;
    } else if constexpr (MODE == M_SWEEP) {
        // Edge {u, x}, x < u, in both directions.  Pull: the owner's candidate
        // (reduced per owner by the caller).  aux0 is the owner's distance
        // (DIST_INF when unreachable).
        S.gath++;
        if (p.sx != ST_INF) {
            const uint64_t nd = (uint64_t)st_dist(p.sx) + w;
            if (nd < DIST_INF)
                pull = min(pull, pack_state((uint32_t)nd, x));
            else if (aux0 == DIST_INF)
                atomicOr(&A.c->err, ERR_DIST_SUSPECT);
        }
        // Push, as RELAX: the neighbour's candidate through the owner.
        if (aux0 == DIST_INF) return NO_SLOT;
        const uint64_t nd = (uint64_t)aux0 + w;
        if (nd >= DIST_INF) {
            if (p.sx == ST_INF) atomicOr(&A.c->err, ERR_DIST_SUSPECT);
            return NO_SLOT;
        }
        const uint64_t cand = pack_state((uint32_t)nd, u);
        if (cand >= p.sx || st_parent(p.sx) == x) return NO_SLOT;
        atomicMin((unsigned long long *)&A.st[xs], (unsigned long long)cand);
        S.marks++;
        return xs;
    } else if constexpr (is_inval(MODE)) {
        // x is a child of u in the pre-update tree: reset it and descend.
        // Tombstoned entries are scanned too, so a child hanging off a deleted
        // parallel copy is still found.  Only ST_INF is ever written in this
        // phase, so a plain store is race-free.
        if (p.sx == ST_INF || st_parent(p.sx) != u) return NO_SLOT;
        if constexpr (MODE == M_INVAL_RESCUE) {
            // A child with a long row keeps its state for now: the rescue stage
            // after this round looks for another tight parent before its
            // subtree is given up.
            const VRec *r = &A.g.vr[x];
            if (r->blen + (uint32_t)r->dbl >= A.rescue_deg) {
                mark(A.cand_bm, xs);
                S.cand = true;
                return NO_SLOT;
            }
        }
        A.st[xs] = ST_INF;
        mark(A.inv_bm, xs);
        return xs;
    } else if constexpr (MODE == M_SEED) {
        // u was invalidated: candidate from a still-valid neighbour, reduced per
        // owner by the caller (every lane of a row targets the same slot).
        if (p.sx == ST_INF) return NO_SLOT;
        const uint64_t nd = (uint64_t)st_dist(p.sx) + w;
        if (nd >= DIST_INF) {
            atomicOr(&A.c->err, ERR_DIST_SUSPECT);  // the owner is unreachable so far
            return NO_SLOT;
        }
        pull = min(pull, pack_state((uint32_t)nd, x));
    } else if constexpr (MODE == M_TBFS) {
        if (p.sx == ST_INF || st_parent(p.sx) == x) return NO_SLOT;
        if ((uint64_t)aux0 + w != (uint64_t)st_dist(p.sx)) return NO_SLOT;
        const uint32_t nl = aux1 + 1;
        if (nl >= p.lx) return NO_SLOT;
        atomicMin(&A.lvl[xs], nl);
        return xs;
    } else if constexpr (MODE == M_OFCHECK) {
        if (p.sx == ST_INF) atomicOr(&A.c->err, ERR_DIST_OVERFLOW);
    } else {  // M_PICK: os is the slot whose parent is being chosen
        if (p.sx == ST_INF) return NO_SLOT;
        if ((uint64_t)st_dist(p.sx) + w != (uint64_t)aux0) return NO_SLOT;
        if (A.use_lvl && p.lx + 1 != aux1) return NO_SLOT;
        pull = min(pull, (uint64_t)x);
    }
    return NO_SLOT;
}

constexpr uint32_t LSTK = 256;  // slots a wave keeps to visit itself in the same launch

// A slot to visit next: kept on the wave's LDS stack when local continuation is
// on and there is room (it is then visited later in this launch, unmarked),
// else marked in the output bitmap for the next round.  `top` is wave-uniform;
// td is the target's new distance (RELAX).
template <int MODE>
__device__ __forceinline__ void visit_next(const FArgs &A, uint32_t tgt, uint32_t td, uint32_t *stk, uint32_t &top,
                                           WaveStats &S)
{
    const unsigned long long m = __ballot(tgt != NO_SLOT);
    if (!m) return;
    const uint32_t cnt = (uint32_t)__popcll(m);
    if (MODE != M_SWEEP && A.local && top + cnt <= LSTK) {  // (a sweep's marks all go to the bitmap)
        if (tgt != NO_SLOT) stk[top + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1ull))] = tgt;
        top += cnt;
    } else {
        if (tgt != NO_SLOT) {
            mark(A.out_bm, tgt);
            if constexpr (MODE == M_RELAX) {
                if (A.delta) note_mark(S, td);
            }
        }
        S.outm += cnt;
        S.any = true;
    }
}

// One wave's LDS: the fields of its tile's up to 64 slots, and its stack.
struct WaveSmem {
    uint32_t off[64];
    uint32_t u[64];
    uint32_t aux0[64];
    uint32_t aux1[64];
    uint32_t blen[64];
    uint32_t dbeg[64];
    uint64_t b0[64];
    unsigned long long pull[64];  // SEED / PICK: per-owner minimum candidate
    uint32_t stk[LSTK];           // push modes: slots to visit in this launch
};

constexpr bool is_pull(int mode) { return mode == M_SEED || mode == M_PICK || mode == M_SWEEP; }

// Owner side of a pull mode: the reduced candidate `best` of slot os.  Returns
// true when it lowered the slot (SEED), so that the slot is marked.
template <int MODE>
__device__ __forceinline__ bool pull_commit(const FArgs &A, uint32_t os, uint64_t best)
{
    if (best == ~0ull) return false;
    if constexpr (MODE == M_SEED) {
        if (best >= A.st[os]) return false;
        atomicMin((unsigned long long *)&A.st[os], (unsigned long long)best);
        return true;
    } else if constexpr (MODE == M_SWEEP) {
        const uint64_t cur = A.st[os];
        // a root (parent itself) never changes
        if (best >= cur || st_parent(cur) == (os >> A.kshift)) return false;
        atomicMin((unsigned long long *)&A.st[os], (unsigned long long)best);
        return true;
    } else {
        atomicMin(&A.best[os], (uint32_t)best);
        return false;
    }
}

// Owner-side fields of slot os; returns false when the owner has nothing to scan.
template <int MODE>
__device__ __forceinline__ bool owner_prep(const FArgs &A, uint32_t os, uint32_t &aux0, uint32_t &aux1)
{
    if constexpr (MODE == M_RELAX) {
        const uint64_t su = A.st[os];
        aux0 = st_dist(su);
        return su != ST_INF;
    } else if constexpr (MODE == M_TBFS) {
        const uint64_t su = A.st[os];
        aux0 = st_dist(su);
        aux1 = A.lvl[os];
        return su != ST_INF && aux1 != LVL_INF;
    } else if constexpr (MODE == M_OFCHECK) {
        return A.st[os] != ST_INF;
    } else if constexpr (MODE == M_SWEEP) {
        aux0 = st_dist(A.st[os]);  // an unreachable owner still pulls
        return true;
    } else if constexpr (MODE == M_PICK) {
        const uint64_t su = A.st[os];
        aux0 = st_dist(su);
        aux1 = A.use_lvl ? A.lvl[os] : 0u;
        return su != ST_INF && st_parent(su) != (os >> A.kshift);
    } else {
        return true;
    }
}

// A reservation [base, base + n) of the item list that does not fit: the part
// below the capacity is read by the next launch, so it gets empty items.
__device__ __forceinline__ void void_items(const FArgs &A, uint32_t base, uint32_t n)
{
    const uint32_t end = min(base + n, A.hcap);
    for (uint32_t i = base + lane_id(); i < end; i += 64) {
        A.out_hq[2 * (uint64_t)i] = 0ull;
        A.out_hq[2 * (uint64_t)i + 1] = 0ull;
    }
}

// Work of a launch: the set slots of its input bitmap and the items left by the
// previous launch.  Set slots are compacted across bitmap words into tiles of up
// to 64 slots, one per lane; a tile's rows are concatenated into one range of
// entries that the lanes share.  A slot whose row has more than A.heavy entries,
// or whose row would end past the tile's first A.budget entries, stays out of
// the tile: its row becomes items of A.chunk entries for the next launch, where
// any wave takes one.  So every wave's share of a launch is bounded.  An item is
// two words: a slot mask within one bitmap word, and word index << 32 | first
// entry.

// Position of the k-th set bit of w (k < popcount(w)).
__device__ __forceinline__ uint32_t select_bit(uint64_t w, uint32_t k)
{
    uint32_t pos = 0;
#pragma unroll
    for (uint32_t half = 32; half > 0; half >>= 1) {
        const uint32_t c = (uint32_t)__popcll(w & ((1ull << half) - 1ull));
        if (k >= c) {
            k -= c;
            w >>= half;
            pos += half;
        }
    }
    return pos;
}

// Entries [start, ...) of the rows of slots os (one per lane, where valid).  A
// fresh tile (from the input bitmap) hands its heavy rows to items and covers all
// its other entries; an item covers A.chunk entries of its range.
template <int MODE, int EPT>
__device__ void do_slots(const FArgs &A, uint32_t os, bool valid, uint32_t start, bool fresh, WaveSmem &sm,
                         WaveStats &S, uint32_t &top)
{
    const uint32_t lane = lane_id();
    uint32_t deg = 0, blen = 0, dbeg = 0, aux0 = 0, aux1 = 0;
    uint64_t b0 = 0;
    if (valid) {
        const VRec *r = &A.g.vr[os >> A.kshift];
        b0 = r->boff;
        blen = r->blen;
        const uint64_t dd = r->dbl;
        dbeg = (uint32_t)(dd >> 32);
        if constexpr (MODE == M_SWEEP) blen = r->blo;  // base entries below the owner's id
        deg = blen + (uint32_t)dd;
        if (!owner_prep<MODE>(A, os, aux0, aux1)) deg = 0;
    }
    bool later = false;  // at or above the round's threshold: left for a later round
    if constexpr (MODE == M_RELAX) {
        if (fresh && A.delta && deg && aux0 >= S.thr) {
            mark(A.out_bm, os);
            S.any = true;
            S.fenc = max(S.fenc, ~aux0);
            later = true;
            deg = 0;
        }
    }
    if (fresh) {
        S.front += (uint32_t)__popcll(__ballot(valid && !later));
        bool defer = deg > A.heavy;
        {  // rows ending past the budget, the tile's first row excepted
            uint32_t ib = defer ? 0u : deg;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t y = __shfl_up(ib, d);
                if ((int)lane >= d) ib += y;
            }
            if (!defer && deg && ib > A.budget && ib > deg) defer = true;
        }
        const uint32_t nch = defer ? (deg + A.chunk - 1) / A.chunk : 0u;
        uint32_t ic = nch;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(ic, d);
            if ((int)lane >= d) ic += y;
        }
        const uint32_t tot = __shfl(ic, 63);
        if (tot) {  // one list reservation per tile
            uint32_t base = 0;
            if (lane == 0) base = atomicAdd(A.out_hcnt, tot);
            base = __shfl(base, 0);
            if (base + tot > A.hcap) {
                // Push modes may visit a slot more than once per launch (its own
                // continuation): the list can fill; the slot is then visited in
                // the next round instead.  Pull modes visit each slot once.
                void_items(A, base, tot);
                if constexpr (MODE == M_RELAX || is_inval(MODE) || MODE == M_TBFS) {
                    if (defer) {
                        mark(A.out_bm, os);
                        if constexpr (MODE == M_RELAX) {
                            if (A.delta) note_mark(S, aux0);
                        }
                    }
                    S.any = true;
                } else {
                    if (lane == 0) atomicOr(&A.c->err, ERR_QUEUE);
                }
            } else {
                const uint64_t it0 = base + ic - nch;
                for (uint32_t c = 0; c < nch; ++c) {
                    A.out_hq[2 * (it0 + c)] = 1ull << (os & 63u);
                    A.out_hq[2 * (it0 + c) + 1] = ((uint64_t)(os >> 6) << 32) | (c * A.chunk);
                }
            }
        }
        if (defer) deg = 0;
    }
    uint32_t incl = deg;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d);
        if ((int)lane >= d) incl += y;
    }
    const uint32_t total = __shfl(incl, 63);
    const uint32_t end = fresh ? total : min(start + A.chunk, total);
    if (end <= start) return;
    sm.off[lane] = incl - deg;
    sm.u[lane] = os;
    sm.aux0[lane] = aux0;
    sm.aux1[lane] = aux1;
    sm.blen[lane] = blen;
    sm.dbeg[lane] = dbeg;
    sm.b0[lane] = b0;
    if constexpr (is_pull(MODE)) sm.pull[lane] = ~0ull;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // An item is a piece of one row (exactly one lane has entries): every entry
    // belongs to lane ol, so no owner search, and a pull mode keeps its
    // candidates in registers and reduces them over the wave once.
    const bool single = !fresh;
    const uint32_t ol = single ? (uint32_t)__builtin_ctzll(__ballot(deg != 0)) : 0u;
    uint64_t acc = ~0ull;
    for (uint32_t eb = start; eb < end; eb += 64 * EPT) {
        uint64_t ent[EPT];
        uint32_t ows[EPT], a0[EPT], a1[EPT], oi[EPT], pos[EPT];
        bool act[EPT];
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const uint32_t e = eb + j * 64 + lane;
            act[j] = e < end;
            ent[j] = 0;
            ows[j] = a0[j] = a1[j] = oi[j] = pos[j] = 0;
            if (act[j]) {
                uint32_t o = ol;
                if (!single) {
#pragma unroll
                    for (uint32_t st = 32; st > 0; st >>= 1)
                        if (sm.off[o + st] <= e) o += st;
                }
                const uint32_t jj = e - sm.off[o];
                const uint32_t bl = sm.blen[o];
                oi[j] = o;
                pos[j] = jj;
                ows[j] = sm.u[o];
                a0[j] = sm.aux0[o];
                a1[j] = sm.aux1[o];
                ent[j] = jj < bl ? A.g.adj[sm.b0[o] + jj] : A.g.pool[(uint64_t)sm.dbeg[o] + (jj - bl)];
            }
        }
        if constexpr (MODE == M_EMAP) {
#pragma unroll
            for (int j = 0; j < EPT; ++j)
                if (act[j] && !(ent[j] & E_DEAD) && emap_keeps(pos[j], sm.blen[oi[j]]))
                    emap_insert(A.g, ows[j], e_col(ent[j]), e_w(ent[j]), pos[j]);
            continue;
        }
        Pre pre[EPT];
#pragma unroll
        for (int j = 0; j < EPT; ++j) pre[j] = act[j] ? gather<MODE>(A, ows[j], ent[j]) : Pre{ST_INF, LVL_INF, true};
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            uint64_t pull = ~0ull;
            uint32_t td = 0;
            const uint32_t tgt = apply<MODE>(A, ows[j], a0[j], a1[j], ent[j], pre[j], S, pull, td);
            if constexpr (is_pull(MODE)) {
                if (single)
                    acc = min(acc, pull);
                else if (pull != ~0ull)
                    atomicMin(&sm.pull[oi[j]], (unsigned long long)pull);
            }
            if constexpr (MODE == M_RELAX || is_inval(MODE) || MODE == M_TBFS || MODE == M_SWEEP)
                visit_next<MODE>(A, tgt, td, sm.stk, top, S);
        }
    }
    S.edges += end - start;
    if constexpr (is_pull(MODE)) {
        if (single) {
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) acc = min(acc, (uint64_t)__shfl_xor((unsigned long long)acc, d));
            if (lane == ol) sm.pull[ol] = acc;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if constexpr (is_pull(MODE)) {
        // One commit per owner with entries in this range.
        const uint32_t lo = sm.off[lane];
        const bool mine = deg != 0 && lo < end && lo + deg > start;
        const uint64_t best = sm.pull[lane];
        if (mine && pull_commit<MODE>(A, os, best)) {
            mark(A.out_bm, os);
            S.any = true;
            if constexpr (MODE == M_SWEEP) S.marks++;
            if (A.delta) note_mark(S, st_dist(best));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Forests of 64 trees (K' = 64): bitmap word v is the tree mask of vertex v, so a
// word is processed with lane = tree.  Each row entry is read once (the same
// address in every lane) and the neighbour's 64 tree states are one contiguous
// 512-byte gather; the marks for a neighbour are one ballot and one atomicOr.
// A task covers FCH entries of the row; longer rows leave items for the rest.
constexpr uint32_t FCH = 64;              // row entries per forest task
constexpr int FE = 8;                     // row entries in flight per forest iteration
constexpr uint32_t FMIN = 24;             // trees of a vertex that make the tree-vector path pay
constexpr uint32_t FOREST_ITEM = 1u << 31;  // item tag: a piece of a tree-vector task

template <int MODE>
__device__ void do_forest(const FArgs &A, uint32_t v, uint64_t mask, uint32_t start, bool fresh, WaveStats &S)
{
    const uint32_t lane = lane_id();
    v = __builtin_amdgcn_readfirstlane(v);
    const uint32_t os = v * 64 + lane;
    const VRec *r = &A.g.vr[v];
    const uint64_t boff = r->boff;
    const uint32_t blen = r->blen;
    const uint64_t dd = r->dbl;
    const uint64_t dbeg = dd >> 32;
    const uint32_t deg = blen + (uint32_t)dd;
    bool act = (mask >> lane) & 1ull;
    uint32_t d = 0, l = 0;
    if (act) {
        if constexpr (MODE == M_RELAX || MODE == M_TBFS || MODE == M_PICK || MODE == M_OFCHECK) {
            const uint64_t su = A.st[os];
            d = st_dist(su);
            act = su != ST_INF;
            if constexpr (MODE == M_TBFS) {
                l = A.lvl[os];
                act = act && l != LVL_INF;
            }
            if constexpr (MODE == M_PICK) {
                act = act && st_parent(su) != v;
                l = A.use_lvl ? A.lvl[os] : 0u;
            }
        }
    }
    if constexpr (MODE == M_RELAX) {
        if (fresh && A.delta) {  // trees at or above the round's threshold: left for a later round
            const bool later = act && d >= S.thr;
            const unsigned long long lm = __ballot(later);
            if (lm) {
                if (lane == 0) atomicOr((unsigned long long *)&A.out_bm[v], lm);
                S.any = true;
                if (later) S.fenc = max(S.fenc, ~d);
                act = act && !later;
            }
        }
    }
    const unsigned long long amask = __ballot(act);
    if (fresh) S.front += (uint32_t)__popcll(MODE == M_RELAX ? amask : mask);
    if (!amask) return;
    uint32_t end = min(start + FCH, deg);
    if (fresh && deg > FCH) {  // the rest of the row as items for the next launch
        const uint32_t nit = (deg - 1) / FCH;
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(A.out_hcnt, nit);
        base = __shfl(base, 0);
        if (base + nit > A.hcap) {
            void_items(A, base, nit);
            if constexpr (MODE == M_RELAX || is_inval(MODE) || MODE == M_TBFS) {
                if (lane == 0) atomicOr((unsigned long long *)&A.out_bm[v], amask);
                S.any = true;
                if constexpr (MODE == M_RELAX) {
                    if (A.delta && act) note_mark(S, d);
                }
                end = start;  // the whole row is visited in the next round
            } else {
                if (lane == 0) atomicOr(&A.c->err, ERR_QUEUE);
            }
        } else {
            for (uint32_t i = lane; i < nit; i += 64) {
                A.out_hq[2 * (uint64_t)(base + i)] = amask;
                A.out_hq[2 * (uint64_t)(base + i) + 1] = ((uint64_t)v << 32) | FOREST_ITEM | ((i + 1) * FCH);
            }
        }
    }
    uint64_t best = ~0ull;  // pull modes: this lane's best candidate
    for (uint32_t eb = start; eb < end; eb += FE) {
        uint64_t ent[FE], sx[FE];
        uint32_t lx[FE];
#pragma unroll
        for (int j = 0; j < FE; ++j) {
            const uint32_t e = eb + j;
            ent[j] = e < end ? (e < blen ? A.g.adj[boff + e] : A.g.pool[dbeg + (e - blen)]) : ((uint64_t)v << 32);
        }
#pragma unroll
        for (int j = 0; j < FE; ++j) {
            const uint32_t x = e_col(ent[j]);
            const bool use = x != v && (MODE == M_INVAL || !(e_w(ent[j]) & W_DEAD));
            sx[j] = act && use ? A.st[x * 64 + lane] : ST_INF;
            lx[j] = LVL_INF;
            if constexpr (MODE == M_TBFS) lx[j] = act && use ? A.lvl[x * 64 + lane] : LVL_INF;
            if constexpr (MODE == M_PICK) {
                if (A.use_lvl) lx[j] = act && use ? A.lvl[x * 64 + lane] : LVL_INF;
            }
        }
#pragma unroll
        for (int j = 0; j < FE; ++j) {
            const uint32_t x = e_col(ent[j]);
            const uint32_t w = e_w(ent[j]);
            const bool use = act && x != v && (MODE == M_INVAL || !(w & W_DEAD));
            const uint32_t xs = x * 64 + lane;
            bool hit = false;
            if constexpr (MODE == M_RELAX) {
                if (use) {
                    const uint64_t nd = (uint64_t)d + w;
                    if (nd >= DIST_INF) {
                        if (sx[j] == ST_INF) atomicOr(&A.c->err, ERR_DIST_SUSPECT);
                    } else {
                        const uint64_t cand = pack_state((uint32_t)nd, v);
                        if (cand < sx[j] && st_parent(sx[j]) != x) {
                            atomicMin((unsigned long long *)&A.st[xs], (unsigned long long)cand);
                            hit = true;
                            S.marks++;
                            if (A.delta) note_mark(S, (uint32_t)nd);
                        }
                    }
                }
            } else if constexpr (MODE == M_INVAL) {
                if (use && sx[j] != ST_INF && st_parent(sx[j]) == v) {
                    A.st[xs] = ST_INF;
                    hit = true;
                }
            } else if constexpr (MODE == M_SEED) {
                if (use && sx[j] != ST_INF) {
                    const uint64_t nd = (uint64_t)st_dist(sx[j]) + w;
                    if (nd >= DIST_INF)
                        atomicOr(&A.c->err, ERR_DIST_SUSPECT);
                    else
                        best = min(best, pack_state((uint32_t)nd, x));
                }
            } else if constexpr (MODE == M_TBFS) {
                if (use && sx[j] != ST_INF && st_parent(sx[j]) != x && (uint64_t)d + w == (uint64_t)st_dist(sx[j]) &&
                    l + 1 < lx[j]) {
                    atomicMin(&A.lvl[xs], l + 1);
                    hit = true;
                }
            } else if constexpr (MODE == M_PICK) {
                if (use && sx[j] != ST_INF && (uint64_t)st_dist(sx[j]) + w == (uint64_t)d &&
                    (!A.use_lvl || lx[j] + 1 == l))
                    best = min(best, (uint64_t)x);
            } else if constexpr (MODE == M_OFCHECK) {
                if (use && sx[j] == ST_INF) atomicOr(&A.c->err, ERR_DIST_OVERFLOW);
            }
            if constexpr (MODE == M_RELAX || is_inval(MODE) || MODE == M_TBFS) {
                const unsigned long long m = __ballot(hit);
                if (m) {
                    if (lane == 0) {
                        atomicOr((unsigned long long *)&A.out_bm[x], m);
                        if constexpr (MODE == M_INVAL) atomicOr((unsigned long long *)&A.inv_bm[x], m);
                    }
                    S.any = true;
                }
            }
        }
    }
    S.edges += (uint64_t)(end - start) * (uint32_t)__popcll(amask);
    if constexpr (MODE == M_SEED) {
        const bool hit = act && best != ~0ull && best < A.st[os];
        if (hit) {
            atomicMin((unsigned long long *)&A.st[os], (unsigned long long)best);
            if (A.delta) note_mark(S, st_dist(best));
        }
        const unsigned long long m = __ballot(hit);
        if (m) {
            if (lane == 0) atomicOr((unsigned long long *)&A.out_bm[v], m);
            S.any = true;
        }
    } else if constexpr (MODE == M_PICK) {
        if (act && best != ~0ull) atomicMin(&A.best[os], (uint32_t)best);
    }
}

// Visits the slots on the wave's stack, 64 at a time, for at most A.local
// tiles; what is left is marked for the next round.
template <int MODE, int EPT>
__device__ void drain(const FArgs &A, WaveSmem &sm, WaveStats &S, uint32_t &top)
{
    if constexpr (MODE == M_RELAX || is_inval(MODE) || MODE == M_TBFS) {
        const uint32_t lane = lane_id();
        uint32_t budget = A.local;
        while (top) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            if (budget == 0) {
                for (uint32_t i = lane; i < top; i += 64) {
                    const uint32_t slot = sm.stk[i];
                    mark(A.out_bm, slot);
                    if constexpr (MODE == M_RELAX) {
                        if (A.delta) note_mark(S, st_dist(A.st[slot]));
                    }
                }
                S.any = true;
                S.outm += top;
                top = 0;
                break;
            }
            --budget;
            const uint32_t n = min(64u, top);
            const bool valid = lane < n;
            const uint32_t os = valid ? sm.stk[top - n + lane] : 0u;
            top -= n;
            do_slots<MODE, EPT>(A, os, valid, 0u, true, sm, S, top);
        }
    }
}

// Least waves per SIMD, which caps the register budget: 5 (up to 102
// registers; relaxation takes 96 and spills 4 values instead of 20) for the
// frontier modes, 6 (up to 80) for the full edge pass, whose grid is sized for
// 6 per SIMD.  Measured on MI355X: relaxation 0.327 -> 0.320 ms per update of
// the scale-24 headline at 5, the 64-tree forest 119.5 -> 115.8 ms per batch,
// the edge pass 3.64 -> 4.9 ms at 5 with the same grid.
constexpr int frontier_waves(int mode, bool) { return mode == M_SWEEP ? 6 : 5; }

template <int MODE, int EPT, bool FOREST, int FB>
__global__ __launch_bounds__(FB) __attribute__((amdgpu_waves_per_eu(frontier_waves(MODE, FOREST)))) void k_frontier(FArgs A)
{
    constexpr int FW = FB / 64;  // waves per workgroup
    __shared__ WaveSmem smem[FW];
    const uint32_t lane = lane_id();
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t gw = blockIdx.x * FW + wv;
    const uint32_t nw = gridDim.x * FW;
    // The bitmap is read in blocks of 8 words (64 bytes); in one pass a wave
    // reads 8 blocks nw apart, so that a cluster of non-empty words (frontiers
    // cluster by vertex id) spreads over many waves.  The set slots of the 64
    // words are then taken 64 at a time, in word order.
    const uint32_t nblk = A.tiles ? (A.nwords + 7) / 8 : 0u;
    // The first pass's words, the item count and the input flag are loaded
    // before this launch stores anything: a store in between (none of them to
    // these addresses) would otherwise hold the later loads back, one memory
    // latency each.
    uint32_t wi = 0;
    bool in0 = false;
    uint64_t raw0 = 0;
    if (nblk) {
        const uint32_t blk = (lane >> 3) * nw + gw;
        wi = blk * 8 + (lane & 7);
        in0 = blk < nblk && wi < A.nwords;
        if (in0 && !A.identity) raw0 = A.in_bm[wi];
    }
    // (without a count or flag, the launch's error word stands in for the
    // load, so that both loads issue unconditionally, together)
    const uint32_t hraw = *(A.in_hcnt ? A.in_hcnt : &A.c->err);
    const uint32_t act_in = *(A.in_act ? A.in_act : &A.c->err);
    if (blockIdx.x == 0 && threadIdx.x < 64) {
        if (threadIdx.x == 0) {
            if (A.zero_act) *A.zero_act = 0;
            if (A.zero_hcnt) *A.zero_hcnt = 0;
        }
        if (A.delta) {
            const uint32_t z = (A.ring + 2) % 3;
            A.c->fenc[z][lane] = 0;
            if (lane == 0) A.c->near[z] = 0;
        }
        if (MODE == M_RELAX && A.sweep) A.c->nmark[(A.ring + 2) % 3][lane] = 0;
    }
    // This round's threshold: the previous one, or, when the previous round
    // marked nothing below it, the smallest distance it marked plus delta.
    WaveStats S{0, 0, 0, false, 0u, false, 0u, 0u, 0u, false};
    if constexpr (MODE == M_RELAX) {
        if (A.delta) {
            uint32_t fe = A.c->fenc[A.ring][lane];
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) fe = max(fe, (uint32_t)__shfl_xor((int)fe, d));
            uint32_t thr = A.c->thr[A.ring];
            // (with nothing recorded at all, every slot is visited)
            if (!A.c->near[A.ring]) thr = fe ? (uint32_t)min((uint64_t)~fe + A.delta, (uint64_t)DIST_INF) : DIST_INF;
            S.thr = __builtin_amdgcn_readfirstlane(thr);
            if (blockIdx.x == 0 && threadIdx.x == 0) A.c->thr[(A.ring + 1) % 3] = S.thr;
        }
    }
    // Push-vs-full-sweep: after a round that marked at least A.sweep slots,
    // every slot is relaxed (rows in vertex order, tiles of 64 consecutive
    // slots) instead of the marked ones.  Every wave sums the same 64 words,
    // written by the previous launch, so the choice is the same grid-wide.
    bool sweep = false;
    if constexpr (MODE == M_RELAX && !FOREST) {
        if (A.sweep) {
            unsigned long long tot = A.c->nmark[A.ring][lane];
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) tot += __shfl_xor(tot, d);
            sweep = __builtin_amdgcn_readfirstlane((uint32_t)(tot >= A.sweep)) != 0;
            if (sweep && blockIdx.x == 0 && threadIdx.x == 0) add_stat(A.bstats, ST_SWEEPS, 1);
        }
    }
    WaveSmem &sm = smem[wv];
    auto load_word = [&](uint32_t pb, uint32_t &wi) -> uint64_t {
        const uint32_t blk = pb + (lane >> 3) * nw + gw;
        wi = blk * 8 + (lane & 7);
        uint64_t word = 0;
        if (blk < nblk && wi < A.nwords) {
            if (A.identity || sweep) {
                const uint32_t lo = wi * 64;
                word = A.nslots - lo >= 64 ? ~0ull : ((1ull << (A.nslots - lo)) - 1ull);
                if (sweep && A.in_bm[wi]) A.in_bm[wi] = 0;  // consumed all the same
            } else {
                word = A.in_bm[wi];
                if (word && A.consume) A.in_bm[wi] = 0;
            }
        }
        return word;
    };
    // The first pass's words (loaded above), as load_word takes them.
    uint64_t word = 0;
    if (in0) {
        if (A.identity || sweep) {
            const uint32_t lo = wi * 64;
            word = A.nslots - lo >= 64 ? ~0ull : ((1ull << (A.nslots - lo)) - 1ull);
            if (sweep && (A.identity ? A.in_bm[wi] : raw0)) A.in_bm[wi] = 0;  // consumed all the same
        } else {
            word = raw0;
            if (word && A.consume) A.in_bm[wi] = 0;
        }
    }
    // Items left by the previous launch; count clamped to the list capacity.
    const uint32_t H = A.in_hcnt ? min(hraw, A.hcap) : 0u;
    // A round after convergence has an empty bitmap and no items: it returns
    // without scanning further.
    if (H == 0 && A.in_act && act_in == 0) return;
    if (A.count_round && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&A.c->worked, 1u);
    uint32_t top = 0;  // slots on this wave's stack (wave-uniform)
    for (uint32_t pb = 0; pb < nblk; pb += nw * 8) {
        if (pb) word = load_word(pb, wi);
        if constexpr (FOREST && MODE != M_EMAP) {
            {  // forests of 64 trees: one vertex per word
                const bool wide = (uint32_t)__popcll(word) >= FMIN;
                unsigned long long nz = __ballot(wide);
                while (nz) {
                    const uint32_t l = (uint32_t)__builtin_ctzll(nz);
                    nz &= nz - 1;
                    do_forest<MODE>(A, (uint32_t)__shfl((int)wi, (int)l), __shfl(word, (int)l), 0u, true, S);
                }
                if (wide) word = 0;  // the other words go through slot compaction
            }
        }
        const uint32_t cnt = (uint32_t)__popcll(word);
        uint32_t incl = cnt;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl, d);
            if ((int)lane >= d) incl += y;
        }
        const uint32_t T = __shfl(incl, 63);
        for (uint32_t r0 = 0; r0 < T; r0 += 64) {
            const uint32_t q = r0 + lane;  // rank of this lane's slot
            uint32_t o = 0;                // owner lane: the number of lanes with incl <= q
#pragma unroll
            for (uint32_t st = 32; st > 0; st >>= 1) {
                const uint32_t v = (uint32_t)__shfl((int)incl, (int)(o + st - 1));
                if (v <= q) o += st;
            }
            o = min(o, 63u);
            const uint64_t ow = __shfl(word, (int)o);
            const uint32_t oex = (uint32_t)__shfl((int)(incl - cnt), (int)o);
            const uint32_t owi = (uint32_t)__shfl((int)wi, (int)o);
            const bool valid = q < T;
            const uint32_t os = valid ? owi * 64 + select_bit(ow, q - oex) : 0u;
            do_slots<MODE, EPT>(A, os, valid, 0u, true, sm, S, top);
        }
        drain<MODE, EPT>(A, sm, S, top);
    }
    for (uint32_t it = gw; it < H; it += nw) {
        const uint64_t mask = A.in_hq[2 * (uint64_t)it];
        const uint64_t ws = A.in_hq[2 * (uint64_t)it + 1];
        const uint32_t widx = (uint32_t)(ws >> 32);
        if constexpr (FOREST && MODE != M_EMAP) {
            if ((uint32_t)ws & FOREST_ITEM) {
                do_forest<MODE>(A, widx, mask, (uint32_t)ws & ~FOREST_ITEM, false, S);
                continue;
            }
        }
        do_slots<MODE, EPT>(A, widx * 64 + lane, (mask >> lane) & 1ull, (uint32_t)ws, false, sm, S, top);
    }
    drain<MODE, EPT>(A, sm, S, top);
    const uint32_t marks = wave_sum(S.marks);
    const uint32_t gath = MODE == M_SWEEP ? wave_sum(S.gath) : 0u;
    const bool any = __ballot(S.any) != 0;
    if constexpr (MODE == M_RELAX || MODE == M_SEED) {
        if (A.delta) {
            const bool nr = __ballot(S.near) != 0;
            uint32_t fe = S.fenc;
#pragma unroll
            for (int d = 32; d > 0; d >>= 1) fe = max(fe, (uint32_t)__shfl_xor((int)fe, d));
            if (lane == 0) {
                const uint32_t o = (A.ring + 1) % 3;
                if (nr) A.c->near[o] = 1u;
                if (fe) raise_u32(&A.c->fenc[o][blockIdx.x & 63u], fe);
            }
        }
    }
    if constexpr (MODE == M_INVAL_RESCUE) {
        if (__ballot(S.cand) != 0 && lane == 0) *A.cand_flag = 1u;
    }
    if (lane == 0) {
        if (any && A.out_act) *A.out_act = 1u;
        if constexpr (MODE == M_RELAX && !FOREST) {
            if (A.sweep && S.outm) atomicAdd(&A.c->nmark[(A.ring + 1) % 3][gw & 63u], S.outm);
        }
        if constexpr (MODE == M_RELAX) {
            add_stat(A.bstats, ST_RELAX_FRONT, S.front);
            add_stat(A.bstats, ST_RELAX_EDGES, S.edges);
            add_stat(A.bstats, ST_RELAX_MARKS, marks);
        } else if constexpr (MODE == M_SWEEP) {
            add_stat(A.bstats, ST_RELAX_FRONT, S.front);
            add_stat(A.bstats, ST_RELAX_EDGES, S.edges);
            add_stat(A.bstats, ST_RELAX_MARKS, marks);
            add_stat(A.bstats, ST_SWEEP_GATHERS, gath);
        } else if constexpr (MODE == M_EMAP) {
            // no statistics: the map is built outside updates
        } else {
            add_stat(A.bstats, ST_OTHER_EDGES, S.edges);
            if constexpr (MODE == M_SEED) add_stat(A.bstats, ST_INVALIDATED, S.front);
        }
    }
}

// Sums the per-workgroup statistics into the counters and clears the slots:
// REDUCE_WG workgroups share the slots, one atomicAdd per statistic each.
constexpr int REDUCE_WG = 8;
__global__ __launch_bounds__(256) void k_reduce_stats(Counters *c, unsigned long long *bstats, const GraphCtl *gc)
{
    if (gc && blockIdx.x == 0 && threadIdx.x == 0) {
        c->g_pool_top = gc->pool_top;
        c->g_err = gc->err;
    }
    __shared__ unsigned long long part[256][NSTAT];
    unsigned long long acc[NSTAT] = {};
    for (int i = blockIdx.x * 256 + threadIdx.x; i < STAT_SLOTS; i += REDUCE_WG * 256) {
#pragma unroll
        for (int k = 0; k < NSTAT; ++k) {
            acc[k] += bstats[i * NSTAT + k];
            bstats[i * NSTAT + k] = 0;
        }
    }
#pragma unroll
    for (int k = 0; k < NSTAT; ++k) part[threadIdx.x][k] = acc[k];
    __syncthreads();
    if (threadIdx.x < NSTAT) {
        unsigned long long s = 0;
        for (int i = 0; i < 256; ++i) s += part[i][threadIdx.x];
        if (s) atomicAdd(&c->stats[threadIdx.x], s);
    }
}

// ---------------------------------------------------------------------------
// Insertion of a batch into the resident CSR (delta blocks in a pool).  Most
// copies land in spare capacity of their row's block; only rows that overflow
// get a new block, carved deterministically: per-word sizes, one scan,
// per-word bases.
// ---------------------------------------------------------------------------

// Copy s of insertion record i: row a (s = 0) or row b (s = 1); a self-loop's
// two copies both go to row a.
__device__ __forceinline__ bool ins_copy(const int4 *rec, int64_t t, uint32_t &row, uint32_t &col, uint32_t &w)
{
    const int4 r = rec[t >> 1];
    if (r.w != 1) return false;
    const bool side = t & 1;
    row = (uint32_t)(side && r.x != r.y ? r.y : r.x);
    col = (uint32_t)(side ? r.x : r.y);
    w = (uint32_t)r.z;
    return true;
}

// Reserves a position with a returning atomicAdd on the row's delta length;
// with room in the block the copy is written (and entered in the edge map past
// the row's first EMAP_DSKIP delta entries), else
// it waits for the block to grow (the thread that reserved position == capacity
// marks the row).
__global__ __launch_bounds__(RB) void k_ins_place(const int4 *rec, int64_t k, GView g, uint32_t *tslot,
                                                  uint64_t *tbm)
{
    const int64_t t = (int64_t)blockIdx.x * RB + threadIdx.x;
    uint32_t row, col, w;
    if (t >= 2 * k || !ins_copy(rec, t, row, col, w)) return;
    VRec *r = &g.vr[row];
    const uint64_t o = atomicAdd((unsigned long long *)&r->dbl, 1ull);
    const uint32_t pos = (uint32_t)o, cap = r->dcap;
    if (pos < cap) {
        g.pool[(o >> 32) + pos] = pack_entry(col, w);
        if (pos >= EMAP_DSKIP) emap_insert(g, row, col, w, r->blen + pos);
        tslot[t] = pos;
    } else {
        tslot[t] = pos | SLOT_PENDING;
        if (pos == cap) mark(tbm, row);
    }
}

// Exclusive prefix sum over the 256 threads of a workgroup; *total gets the sum.
__device__ __forceinline__ uint64_t block_exclusive_scan(uint64_t v, uint64_t *total)
{
    __shared__ uint64_t wtot[RB / 64];
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    uint64_t incl = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(incl, d);
        if ((int)lane >= d) incl += y;
    }
    if (lane == 63) wtot[wv] = incl;
    __syncthreads();
    uint64_t before = 0, all = 0;
#pragma unroll
    for (int i = 0; i < RB / 64; ++i) {
        before += i < (int)wv ? wtot[i] : 0ull;
        all += wtot[i];
    }
    *total = all;
    return before + incl - v;
}

// Sum of v over the workgroup's RB threads, valid in thread 0 (so that a
// counter shared by the whole grid takes one atomic per workgroup).
__device__ __forceinline__ unsigned long long block_sum(unsigned long long v)
{
    __shared__ unsigned long long part[RB / 64];
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
    if (lane_id() == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    unsigned long long s = 0;
    if (threadIdx.x == 0)
        for (int i = 0; i < RB / 64; ++i) s += part[i];
    __syncthreads();
    return s;
}

// Capacity of a delta block that must hold `need` entries and holds `cap`:
// doubling, so that a row is moved O(log length) times over its lifetime.
__device__ __forceinline__ uint32_t grown_cap(uint32_t cap, uint32_t need)
{
    return need <= cap ? cap : max(max(2 * cap, need), 4u);
}

// wsum[w] = pool entries for the new blocks of the rows marked in bitmap word w;
// bsum[b] = the sum over the 256 words of workgroup b.
__global__ __launch_bounds__(RB) void k_ins_wordsum(const uint64_t *tbm, uint32_t nwords, const VRec *vr,
                                                    uint32_t *wsum, uint64_t *bsum)
{
    const uint32_t w = blockIdx.x * RB + threadIdx.x;
    uint32_t sum = 0;
    if (w < nwords) {
        uint64_t word = tbm[w];
        while (word) {
            const uint32_t v = w * 64 + (uint32_t)__builtin_ctzll(word);
            word &= word - 1;
            sum += grown_cap(vr[v].dcap, (uint32_t)vr[v].dbl);
        }
        wsum[w] = sum;
    }
    uint64_t total;
    (void)block_exclusive_scan(sum, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// For graphs of more than INS_FOLD_GROUPS workgroup sums: a single-workgroup
// exclusive scan of them into pool offsets bbase (from the current pool top);
// the total goes to gc->batch_new (the pool top advances in k_ins_pending).
__global__ __launch_bounds__(1024) void k_ins_scan(const uint64_t *bsum, uint32_t nb, uint64_t *bbase,
                                                   uint64_t pool_cap, GraphCtl *gc)
{
    __shared__ unsigned long long part[1024];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t lo = min(t * per, nb), hi = min(lo + per, nb);
    unsigned long long s = 0;
    for (uint32_t i = lo; i < hi; ++i) s += bsum[i];
    part[t] = s;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const unsigned long long y = t >= d ? part[t - d] : 0ull;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    const unsigned long long top = gc->pool_top;
    unsigned long long run = top + part[t] - s;
    for (uint32_t i = lo; i < hi; ++i) {
        bbase[i] = run;
        run += bsum[i];
    }
    if (t == 1023) {
        const unsigned long long total = part[1023];
        gc->batch_new = total;
        if (top + total > pool_cap) gc->err |= ERR_POOL;
    }
}

// Moves each marked row's block (its entries written so far) to a new block;
// the length already counts the pending copies.  Clears the marks.  Word w's
// new blocks start at its workgroup's base plus the workgroup-local prefix of
// wsum.  The base is bbase[blockIdx.x] when k_ins_scan ran; otherwise (at most
// INS_FOLD_GROUPS workgroups) every workgroup takes the pool top plus the sums
// of the workgroups before its own, and sums all nb of them: new blocks past
// the pool's capacity are an error (ERR_POOL, raised by workgroup 0), and then
// nothing moves.  The pool top itself advances in k_ins_pending, after every
// workgroup here has read it.
__global__ __launch_bounds__(RB) void k_ins_alloc(uint64_t *tbm, uint32_t nwords, const uint32_t *wsum,
                                                  const uint64_t *bsum, uint32_t nb, const uint64_t *bbase,
                                                  VRec *vr, uint64_t *pool, uint64_t pool_cap, GraphCtl *gc)
{
    __shared__ unsigned long long base_all[2];
    bool over;
    if (bbase) {
        if (threadIdx.x == 0) base_all[0] = bbase[blockIdx.x];
        __syncthreads();
        over = (gc->err & ERR_POOL) != 0;
    } else {
        unsigned long long before = 0, all = 0;
        for (uint32_t i = threadIdx.x; i < nb; i += RB) {
            const unsigned long long v = bsum[i];
            all += v;
            if (i < blockIdx.x) before += v;
        }
        before = block_sum(before);  // (valid in thread 0)
        all = block_sum(all);
        const unsigned long long top = gc->pool_top;
        if (threadIdx.x == 0) {
            base_all[0] = top + before;
            base_all[1] = all;
            if (blockIdx.x == 0) {
                gc->batch_new = all;
                if (top + all > pool_cap) gc->err |= ERR_POOL;
            }
        }
        __syncthreads();
        over = top + base_all[1] > pool_cap;
    }
    const uint32_t w = blockIdx.x * RB + threadIdx.x;
    const uint32_t mine = w < nwords ? wsum[w] : 0u;
    uint64_t total;
    uint64_t nb_ = base_all[0] + block_exclusive_scan(mine, &total);
    if (w >= nwords || !mine) return;
    uint64_t word = tbm[w];
    tbm[w] = 0;
    if (over) return;
    while (word) {
        const uint32_t v = w * 64 + (uint32_t)__builtin_ctzll(word);
        word &= word - 1;
        VRec *r = &vr[v];
        const uint64_t d = r->dbl;
        const uint32_t len = (uint32_t)d, cap = r->dcap;
        const uint64_t beg = d >> 32;
        for (uint32_t j = 0; j < cap; ++j) pool[nb_ + j] = pool[beg + j];
        r->dbl = (nb_ << 32) | len;
        r->dcap = grown_cap(cap, len);
        nb_ += r->dcap;
    }
}

// Writes the copies that waited for their row's new block; thread 0 advances
// the pool top past the new blocks (k_ins_alloc has read it).
__global__ __launch_bounds__(RB) void k_ins_pending(const int4 *rec, int64_t k, GView g, uint32_t *tslot,
                                                    GraphCtl *gc)
{
    const int64_t t = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (t == 0 && !(gc->err & ERR_POOL)) gc->pool_top += gc->batch_new;
    uint32_t row, col, w;
    if (t >= 2 * k || !(tslot[t] & SLOT_PENDING) || !ins_copy(rec, t, row, col, w)) return;
    if (gc->err & ERR_POOL) return;  // reported to the host; nothing was placed
    const uint32_t pos = tslot[t] & ~SLOT_PENDING;
    const VRec *r = &g.vr[row];
    g.pool[(r->dbl >> 32) + pos] = pack_entry(col, w);
    if (pos >= EMAP_DSKIP) emap_insert(g, row, col, w, r->blen + pos);
    tslot[t] = pos;
}

// ---------------------------------------------------------------------------
// Deletions: tree-edge test + tombstones (binary search in the sorted base row,
// scan of the short delta row).
// ---------------------------------------------------------------------------

// Deletion records: thread 2i+s tombstones the copies of record i in row a
// (s = 0) or row b (s = 1); a self-loop's two copies both sit in row a.
__global__ __launch_bounds__(RB) void k_tomb(const int4 *rec, int64_t k, GView g, unsigned long long *bstats)
{
    const int64_t t = (int64_t)blockIdx.x * RB + threadIdx.x;
    const int64_t ri = t >> 1;
    uint32_t probes = 0, kills = 0;
    if (ri < k) {
        const int4 r = rec[ri];
        const bool side = t & 1;
        if (r.w == 0 && !(side && r.x == r.y))
            probes = emap_kill(g, (uint32_t)(side ? r.y : r.x), (uint32_t)(side ? r.x : r.y), (uint32_t)r.z, kills);
    }
    const uint32_t p = wave_sum(probes);
    const uint32_t kl = wave_sum(kills);
    if (lane_id() == 0) {
        add_stat(bstats, ST_PROBES, p);
        add_stat(bstats, ST_KILLS, kl);
    }
}

// deleteEdge's tree-edge test (gpuFunctions_undir.cuh:189-198), weight ignored:
// the child endpoint of a deleted tree edge roots an invalidated subtree.
// Thread i handles record i >> kshift for tree i & mask.
__global__ __launch_bounds__(RB) void k_classify(const int4 *rec, int64_t k, uint32_t kshift, uint64_t *st,
                                                 uint64_t *out_bm, uint64_t *inv_bm, uint32_t *out_act,
                                                 const VRec *vr, uint32_t rescue_deg, uint64_t *cand_bm,
                                                 uint32_t *cand_flag)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    const int64_t ri = i >> kshift;
    const uint32_t tree = (uint32_t)i & ((1u << kshift) - 1u);
    bool hit = false, cand = false;
    if (ri < k) {
        const int4 r = rec[ri];
        const uint32_t a = (uint32_t)r.x, b = (uint32_t)r.y;
        if (r.w == 0 && a != b) {
            const uint32_t sa = (a << kshift) | tree, sb = (b << kshift) | tree;
            const uint64_t vb = st[sb], va = st[sa];  // both at once: most deleted edges are not tree edges
            uint32_t victim = 0xFFFFFFFFu;
            if (vb != ST_INF && st_parent(vb) == a)
                victim = sb;
            else if (va != ST_INF && st_parent(va) == b)
                victim = sa;
            if (victim != 0xFFFFFFFFu) {
                const VRec *r = &vr[victim >> kshift];
                if (rescue_deg && r->blen + (uint32_t)r->dbl >= rescue_deg) {
                    mark(cand_bm, victim);  // kept for the rescue stage
                    cand = true;
                } else {
                    st[victim] = ST_INF;  // only ST_INF is written by this kernel
                    mark(out_bm, victim);
                    mark(inv_bm, victim);
                    hit = true;
                }
            }
        }
    }
    if (__ballot(hit) != 0 && lane_id() == 0) *out_act = 1u;
    if (__ballot(cand) != 0 && lane_id() == 0) *cand_flag = 1u;
}

// Rescue commit: every candidate slot of the bitmap either keeps its distance
// under the parent the pick pass found (the smallest-id live neighbour y with
// d[y] + w = d[slot] that was not reset) or is reset like any invalidated slot.
// One thread per bitmap word.  A kept slot is exact with its new parent: that
// parent's distance is strictly smaller (positive weights), so parents never
// form a cycle, and if the parent is reset later its row scan finds the slot
// again as its child.
__global__ __launch_bounds__(RB) void k_rescue_commit(uint64_t *cand_bm, uint32_t nwords, uint64_t *st, uint32_t *best,
                                                      uint64_t *inv_bm, uint64_t *out_bm, uint32_t *out_act,
                                                      const uint32_t *cand_flag, uint32_t *clear_flag,
                                                      uint32_t *aux_hcnt, unsigned long long *bstats)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *clear_flag = 0u;
        *aux_hcnt = 0u;
    }
    if (*cand_flag == 0) return;
    const uint64_t wi = (uint64_t)blockIdx.x * RB + threadIdx.x;
    uint64_t word = 0;
    if (wi < nwords) {
        word = cand_bm[wi];
        if (word) cand_bm[wi] = 0;
    }
    uint64_t reset = 0;
    uint32_t kept = 0;
    for (uint64_t m = word; m; m &= m - 1) {
        const uint32_t slot = (uint32_t)(wi * 64 + (uint64_t)__builtin_ctzll(m));
        const uint32_t p = best[slot];
        if (p != 0xFFFFFFFFu) {
            st[slot] = pack_state(st_dist(st[slot]), p);
            best[slot] = 0xFFFFFFFFu;
            ++kept;
        } else {
            st[slot] = ST_INF;
            reset |= 1ull << (slot & 63u);
        }
    }
    if (reset) {
        atomicOr((unsigned long long *)&inv_bm[wi], (unsigned long long)reset);
        atomicOr((unsigned long long *)&out_bm[wi], (unsigned long long)reset);
    }
    if (__ballot(reset != 0) != 0 && lane_id() == 0) *out_act = 1u;
    const uint32_t k = wave_sum(kept);
    if (lane_id() == 0) add_stat(bstats, ST_RESCUED, k);
}

// Offers every inserted edge that survived the batch's deletions in both
// directions (the role of insertEdge, gpuFunctions_undir.cuh:217-251).
__global__ __launch_bounds__(RB) void k_ins_offer(const int4 *rec, int64_t k, uint32_t kshift, const uint32_t *tslot,
                                                  GView g, uint64_t *st, uint64_t *out_bm, uint32_t *out_act,
                                                  Counters *c, uint32_t far)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    const int64_t ri = i >> kshift;
    const uint32_t tree = (uint32_t)i & ((1u << kshift) - 1u);
    bool hit = false;
    uint32_t fe = 0;  // ~(smallest distance marked)
    const int4 r = ri < k ? rec[ri] : make_int4(0, 0, 0, 0);
    if (r.w == 1 && r.x != r.y) {
        const uint32_t a = (uint32_t)r.x, b = (uint32_t)r.y;
        const uint32_t sa_slot = (a << kshift) | tree, sb_slot = (b << kshift) | tree;
        const uint32_t w = (uint32_t)r.z;
        const uint64_t sa = st[sa_slot], sb = st[sb_slot];
        const uint32_t ta = tslot[2 * ri];
        const bool a_root = sa != ST_INF && st_parent(sa) == a;  // roots never change
        const bool b_root = sb != ST_INF && st_parent(sb) == b;
        // Whether the edge would lower an endpoint (or overflow): only then is
        // it checked against the batch's deletions (the copy in row a is
        // tombstoned when one removed it), which costs two more dependent
        // random reads that most insertions, lowering nothing, skip.
        auto acts = [&](uint64_t sf, uint64_t sto, uint32_t from, bool to_root) {
            if (sf == ST_INF || to_root) return false;
            const uint64_t nd = (uint64_t)st_dist(sf) + w;
            return nd >= DIST_INF ? sto == ST_INF : pack_state((uint32_t)nd, from) < sto;
        };
        const bool want = acts(sa, sb, a, b_root) || acts(sb, sa, b, a_root);
        if (want && !(ta & SLOT_PENDING) && !(g.pool[(g.vr[a].dbl >> 32) + ta] & E_DEAD)) {
            if (sa != ST_INF && !b_root) {
                const uint64_t nd = (uint64_t)st_dist(sa) + w;
                if (nd >= DIST_INF) {
                    if (sb == ST_INF) atomicOr(&c->err, ERR_DIST_SUSPECT);
                } else {
                    const uint64_t cand = pack_state((uint32_t)nd, a);
                    if (cand < sb) {
                        atomicMin((unsigned long long *)&st[sb_slot], (unsigned long long)cand);
                        mark(out_bm, sb_slot);
                        hit = true;
                        fe = max(fe, ~(uint32_t)nd);
                    }
                }
            }
            if (sb != ST_INF && !a_root) {
                const uint64_t nd = (uint64_t)st_dist(sb) + w;
                if (nd >= DIST_INF) {
                    if (sa == ST_INF) atomicOr(&c->err, ERR_DIST_SUSPECT);
                } else {
                    const uint64_t cand = pack_state((uint32_t)nd, b);
                    if (cand < sa) {
                        atomicMin((unsigned long long *)&st[sa_slot], (unsigned long long)cand);
                        mark(out_bm, sa_slot);
                        hit = true;
                        fe = max(fe, ~(uint32_t)nd);
                    }
                }
            }
        }
    }
    if (__ballot(hit) != 0 && lane_id() == 0) *out_act = 1u;
    if (far) {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) fe = max(fe, (uint32_t)__shfl_xor((int)fe, d));
        if (fe && lane_id() == 0) raise_u32(&c->fenc[0][blockIdx.x & 63u], fe);
    }
}

// ---------------------------------------------------------------------------
// Merge check of a batch made of several requests applied as one.  One batch
// applies its insertions, then removes every copy of each deleted (a, b, w),
// so it leaves the graph that the requests applied one by one leave unless a
// request inserts a key that an earlier request deletes.  The deletions enter
// a hash table of key -> earliest request index (atomicMin); each insertion
// looks its key up; *first = the smallest index of a request that inserts a
// key an earlier request deletes (~0: none).  Keys are 64-bit hashes of
// (min(a,b), max(a,b), w): a collision can only report a conflict that is not
// there, which splits the merge and costs time, never exactness.
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint64_t change_key(const int4 &r)
{
    const uint64_t lo = (uint32_t)min(r.x, r.y), hi = (uint32_t)max(r.x, r.y);
    uint64_t h = ((lo << 32) | hi) * 0x9E3779B97F4A7C15ull;
    h ^= ((uint64_t)(uint32_t)r.z + 0x51ull) * 0xC2B2AE3D27D4EB4Full;
    h ^= h >> 31;
    h *= 0xBF58476D1CE4E5B9ull;
    h ^= h >> 29;
    return h ? h : 1ull;
}

// Request index of record i: the last request whose first record is <= i.
__device__ __forceinline__ uint32_t request_of(const uint64_t *first_rec, uint32_t nreq, uint64_t i)
{
    uint32_t lo = 0, hi = nreq;  // first_rec[lo] <= i < first_rec[hi]
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) / 2;
        if (first_rec[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

__global__ __launch_bounds__(RB) void k_conflict_del(const int4 *rec, int64_t k, const uint64_t *first_rec,
                                                     uint32_t nreq, unsigned long long *keys, uint32_t *req,
                                                     uint64_t mask)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= k) return;
    const int4 r = rec[i];
    if (r.w != 0) return;
    const unsigned long long h = change_key(r);
    const uint32_t q = request_of(first_rec, nreq, (uint64_t)i);
    for (uint64_t sl = h & mask;; sl = (sl + 1) & mask) {
        const unsigned long long prev = atomicCAS(&keys[sl], 0ull, h);
        if (prev == 0ull || prev == h) {
            atomicMin(&req[sl], q);
            return;
        }
    }
}

__global__ __launch_bounds__(RB) void k_conflict_ins(const int4 *rec, int64_t k, const uint64_t *first_rec,
                                                     uint32_t nreq, const unsigned long long *keys,
                                                     const uint32_t *req, uint64_t mask, uint32_t *first)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= k) return;
    const int4 r = rec[i];
    if (r.w != 1) return;
    const unsigned long long h = change_key(r);
    for (uint64_t sl = h & mask;; sl = (sl + 1) & mask) {
        const unsigned long long kk = keys[sl];
        if (kk == 0ull) return;
        if (kk == h) {
            const uint32_t q = request_of(first_rec, nreq, (uint64_t)i);
            if (req[sl] < q) atomicMin(first, q);
            return;
        }
    }
}

// ---------------------------------------------------------------------------
// Repack of the resident CSR (tombstone reclamation and pool compaction), in
// three steps, one wave per row in the first and last:
//   rows   a row's live base entries move to the front of its base segment in
//          order (so it stays sorted), its base length and the count of them
//          below the vertex's own id are rewritten, and the capacity of its new
//          delta block (at least min_cap, at least its live delta entries) is
//          recorded; live totals go to gc->live
//   scan   new pool offsets: exclusive prefix sum of the capacities
//   delta  the row's live delta entries are copied, in order, to its new block
//          in the new pool
// The edge map is rebuilt afterwards (entry positions changed).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(RB) void k_repack_rows(VRec *vr, uint64_t *adj, const uint64_t *pool, uint64_t n,
                                                    uint32_t min_cap, uint32_t *cap, GraphCtl *gc)
{
    const uint32_t lane = lane_id();
    const uint64_t gw = ((uint64_t)blockIdx.x * RB + threadIdx.x) >> 6;
    const uint64_t nw = ((uint64_t)gridDim.x * RB) >> 6;
    unsigned long long lb = 0, ld = 0;  // wave-uniform totals
    for (uint64_t v = gw; v < n; v += nw) {
        const uint64_t boff = vr[v].boff;
        const uint32_t blen = vr[v].blen;
        const uint64_t dd = vr[v].dbl;
        uint32_t w = 0, below = 0;
        for (uint32_t i0 = 0; i0 < blen; i0 += 64) {
            const uint32_t i = i0 + lane;
            uint64_t e = 0;
            bool live = false;
            if (i < blen) {
                e = adj[boff + i];
                live = !(e & E_DEAD);
            }
            // every lane's entry is loaded before any is stored (the ballot
            // needs them all), and stores go at or below the chunk read
            const unsigned long long m = __ballot(live);
            if (live) {
                adj[boff + w + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = e;
                below += e_col(e) < (uint32_t)v;
            }
            w += (uint32_t)__popcll(m);
        }
        below = wave_sum(below);
        const uint64_t dbeg = dd >> 32;
        const uint32_t dl = (uint32_t)dd;
        uint32_t live_d = 0;
        for (uint32_t i0 = 0; i0 < dl; i0 += 64) {
            const uint32_t i = i0 + lane;
            live_d += (uint32_t)__popcll(__ballot(i < dl && !(pool[dbeg + i] & E_DEAD)));
        }
        if (lane == 0) {
            vr[v].blen = w;
            vr[v].blo = below;
            cap[v] = max(min_cap, live_d);
        }
        lb += w;
        ld += live_d;
    }
    // lb, ld are wave-uniform: count them once per wave
    lb = block_sum(lane == 0 ? lb : 0ull);
    ld = block_sum(lane == 0 ? ld : 0ull);
    if (threadIdx.x == 0 && (lb || ld)) {
        atomicAdd(&gc->live[0], lb);
        atomicAdd(&gc->live[1], ld);
    }
}

// Scan, step 1: bsum[b] = sum of cap over the RB rows of workgroup b.
__global__ __launch_bounds__(RB) void k_scan_blocks(const uint32_t *cap, uint64_t n, uint64_t *bsum)
{
    const uint64_t v = (uint64_t)blockIdx.x * RB + threadIdx.x;
    uint64_t total;
    (void)block_exclusive_scan(v < n ? cap[v] : 0u, &total);
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Scan, step 2 (one workgroup): bbase = exclusive prefix sum of bsum[0, nb);
// *total = the sum.
__global__ __launch_bounds__(1024) void k_scan_top(const uint64_t *bsum, uint64_t nb, uint64_t *bbase,
                                                   unsigned long long *total)
{
    __shared__ unsigned long long part[1024];
    const uint32_t t = threadIdx.x;
    const uint64_t per = (nb + 1023) / 1024;
    const uint64_t lo = min((uint64_t)t * per, nb), hi = min(lo + per, nb);
    unsigned long long sum = 0;
    for (uint64_t i = lo; i < hi; ++i) sum += bsum[i];
    part[t] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const unsigned long long y = t >= d ? part[t - d] : 0ull;
        __syncthreads();
        part[t] += y;
        __syncthreads();
    }
    unsigned long long run = part[t] - sum;
    for (uint64_t i = lo; i < hi; ++i) {
        bbase[i] = run;
        run += bsum[i];
    }
    if (t == 1023) *total = part[1023];
}

// Scan, step 3 and the move: row v's new block starts at bbase[block] plus the
// in-block prefix of cap; its live delta entries are copied there in order.
__global__ __launch_bounds__(RB) void k_repack_delta(VRec *vr, const uint64_t *pool, uint64_t *npool, uint64_t n,
                                                     const uint32_t *cap, const uint64_t *bbase)
{
    __shared__ uint64_t off[RB];
    __shared__ uint32_t cp[RB];
    const uint64_t v0 = (uint64_t)blockIdx.x * RB;
    const uint64_t v = v0 + threadIdx.x;
    const uint32_t c = v < n ? cap[v] : 0u;
    uint64_t total;
    off[threadIdx.x] = bbase[blockIdx.x] + block_exclusive_scan(c, &total);
    cp[threadIdx.x] = c;
    __syncthreads();
    // each wave copies the rows of the workgroup in turn, one row at a time
    const uint32_t lane = lane_id(), wv = threadIdx.x >> 6;
    for (uint32_t j = wv; j < RB && v0 + j < n; j += RB / 64) {
        const uint64_t row = v0 + j;
        const uint64_t dd = vr[row].dbl;
        const uint64_t dbeg = dd >> 32;
        const uint32_t dl = (uint32_t)dd;
        uint32_t w = 0;
        for (uint32_t i0 = 0; i0 < dl; i0 += 64) {
            const uint32_t i = i0 + lane;
            uint64_t e = 0;
            bool live = false;
            if (i < dl) {
                e = pool[dbeg + i];
                live = !(e & E_DEAD);
            }
            const unsigned long long m = __ballot(live);
            if (live) npool[off[j] + w + (uint32_t)__popcll(m & ((1ull << lane) - 1ull))] = e;
            w += (uint32_t)__popcll(m);
        }
        if (lane == 0) {
            vr[row].dbl = (off[j] << 32) | w;
            vr[row].dcap = cp[j];
        }
    }
}

// ---------------------------------------------------------------------------
// Small helpers.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(RB) void k_set_sources(uint64_t *st, const uint32_t *src, uint32_t count,
                                                    uint32_t kshift, uint32_t per_tree, uint64_t *out_bm,
                                                    uint32_t *out_act)
{
    const uint32_t j = blockIdx.x * RB + threadIdx.x;
    if (j >= count) return;
    const uint32_t v = src[j];
    const uint32_t slot = (v << kshift) | (per_tree ? j : 0u);
    st[slot] = pack_state(0u, v);
    mark(out_bm, slot);
    *out_act = 1u;
}

__global__ __launch_bounds__(RB) void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) p[i] = v;
}

__global__ __launch_bounds__(RB) void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) p[i] = v;
}

// Zeroes p[0, n) and q[0, m) in one launch.
__global__ __launch_bounds__(RB) void k_zero2_u32(uint32_t *p, uint64_t n, uint32_t *q, uint64_t m)
{
    for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n + m; i += (uint64_t)gridDim.x * RB) {
        if (i < n)
            p[i] = 0u;
        else
            q[i - n] = 0u;
    }
}

// Device-side check of k change records (the host check of dsssp_update,
// for records already in device memory): cnt[0] insertions, cnt[1] deletions,
// cnt[2] zero-weight insertions, cnt[3] the smallest index of a record with
// op 0/1 and a vertex outside [0, n) or a negative weight (~0: none).  Records
// with another op are ignored.
__global__ __launch_bounds__(RB) void k_validate(const int4 *rec, int64_t k, int64_t n, unsigned long long *cnt)
{
    unsigned long long ins = 0, del = 0, zw = 0;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < k; i += (int64_t)gridDim.x * RB) {
        const int4 r = rec[i];
        if (r.w == 0 || r.w == 1) {
            if (r.x < 0 || r.x >= n || r.y < 0 || r.y >= n || r.z < 0)
                atomicMin(&cnt[3], (unsigned long long)i);
            else if (r.w == 1)
                ins += 1, zw += r.z == 0;
            else
                del += 1;
        }
    }
    ins = block_sum(ins);
    del = block_sum(del);
    zw = block_sum(zw);
    if (threadIdx.x == 0) {
        if (ins) atomicAdd(&cnt[0], ins);
        if (del) atomicAdd(&cnt[1], del);
        if (zw) atomicAdd(&cnt[2], zw);
    }
}

// Probe kernels of known footprint (counter calibration).  k_probe_copy: out =
// in over n words.  k_probe_gather: count 8-byte reads at hashed positions of
// in[0, n), summed per lane and written once per lane (so no read is elided).
__global__ __launch_bounds__(RB) void k_probe_copy(const uint64_t *in, uint64_t *out, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) out[i] = in[i];
}

__global__ __launch_bounds__(RB) void k_probe_gather(const uint64_t *in, uint64_t n, uint64_t count, uint64_t *out)
{
    const uint64_t t = (uint64_t)blockIdx.x * RB + threadIdx.x, nt = (uint64_t)gridDim.x * RB;
    uint64_t acc = 0;
    for (uint64_t i = t; i < count; i += nt) {
        uint64_t h = (i + 1) * 0x9E3779B97F4A7C15ull;
        h ^= h >> 31;
        h *= 0xBF58476D1CE4E5B9ull;
        h ^= h >> 29;
        acc += in[h % n];
    }
    out[t] = acc;
}

// dst |= src word by word, src cleared, *count = the bits src held (one
// atomicAdd per wave).
__global__ __launch_bounds__(RB) void k_merge_count(uint64_t *dst, uint64_t *src, uint64_t nwords,
                                                    unsigned long long *count)
{
    uint32_t c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < nwords; i += (uint64_t)gridDim.x * RB) {
        const uint64_t v = src[i];
        if (v) {
            dst[i] |= v;
            src[i] = 0;
            c += (uint32_t)__popcll(v);
        }
    }
    const unsigned long long tot = block_sum(c);
    if (threadIdx.x == 0 && tot) atomicAdd(count, tot);
}

// out[v] = st[v << kshift | tree]: one tree's states, contiguous.
__global__ __launch_bounds__(RB) void k_gather_tree(const uint64_t *st, uint64_t n, uint32_t kshift, uint32_t tree,
                                                    uint64_t *out)
{
    for (uint64_t v = (uint64_t)blockIdx.x * RB + threadIdx.x; v < n; v += (uint64_t)gridDim.x * RB)
        out[v] = st[(v << kshift) | tree];
}

// Order-independent digest of p[0, n): the sum of a 64-bit mix of (index,
// value), one atomicAdd per wave.
__global__ __launch_bounds__(RB) void k_digest(const uint64_t *p, uint64_t n, unsigned long long *out)
{
    uint64_t acc = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) {
        uint64_t h = (p[i] ^ (i * 0x9E3779B97F4A7C15ull)) * 0xBF58476D1CE4E5B9ull;
        acc += h ^ (h >> 31);
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) acc += (uint64_t)__shfl_xor((long long)acc, d);
    if (lane_id() == 0) atomicAdd(out, (unsigned long long)acc);
}

// Roots (parent == self) start the tight-graph BFS at level 0.
__global__ __launch_bounds__(RB) void k_lvl_init(uint32_t nslots, uint32_t kshift, const uint64_t *st, uint32_t *lvl,
                                                 uint32_t *best, uint64_t *out_bm, uint32_t *out_act)
{
    const uint32_t s = blockIdx.x * RB + threadIdx.x;
    bool root = false;
    if (s < nslots) {
        const uint64_t v = st[s];
        root = v != ST_INF && st_parent(v) == (s >> kshift);
        lvl[s] = root ? 0u : LVL_INF;
        best[s] = 0xFFFFFFFFu;
        if (root) mark(out_bm, s);
    }
    if (__ballot(root) != 0 && lane_id() == 0) *out_act = 1u;
}

__global__ __launch_bounds__(RB) void k_apply_best(uint32_t nslots, uint32_t kshift, uint64_t *st, const uint32_t *best,
                                                   Counters *c)
{
    const uint32_t s = blockIdx.x * RB + threadIdx.x;
    if (s >= nslots) return;
    const uint64_t v = st[s];
    if (v == ST_INF || st_parent(v) == (s >> kshift)) return;
    const uint32_t b = best[s];
    if (b == 0xFFFFFFFFu) {
        atomicOr(&c->err, ERR_NO_TIGHT);
        return;
    }
    st[s] = pack_state(st_dist(v), b);
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + RB - 1) / RB); }

template <int MODE, int EPT, bool FOREST, int FB>
void launch_block(const FArgs &a, int grid, hipStream_t s, hipEvent_t start, hipEvent_t stop)
{
    if (start || stop)
        hipExtLaunchKernelGGL((k_frontier<MODE, EPT, FOREST, FB>), dim3(grid), dim3(FB), 0, s, start, stop, 0, a);
    else
        hipLaunchKernelGGL((k_frontier<MODE, EPT, FOREST, FB>), dim3(grid), dim3(FB), 0, s, a);
}

// Workgroup size of the launch: a.block threads (128, 256 or 512).
template <int MODE, int EPT, bool FOREST>
void launch_kernel(const FArgs &a, int grid, hipStream_t s, hipEvent_t start, hipEvent_t stop)
{
    switch (a.block) {
    case 128: launch_block<MODE, EPT, FOREST, 128>(a, grid, s, start, stop); break;
    case 512: launch_block<MODE, EPT, FOREST, 512>(a, grid, s, start, stop); break;
    default: launch_block<MODE, EPT, FOREST, 256>(a, grid, s, start, stop); break;
    }
}

// Slot layouts with 64 trees per vertex use the variant with the tree-vector
// path (its registers would lower the occupancy of every other launch).
template <int MODE, int EPT>
void launch_mode(const FArgs &a, int grid, hipStream_t s, hipEvent_t start, hipEvent_t stop)
{
    if (MODE != M_EMAP && a.kshift == 6)
        launch_kernel<MODE, 4, MODE != M_EMAP>(a, grid, s, start, stop);
    else
        launch_kernel<MODE, EPT, false>(a, grid, s, start, stop);
}

template <int EPT>
void launch_frontier_ept(int mode, const FArgs &a, int grid, hipStream_t s, hipEvent_t start, hipEvent_t stop)
{
    switch (mode) {
    case M_RELAX: launch_mode<M_RELAX, EPT>(a, grid, s, start, stop); break;
    case M_INVAL: launch_mode<M_INVAL, EPT>(a, grid, s, start, stop); break;
    case M_SEED: launch_mode<M_SEED, EPT>(a, grid, s, start, stop); break;
    case M_TBFS: launch_mode<M_TBFS, EPT>(a, grid, s, start, stop); break;
    case M_PICK: launch_mode<M_PICK, EPT>(a, grid, s, start, stop); break;
    case M_OFCHECK: launch_mode<M_OFCHECK, EPT>(a, grid, s, start, stop); break;
    case M_SWEEP: launch_kernel<M_SWEEP, EPT, false>(a, grid, s, start, stop); break;  // never 64 trees
    case M_INVAL_RESCUE: launch_kernel<M_INVAL_RESCUE, EPT, false>(a, grid, s, start, stop); break;  // nor this
    default: launch_mode<M_EMAP, EPT>(a, grid, s, start, stop); break;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------

void launch_frontier(int mode, const FArgs &a, int grid, hipStream_t s, hipEvent_t start, hipEvent_t stop)
{
    switch (a.ept) {
    case 1: launch_frontier_ept<1>(mode, a, grid, s, start, stop); break;
    case 2: launch_frontier_ept<2>(mode, a, grid, s, start, stop); break;
    default: launch_frontier_ept<4>(mode, a, grid, s, start, stop); break;
    }
}

void launch_fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s)
{
    if (n == 0) return;
    const unsigned g = (unsigned)std::min<uint64_t>((n + RB - 1) / RB, 8192);
    hipLaunchKernelGGL(k_fill_u64, dim3(g), dim3(RB), 0, s, p, v, n);
}

void launch_fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s)
{
    if (n == 0) return;
    const unsigned g = (unsigned)std::min<uint64_t>((n + RB - 1) / RB, 8192);
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(RB), 0, s, p, v, n);
}

void launch_zero2_u32(uint32_t *p, uint64_t n, uint32_t *q, uint64_t m, hipStream_t s)
{
    if (n + m == 0) return;
    const unsigned g = (unsigned)std::min<uint64_t>((n + m + RB - 1) / RB, 8192);
    hipLaunchKernelGGL(k_zero2_u32, dim3(g), dim3(RB), 0, s, p, n, q, m);
}

void launch_digest(const uint64_t *p, uint64_t n, unsigned long long *out, hipStream_t s)
{
    launch_fill_u64((uint64_t *)out, 0ull, 1, s);
    if (n == 0) return;
    const unsigned g = (unsigned)std::min<uint64_t>((n + RB - 1) / RB, 4096);
    hipLaunchKernelGGL(k_digest, dim3(g), dim3(RB), 0, s, p, n, out);
}

void launch_validate(const int4 *rec, int64_t k, int64_t n, unsigned long long *cnt, hipStream_t s)
{
    launch_fill_u64((uint64_t *)cnt, 0ull, 3, s);
    launch_fill_u64((uint64_t *)cnt + 3, ~0ull, 1, s);
    if (k > 0) hipLaunchKernelGGL(k_validate, dim3(std::min(blocks_for(k), 1024u)), dim3(RB), 0, s, rec, k, n, cnt);
}

void launch_merge_check(const int4 *rec, int64_t k, const uint64_t *first_rec, uint32_t nreq,
                        unsigned long long *keys, uint32_t *req, uint64_t slots, uint32_t *first, hipStream_t s)
{
    launch_fill_u64((uint64_t *)keys, 0ull, slots, s);
    launch_fill_u32(req, 0xFFFFFFFFu, slots, s);
    launch_fill_u32(first, 0xFFFFFFFFu, 1, s);
    if (k <= 0) return;
    hipLaunchKernelGGL(k_conflict_del, dim3(blocks_for(k)), dim3(RB), 0, s, rec, k, first_rec, nreq, keys, req,
                       slots - 1);
    hipLaunchKernelGGL(k_conflict_ins, dim3(blocks_for(k)), dim3(RB), 0, s, rec, k, first_rec, nreq, keys, req,
                       slots - 1, first);
}

void launch_repack_rows(VRec *vr, uint64_t *adj, const uint64_t *pool, uint64_t n, uint32_t min_cap, uint32_t *cap,
                        GraphCtl *gc, hipStream_t s)
{
    if (n == 0) return;
    const unsigned g = (unsigned)std::min<uint64_t>((n * 64 + RB - 1) / RB, 4096);
    hipLaunchKernelGGL(k_repack_rows, dim3(g), dim3(RB), 0, s, vr, adj, pool, n, min_cap, cap, gc);
}

void launch_repack_scan(const uint32_t *cap, uint64_t n, uint64_t *bsum, uint64_t *bbase, unsigned long long *total,
                        hipStream_t s)
{
    const uint64_t nb = (n + RB - 1) / RB;
    if (nb == 0) return;
    hipLaunchKernelGGL(k_scan_blocks, dim3((unsigned)nb), dim3(RB), 0, s, cap, n, bsum);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(1024), 0, s, bsum, nb, bbase, total);
}

void launch_repack_delta(VRec *vr, const uint64_t *pool, uint64_t *npool, uint64_t n, const uint32_t *cap,
                         const uint64_t *bbase, hipStream_t s)
{
    const uint64_t nb = (n + RB - 1) / RB;
    if (nb) hipLaunchKernelGGL(k_repack_delta, dim3((unsigned)nb), dim3(RB), 0, s, vr, pool, npool, n, cap, bbase);
}

void launch_probe(int kind, const uint64_t *in, uint64_t n, uint64_t count, uint64_t *out, hipStream_t s)
{
    if (kind == 0)
        hipLaunchKernelGGL(k_probe_copy, dim3(8192), dim3(RB), 0, s, in, out, n);
    else
        hipLaunchKernelGGL(k_probe_gather, dim3(8192), dim3(RB), 0, s, in, n, count, out);
}

void launch_merge_count(uint64_t *dst, uint64_t *src, uint64_t nwords, unsigned long long *count, hipStream_t s)
{
    launch_fill_u64((uint64_t *)count, 0ull, 1, s);
    if (nwords == 0) return;
    const unsigned g = (unsigned)std::min<uint64_t>((nwords + RB - 1) / RB, 1024);
    hipLaunchKernelGGL(k_merge_count, dim3(g), dim3(RB), 0, s, dst, src, nwords, count);
}

void launch_gather_tree(const uint64_t *st, uint64_t n, uint32_t kshift, uint32_t tree, uint64_t *out,
                        hipStream_t s)
{
    if (n == 0) return;
    const unsigned g = (unsigned)std::min<uint64_t>((n + RB - 1) / RB, 8192);
    hipLaunchKernelGGL(k_gather_tree, dim3(g), dim3(RB), 0, s, st, n, kshift, tree, out);
}

void launch_reduce_stats(Counters *c, unsigned long long *bstats, hipStream_t s, const GraphCtl *gc)
{
    hipLaunchKernelGGL(k_reduce_stats, dim3(REDUCE_WG), dim3(256), 0, s, c, bstats, gc);
}

void launch_ins_place(const int4 *rec, int64_t k, GView g, uint32_t *tslot, uint64_t *tbm, hipStream_t s)
{
    if (k > 0) hipLaunchKernelGGL(k_ins_place, dim3(blocks_for(2 * k)), dim3(RB), 0, s, rec, k, g, tslot, tbm);
}

void launch_ins_wordsum(const uint64_t *tbm, uint32_t nwords, const VRec *vr, uint32_t *wsum, uint64_t *bsum,
                        hipStream_t s)
{
    if (nwords > 0)
        hipLaunchKernelGGL(k_ins_wordsum, dim3(blocks_for(nwords)), dim3(RB), 0, s, tbm, nwords, vr, wsum, bsum);
}


void launch_ins_alloc(uint64_t *tbm, uint32_t nwords, const uint32_t *wsum, const uint64_t *bsum, uint64_t *bbase,
                      VRec *vr, uint64_t *pool, uint64_t pool_cap, GraphCtl *gc, hipStream_t s)
{
    if (nwords == 0) return;
    const uint32_t nb = blocks_for(nwords);
    const bool scan = nb > INS_FOLD_GROUPS;
    if (scan) hipLaunchKernelGGL(k_ins_scan, dim3(1), dim3(1024), 0, s, bsum, nb, bbase, pool_cap, gc);
    hipLaunchKernelGGL(k_ins_alloc, dim3(nb), dim3(RB), 0, s, tbm, nwords, wsum, bsum, nb,
                       scan ? (const uint64_t *)bbase : nullptr, vr, pool, pool_cap, gc);
}

void launch_ins_pending(const int4 *rec, int64_t k, GView g, uint32_t *tslot, GraphCtl *gc, hipStream_t s)
{
    if (k > 0) hipLaunchKernelGGL(k_ins_pending, dim3(blocks_for(2 * k)), dim3(RB), 0, s, rec, k, g, tslot, gc);
}

void launch_classify(const int4 *rec, int64_t k, uint32_t kshift, uint64_t *st, uint64_t *out_bm, uint64_t *inv_bm,
                     uint32_t *out_act, const VRec *vr, uint32_t rescue_deg, uint64_t *cand_bm, uint32_t *cand_flag,
                     hipStream_t s)
{
    if (k > 0)
        hipLaunchKernelGGL(k_classify, dim3(blocks_for(k << kshift)), dim3(RB), 0, s, rec, k, kshift, st, out_bm,
                           inv_bm, out_act, vr, rescue_deg, cand_bm, cand_flag);
}

void launch_rescue_commit(uint64_t *cand_bm, uint32_t nwords, uint64_t *st, uint32_t *best, uint64_t *inv_bm,
                          uint64_t *out_bm, uint32_t *out_act, const uint32_t *cand_flag, uint32_t *clear_flag,
                          uint32_t *aux_hcnt, unsigned long long *bstats, hipStream_t s)
{
    hipLaunchKernelGGL(k_rescue_commit, dim3(blocks_for((int64_t)nwords + 1)), dim3(RB), 0, s, cand_bm,
                       nwords, st, best, inv_bm, out_bm, out_act, cand_flag, clear_flag, aux_hcnt, bstats);
}

void launch_tomb(const int4 *rec, int64_t k, GView g, unsigned long long *bstats, hipStream_t s)
{
    if (k > 0) hipLaunchKernelGGL(k_tomb, dim3(blocks_for(2 * k)), dim3(RB), 0, s, rec, k, g, bstats);
}

void launch_ins_offer(const int4 *rec, int64_t k, uint32_t kshift, const uint32_t *tslot, GView g, uint64_t *st,
                      uint64_t *out_bm, uint32_t *out_act, Counters *c, bool far, hipStream_t s)
{
    if (k > 0)
        hipLaunchKernelGGL(k_ins_offer, dim3(blocks_for(k << kshift)), dim3(RB), 0, s, rec, k, kshift, tslot, g, st,
                           out_bm, out_act, c, far ? 1u : 0u);
}

void launch_set_sources(uint64_t *st, const uint32_t *src, uint32_t count, uint32_t kshift, bool per_tree,
                        uint64_t *out_bm, uint32_t *out_act, hipStream_t s)
{
    if (count > 0)
        hipLaunchKernelGGL(k_set_sources, dim3(blocks_for(count)), dim3(RB), 0, s, st, src, count, kshift,
                           per_tree ? 1u : 0u, out_bm, out_act);
}

void launch_lvl_init(uint32_t nslots, uint32_t kshift, const uint64_t *st, uint32_t *lvl, uint32_t *best,
                     uint64_t *out_bm, uint32_t *out_act, hipStream_t s)
{
    if (nslots > 0)
        hipLaunchKernelGGL(k_lvl_init, dim3(blocks_for(nslots)), dim3(RB), 0, s, nslots, kshift, st, lvl, best, out_bm,
                           out_act);
}

void launch_apply_best(uint32_t nslots, uint32_t kshift, uint64_t *st, const uint32_t *best, Counters *c,
                       hipStream_t s)
{
    if (nslots > 0)
        hipLaunchKernelGGL(k_apply_best, dim3(blocks_for(nslots)), dim3(RB), 0, s, nslots, kshift, st, best, c);
}

}  // namespace dsssp
