// kernels.hip -- hand-written gfx950 kernels of the dynamic SSSP update.
//
// Work mapping (CDNA4, wave64):
//  * Frontier kernels (k_frontier<MODE>) run a fixed grid of 256-thread
//    workgroups; every wave takes work items in a grid-stride loop.  An item is
//    either a "tile" of 64 light frontier vertices or one chunk of a heavy row.
//    A tile is edge-balanced: each lane loads one vertex's row bounds, the wave
//    forms an exclusive prefix sum of the degrees with DPP/shuffle steps, stages
//    the per-vertex fields in LDS, and then lane l of the wave handles edge
//    (base + l) of the tile's concatenated rows, finding its owner with a
//    6-step binary search in LDS.  Adjacent lanes therefore read adjacent
//    8-byte entries of the same row (coalesced) regardless of the degree mix.
//  * Rows longer than `heavy` entries are not scanned inside a tile; they are
//    split into `chunk`-entry pieces appended to the next round's heavy queue
//    and processed one chunk per wave, so a hub never serialises a wave.
//  * Vertices to visit next are compacted with one wave ballot + popcount and
//    a single atomicAdd per wave; an epoch stamp per vertex removes duplicates.
//  * Distance and parent live in one uint64 (dist << 32 | parent) and change
//    together through one 64-bit atomicMin, so the smaller distance wins and,
//    among equal distances, the smaller parent id.
// The reference does one thread per frontier vertex with unsynchronised
// read-modify-writes (gpuFunctions_undir.cuh:370-433) and one full O(n) filter
// pass per round (:125-174); none of that structure is reused here.
#include <hip/hip_runtime.h>

#include "common.h"
#include "kernels.h"

namespace dsssp {

namespace {

constexpr int FB = FRONTIER_BLOCK;
constexpr int FW = FB / 64;  // waves per workgroup
constexpr int RB = 256;      // block size of per-record / per-vertex kernels

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ uint32_t load_relaxed(const uint32_t *p)
{
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ unsigned long long lanemask_lt() { return (1ull << lane_id()) - 1ull; }

__device__ __forceinline__ uint32_t wave_sum(uint32_t v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v += __shfl_xor(v, d);
    return v;
}

// Appends v for every lane with want set: one atomicAdd per wave.
__device__ __forceinline__ void wave_append(bool want, uint32_t v, uint32_t *q, uint32_t *cnt, uint32_t cap,
                                            Counters *c)
{
    const unsigned long long mask = __ballot(want);
    if (mask == 0) return;
    const uint32_t leader = (uint32_t)__builtin_ctzll(mask);
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(cnt, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    if (want) {
        const uint32_t pos = base + (uint32_t)__popcll(mask & lanemask_lt());
        if (pos < cap)
            q[pos] = v;
        else
            atomicOr(&c->err, ERR_QUEUE);
    }
}

struct WaveStats {
    uint32_t front;     // wave-uniform
    uint64_t edges;     // wave-uniform
    uint32_t improved;  // per lane
    uint32_t enq;       // per lane
};

// One edge of the owner slot os (vertex os >> kshift, tree os & mask) with
// adjacency entry ent.  Returns true when slot outv should be enqueued.
template <int MODE>
__device__ __forceinline__ bool edge_op(const FArgs &A, uint32_t os, uint32_t aux0, uint32_t aux1, uint64_t ent,
                                        uint32_t &outv, WaveStats &S)
{
    const uint32_t u = os >> A.kshift;
    const uint32_t x = e_col(ent);
    const uint32_t w = e_w(ent);
    if (x == u) return false;  // a self-loop never carries a shortest path
    if constexpr (MODE != M_INVAL) {
        if (w & W_DEAD) return false;
    }
    const uint32_t xs = (x << A.kshift) | (os & ((1u << A.kshift) - 1u));
    if constexpr (MODE == M_RELAX) {
        const uint64_t nd = (uint64_t)aux0 + w;
        if (nd >= DIST_INF) {
            atomicOr(&A.c->err, ERR_DIST_OVERFLOW);
            return false;
        }
        const uint64_t cand = pack_state((uint32_t)nd, u);
        const uint64_t cur = A.st[xs];  // possibly stale, only ever too large
        // A root (parent itself, distance 0) never changes: with zero weights a
        // candidate (0, smaller id) would otherwise take its place.
        if (cand >= cur || st_parent(cur) == x) return false;
        const uint64_t old = atomicMin((unsigned long long *)&A.st[xs], (unsigned long long)cand);
        if (cand >= old) return false;
        S.improved++;
        outv = xs;
        return true;
    } else if constexpr (MODE == M_INVAL) {
        // x is a child of u in the pre-update tree: reset it and descend.
        // Tombstoned entries are scanned too, so a child hanging off a deleted
        // parallel copy is still found.
        const uint64_t sx = A.st[xs];
        if (sx == ST_INF || st_parent(sx) != u) return false;
        if (atomicCAS((unsigned long long *)&A.st[xs], (unsigned long long)sx, (unsigned long long)ST_INF) != sx)
            return false;
        outv = xs;
        return true;
    } else if constexpr (MODE == M_SEED) {
        // u was invalidated: take the best still-valid neighbour.
        const uint64_t sx = A.st[xs];
        if (sx == ST_INF) return false;
        const uint64_t nd = (uint64_t)st_dist(sx) + w;
        if (nd >= DIST_INF) {
            atomicOr(&A.c->err, ERR_DIST_OVERFLOW);
            return false;
        }
        const uint64_t cand = pack_state((uint32_t)nd, x);
        if (cand >= A.st[os]) return false;
        const uint64_t old = atomicMin((unsigned long long *)&A.st[os], (unsigned long long)cand);
        if (cand >= old) return false;
        outv = os;
        return true;
    } else if constexpr (MODE == M_TBFS) {
        const uint64_t sx = A.st[xs];
        if (sx == ST_INF || st_parent(sx) == x) return false;
        if ((uint64_t)aux0 + w != (uint64_t)st_dist(sx)) return false;
        const uint32_t nl = aux1 + 1;
        if (nl >= A.lvl[xs]) return false;
        if (nl >= atomicMin(&A.lvl[xs], nl)) return false;
        outv = xs;
        return true;
    } else {  // M_PICK: os is the slot whose parent is being chosen
        const uint64_t sx = A.st[xs];
        if (sx == ST_INF) return false;
        if ((uint64_t)st_dist(sx) + w != (uint64_t)aux0) return false;
        if (A.use_lvl && A.lvl[xs] + 1 != aux1) return false;
        if (x < A.best[os]) atomicMin(&A.best[os], x);
        return false;
    }
}

template <int MODE>
__device__ __forceinline__ void finish_edge(const FArgs &A, bool want, uint32_t v, WaveStats &S)
{
    if constexpr (MODE == M_PICK) {
        return;
    } else {
        if constexpr (MODE != M_INVAL) {
            if (want) {
                if (A.stamp[v] == A.epoch || atomicExch(&A.stamp[v], A.epoch) == A.epoch) want = false;
            }
        }
        wave_append(want, v, A.out_q, A.out_cnt, A.out_cap, A.c);
        if constexpr (MODE == M_INVAL) wave_append(want, v, A.invl, A.inv_total, A.out_cap, A.c);
        S.enq += want ? 1u : 0u;
    }
}

// Every mode splits long rows into chunks: loop modes process them in the next
// round, single-pass modes (seed, parent pick) in a second, chunk-only launch.
constexpr bool defers(int) { return true; }

struct TileSmem {
    uint32_t off[FW][64];
    uint32_t u[FW][64];
    uint32_t aux0[FW][64];
    uint32_t aux1[FW][64];
    uint32_t blen[FW][64];
    uint32_t dbeg[FW][64];
    uint64_t b0[FW][64];
};

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
    } else if constexpr (MODE == M_PICK) {
        const uint64_t su = A.st[os];
        aux0 = st_dist(su);
        aux1 = A.use_lvl ? A.lvl[os] : 0u;
        return su != ST_INF && st_parent(su) != (os >> A.kshift);
    } else {
        return true;
    }
}

template <int MODE>
__device__ void do_tile(const FArgs &A, uint32_t t, uint32_t tv_shift, uint32_t L, TileSmem &sm, uint32_t wv,
                        WaveStats &S)
{
    const uint32_t lane = lane_id();
    const uint32_t idx = (t << tv_shift) + lane;
    const bool valid = lane < (1u << tv_shift) && idx < L;
    uint32_t os = 0, deg = 0, blen = 0, dbeg = 0, aux0 = 0, aux1 = 0;
    uint64_t b0 = 0;
    if (valid) {
        os = A.in_q ? A.in_q[idx] : idx;
        const uint32_t u = os >> A.kshift;
        b0 = A.g.rp[u];
        blen = (uint32_t)(A.g.rp[u + 1] - b0);
        const uint64_t dd = A.g.dbl[u];
        dbeg = (uint32_t)(dd >> 32);
        deg = blen + (uint32_t)dd;
        if (!owner_prep<MODE>(A, os, aux0, aux1)) deg = 0;
        if constexpr (defers(MODE)) {
            if (deg > A.heavy) {
                const uint32_t nch = (deg + A.chunk - 1) / A.chunk;
                const uint32_t base = atomicAdd(A.out_hcnt, nch);
                if (base + nch <= A.out_hcap) {
                    for (uint32_t c = 0; c < nch; ++c) A.out_hq[base + c] = ((uint64_t)os << 32) | c;
                } else {
                    atomicOr(&A.c->err, ERR_QUEUE);
                }
                deg = 0;
            }
        }
    }
    S.front += (uint32_t)__popcll(__ballot(valid));
    uint32_t incl = deg;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(incl, d);
        if ((int)lane >= d) incl += y;
    }
    const uint32_t total = __shfl(incl, 63);
    if (total == 0) return;
    sm.off[wv][lane] = incl - deg;
    sm.u[wv][lane] = os;
    sm.aux0[wv][lane] = aux0;
    sm.aux1[wv][lane] = aux1;
    sm.blen[wv][lane] = blen;
    sm.dbeg[wv][lane] = dbeg;
    sm.b0[wv][lane] = b0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t eb = 0; eb < total; eb += 64) {
        const uint32_t e = eb + lane;
        bool want = false;
        uint32_t outv = 0;
        if (e < total) {
            uint32_t o = 0;
#pragma unroll
            for (uint32_t s = 32; s > 0; s >>= 1)
                if (sm.off[wv][o + s] <= e) o += s;
            const uint32_t j = e - sm.off[wv][o];
            const uint32_t bl = sm.blen[wv][o];
            const uint64_t ent = j < bl ? A.g.adj[sm.b0[wv][o] + j] : A.g.pool[(uint64_t)sm.dbeg[wv][o] + (j - bl)];
            want = edge_op<MODE>(A, sm.u[wv][o], sm.aux0[wv][o], sm.aux1[wv][o], ent, outv, S);
        }
        finish_edge<MODE>(A, want, outv, S);
    }
    S.edges += total;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int MODE>
__device__ void do_chunk(const FArgs &A, uint32_t item, WaveStats &S)
{
    const uint32_t lane = lane_id();
    const uint64_t hv = A.in_hq[item];
    const uint32_t os = (uint32_t)(hv >> 32);
    const uint32_t u = os >> A.kshift;
    const uint32_t c = (uint32_t)hv;
    const uint64_t b0 = A.g.rp[u];
    const uint32_t blen = (uint32_t)(A.g.rp[u + 1] - b0);
    const uint64_t dd = A.g.dbl[u];
    const uint32_t dbeg = (uint32_t)(dd >> 32);
    const uint32_t deg = blen + (uint32_t)dd;
    uint32_t aux0 = 0, aux1 = 0;
    if (c == 0) S.front += 1;
    if (!owner_prep<MODE>(A, os, aux0, aux1)) return;
    const uint32_t start = c * A.chunk;
    const uint32_t end = min(start + A.chunk, deg);
    for (uint32_t jb = start; jb < end; jb += 64) {
        const uint32_t j = jb + lane;
        bool want = false;
        uint32_t outv = 0;
        if (j < end) {
            const uint64_t ent = j < blen ? A.g.adj[b0 + j] : A.g.pool[(uint64_t)dbeg + (j - blen)];
            want = edge_op<MODE>(A, os, aux0, aux1, ent, outv, S);
        }
        finish_edge<MODE>(A, want, outv, S);
    }
    if (end > start) S.edges += end - start;
}

template <int MODE>
__global__ __launch_bounds__(FB) void k_frontier(FArgs A)
{
    __shared__ TileSmem sm;
    if (A.dev_iter) {
        Counters *c = A.c;
        const uint32_t it = load_relaxed(&c->iter);
        const uint32_t cs = (load_relaxed(&c->ring_base) + it) % 3u;
        const uint32_t qs = (load_relaxed(&c->q_base) + it) & 1u;
        A.epoch = load_relaxed(&c->epoch_base) + it;
        A.in_q = A.qv[qs];
        A.in_cnt = &c->ring[cs][0];
        A.in_hq = A.hqv[qs];
        A.in_hcnt = &c->ring[cs][1];
        A.out_q = A.qv[qs ^ 1u];
        A.out_cnt = &c->ring[(cs + 1u) % 3u][0];
        A.out_hq = A.hqv[qs ^ 1u];
        A.out_hcnt = &c->ring[(cs + 1u) % 3u][1];
        A.zero_cnt = &c->ring[(cs + 2u) % 3u][0];
    }
    if (A.zero_cnt && blockIdx.x == 0 && threadIdx.x == 0) {
        A.zero_cnt[0] = 0;
        A.zero_cnt[1] = 0;
    }
    // Input buffers have the same capacities as the output ones; a count above
    // capacity (only after a reported overflow) is clamped so no read leaves them.
    const uint32_t L = min(A.in_cnt ? load_relaxed(A.in_cnt) : A.in_n, A.out_cap);
    const uint32_t H = min(A.in_hcnt ? load_relaxed(A.in_hcnt) : 0u, A.out_hcap);
    const uint32_t wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const uint32_t nw = gridDim.x * FW;
    // Vertices per tile: 64 when the frontier fills every wave several times,
    // fewer (down to 1) for short frontiers, so that the rows of a short
    // frontier are spread over all waves instead of walked by a few.
    uint32_t tv_shift = 6;
    while (tv_shift > 0 && (L >> tv_shift) < 2u * nw) --tv_shift;
    const uint32_t T = (L + (1u << tv_shift) - 1u) >> tv_shift;
    const uint32_t items = H + T;
    WaveStats S{0, 0, 0, 0};
    for (uint32_t it = blockIdx.x * FW + wv; it < items; it += nw) {
        if (it < H)
            do_chunk<MODE>(A, it, S);
        else
            do_tile<MODE>(A, it - H, tv_shift, L, sm, wv, S);
    }
    const uint32_t imp = wave_sum(S.improved);
    const uint32_t enq = wave_sum(S.enq);
    if (lane_id() == 0 && (S.edges | S.front | imp | enq)) {
        if constexpr (MODE == M_RELAX) {
            atomicAdd(&A.c->relax_frontier, (unsigned long long)S.front);
            atomicAdd(&A.c->relax_edges, (unsigned long long)S.edges);
            atomicAdd(&A.c->relax_improved, (unsigned long long)imp);
            atomicAdd(&A.c->relax_enq, (unsigned long long)enq);
        } else {
            atomicAdd(&A.c->other_edges, (unsigned long long)S.edges);
        }
    }
    if (A.dev_iter) {
        // The last workgroup to finish advances the round index; the next
        // launch on the stream starts after this one has completed.
        __syncthreads();
        if (threadIdx.x == 0) {
            __threadfence();
            const uint32_t prev = atomicAdd(&A.c->blocks_done, 1u);
            if (prev == gridDim.x - 1u) {
                atomicExch(&A.c->blocks_done, 0u);
                atomicAdd(&A.c->iter, 1u);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Insertion of a batch into the resident CSR (delta rows in a pool).
// ---------------------------------------------------------------------------

// Counts new entries per endpoint; the first touch of a vertex lists it.
__device__ __forceinline__ void wave_append_plain(bool want, uint32_t v, uint32_t *q, uint32_t *cnt)
{
    const unsigned long long mask = __ballot(want);
    if (mask == 0) return;
    const uint32_t leader = (uint32_t)__builtin_ctzll(mask);
    uint32_t base = 0;
    if (lane_id() == leader) base = atomicAdd(cnt, (uint32_t)__popcll(mask));
    base = __shfl(base, (int)leader);
    if (want) q[base + (uint32_t)__popcll(mask & lanemask_lt())] = v;
}

__global__ __launch_bounds__(RB) void k_ins_count(const int4 *rec, int64_t k, uint32_t *tcnt, uint64_t *slot,
                                                  uint32_t *touched, GraphCtl *gc)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    bool ta = false, tb = false;
    int4 r = make_int4(0, 0, 0, -1);
    if (i < k) {
        r = rec[i];
        if (r.w == 1) {
            const uint32_t ra = atomicAdd(&tcnt[r.x], 1u);
            const uint32_t rb = atomicAdd(&tcnt[r.y], 1u);
            ta = ra == 0;
            tb = rb == 0;
            slot[i] = ((uint64_t)ra << 32) | rb;
        }
    }
    // the touched list holds each vertex once (first touch); at most 2k entries
    wave_append_plain(ta, (uint32_t)r.x, touched, &gc->touched);
    wave_append_plain(tb, (uint32_t)r.y, touched, &gc->touched);
}

// Moves each touched vertex's delta row to a fresh block with room for its new
// entries; leaves the old length in tcnt as the write base.  Blocks are carved
// from the pool with one atomicAdd per wave (prefix sum of the lanes' sizes).
__global__ __launch_bounds__(RB) void k_ins_alloc(const uint32_t *touched, uint32_t max_touched, uint64_t *dbl,
                                                  uint32_t *tcnt, uint64_t *pool, uint64_t pool_cap, GraphCtl *gc)
{
    const uint32_t i = blockIdx.x * RB + threadIdx.x;
    const uint32_t nt = min(load_relaxed(&gc->touched), max_touched);
    const uint32_t lane = lane_id();
    uint32_t v = 0, len = 0, add = 0;
    uint64_t beg = 0;
    if (i < nt) {
        v = touched[i];
        const uint64_t d = dbl[v];
        beg = d >> 32;
        len = (uint32_t)d;
        add = tcnt[v];
    }
    uint64_t size = (uint64_t)len + add, incl = size;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint64_t y = __shfl_up(incl, d);
        if ((int)lane >= d) incl += y;
    }
    const uint64_t total = __shfl(incl, 63);
    if (total == 0) return;  // wave-uniform
    unsigned long long base = 0;
    if (lane == 63) base = atomicAdd(&gc->pool_top, (unsigned long long)total);
    base = __shfl(base, 63);
    if (i >= nt) return;
    const uint64_t nb = base + incl - size;
    if (nb + size > pool_cap) {
        atomicOr(&gc->err, ERR_POOL);
        tcnt[v] = 0xFFFFFFFFu;
        return;
    }
    for (uint32_t j = 0; j < len; ++j) pool[nb + j] = pool[beg + j];
    dbl[v] = (nb << 32) | size;
    tcnt[v] = len;
}

__global__ __launch_bounds__(RB) void k_ins_scatter(const int4 *rec, int64_t k, const uint64_t *slot,
                                                    const uint64_t *dbl, const uint32_t *tcnt, uint64_t *pool,
                                                    uint64_t *pos)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    if (i >= k) return;
    const int4 r = rec[i];
    if (r.w != 1) return;
    const uint64_t s = slot[i];
    const uint32_t ba = tcnt[r.x], bb = tcnt[r.y];
    if (ba == 0xFFFFFFFFu || bb == 0xFFFFFFFFu) {  // pool overflow reported by k_ins_alloc
        pos[i] = ~0ull;
        return;
    }
    const uint64_t pa = (dbl[r.x] >> 32) + ba + (s >> 32);
    const uint64_t pb = (dbl[r.y] >> 32) + bb + (uint32_t)s;
    pool[pa] = pack_entry((uint32_t)r.y, (uint32_t)r.z);
    pool[pb] = pack_entry((uint32_t)r.x, (uint32_t)r.z);
    pos[i] = (pa << 32) | pb;
}

__global__ __launch_bounds__(RB) void k_ins_reset(const uint32_t *touched, uint32_t max_touched, uint32_t *tcnt,
                                                  GraphCtl *gc)
{
    const uint32_t i = blockIdx.x * RB + threadIdx.x;
    const uint32_t nt = min(load_relaxed(&gc->touched), max_touched);
    if (i < nt) tcnt[touched[i]] = 0;
}

// ---------------------------------------------------------------------------
// Deletions: tree-edge test + tombstones (binary search in the sorted base row,
// scan of the short delta row).
// ---------------------------------------------------------------------------

__device__ __forceinline__ uint32_t tomb_row(GView g, uint32_t r, uint32_t c, uint32_t w)
{
    const uint64_t key = pack_entry(c, w);
    uint64_t lo = g.rp[r], hi = g.rp[r + 1];
    const uint64_t end = hi;
    uint32_t probes = 0;
    while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        ++probes;
        if (e_key(g.adj[mid]) < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    for (uint64_t j = lo; j < end; ++j) {
        const uint64_t e = g.adj[j];
        ++probes;
        if (e_key(e) != key) break;
        if (!(e & E_DEAD)) g.adj[j] = e | E_DEAD;
    }
    const uint64_t d = g.dbl[r];
    const uint64_t beg = d >> 32;
    const uint32_t len = (uint32_t)d;
    for (uint32_t j = 0; j < len; ++j) {
        const uint64_t e = g.pool[beg + j];
        if (e_key(e) == key && !(e & E_DEAD)) g.pool[beg + j] = e | E_DEAD;
    }
    return probes + len;
}

// Thread i handles record i >> kshift for tree i & mask; tree 0 also tombstones.
__global__ __launch_bounds__(RB) void k_classify(const int4 *rec, int64_t k, uint32_t kshift, GView g, uint64_t *st,
                                                 uint32_t *out_q, uint32_t *out_cnt, uint32_t *invl,
                                                 uint32_t *inv_total, uint32_t cap, Counters *c)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    const int64_t ri = i >> kshift;
    const uint32_t tree = (uint32_t)i & ((1u << kshift) - 1u);
    bool want = false;
    uint32_t victim = 0, probes = 0;
    if (ri < k) {
        const int4 r = rec[ri];
        if (r.w == 0) {
            const uint32_t a = (uint32_t)r.x, b = (uint32_t)r.y;
            if (a != b) {
                // deleteEdge's tree-edge test (gpuFunctions_undir.cuh:189-198): the
                // child endpoint of a tree edge roots an invalidated subtree.
                const uint32_t sa = (a << kshift) | tree, sb = (b << kshift) | tree;
                const uint64_t vb = st[sb];
                if (vb != ST_INF && st_parent(vb) == a) {
                    victim = sb;
                    want = true;
                } else {
                    const uint64_t va = st[sa];
                    if (va != ST_INF && st_parent(va) == b) {
                        victim = sa;
                        want = true;
                    }
                }
                if (want) want = atomicExch((unsigned long long *)&st[victim], (unsigned long long)ST_INF) != ST_INF;
            }
            if (tree == 0) {
                probes += tomb_row(g, a, b, (uint32_t)r.z);
                if (a != b) probes += tomb_row(g, b, a, (uint32_t)r.z);
            }
        }
    }
    wave_append(want, victim, out_q, out_cnt, cap, c);
    wave_append(want, victim, invl, inv_total, cap, c);
    const uint32_t p = wave_sum(probes);
    if (lane_id() == 0 && p) atomicAdd(&c->probes, (unsigned long long)p);
}

// Offers every inserted edge that survived the batch's deletions in both
// directions (the role of insertEdge, gpuFunctions_undir.cuh:217-251).
__global__ __launch_bounds__(RB) void k_ins_offer(const int4 *rec, int64_t k, uint32_t kshift, const uint64_t *pos,
                                                  const uint64_t *pool, uint64_t *st, uint32_t *stamp,
                                                  uint32_t epoch, uint32_t *out_q, uint32_t *out_cnt,
                                                  uint32_t cap, Counters *c)
{
    const int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x;
    const int64_t ri = i >> kshift;
    const uint32_t tree = (uint32_t)i & ((1u << kshift) - 1u);
    bool wa = false, wb = false;
    uint32_t sa_slot = 0, sb_slot = 0;
    if (ri < k) {
        const int4 r = rec[ri];
        if (r.w == 1 && r.x != r.y && pos[ri] != ~0ull && !(pool[pos[ri] >> 32] & E_DEAD)) {
            const uint32_t a = (uint32_t)r.x, b = (uint32_t)r.y;
            sa_slot = (a << kshift) | tree;
            sb_slot = (b << kshift) | tree;
            const uint32_t w = (uint32_t)r.z;
            const uint64_t sa = st[sa_slot], sb = st[sb_slot];
            const bool a_root = sa != ST_INF && st_parent(sa) == a;  // roots never change
            const bool b_root = sb != ST_INF && st_parent(sb) == b;
            if (sa != ST_INF && !b_root) {
                const uint64_t nd = (uint64_t)st_dist(sa) + w;
                if (nd >= DIST_INF) {
                    atomicOr(&c->err, ERR_DIST_OVERFLOW);
                } else {
                    const uint64_t cand = pack_state((uint32_t)nd, a);
                    if (cand < sb && cand < atomicMin((unsigned long long *)&st[sb_slot], (unsigned long long)cand))
                        wb = true;
                }
            }
            if (sb != ST_INF && !a_root) {
                const uint64_t nd = (uint64_t)st_dist(sb) + w;
                if (nd >= DIST_INF) {
                    atomicOr(&c->err, ERR_DIST_OVERFLOW);
                } else {
                    const uint64_t cand = pack_state((uint32_t)nd, b);
                    if (cand < sa && cand < atomicMin((unsigned long long *)&st[sa_slot], (unsigned long long)cand))
                        wa = true;
                }
            }
            if (wa && atomicExch(&stamp[sa_slot], epoch) == epoch) wa = false;
            if (wb && atomicExch(&stamp[sb_slot], epoch) == epoch) wb = false;
        }
    }
    wave_append(wa, sa_slot, out_q, out_cnt, cap, c);
    wave_append(wb, sb_slot, out_q, out_cnt, cap, c);
}

// ---------------------------------------------------------------------------
// Small helpers.
// ---------------------------------------------------------------------------

__global__ void k_reset_counters(Counters *c)
{
    if (threadIdx.x == 0) {
        Counters z{};
        *c = z;
    }
}

__global__ void k_reset_graphctl(GraphCtl *gc)
{
    if (threadIdx.x == 0) {
        gc->touched = 0;
        gc->err = 0;
    }
}

__global__ __launch_bounds__(RB) void k_set_sources(uint64_t *st, const uint32_t *src, uint32_t count,
                                                    uint32_t kshift, uint32_t per_tree, uint32_t *stamp,
                                                    uint32_t epoch, uint32_t *q, uint32_t *cnt, uint32_t cap,
                                                    Counters *c)
{
    const uint32_t j = blockIdx.x * RB + threadIdx.x;
    bool want = false;
    uint32_t slot = 0;
    if (j < count) {
        const uint32_t v = src[j];
        slot = (v << kshift) | (per_tree ? j : 0u);
        st[slot] = pack_state(0u, v);
        want = atomicExch(&stamp[slot], epoch) != epoch;
    }
    wave_append(want, slot, q, cnt, cap, c);
}

__global__ __launch_bounds__(RB) void k_fill_u64(uint64_t *p, uint64_t v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) p[i] = v;
}

__global__ __launch_bounds__(RB) void k_fill_u32(uint32_t *p, uint32_t v, uint64_t n)
{
    for (uint64_t i = (uint64_t)blockIdx.x * RB + threadIdx.x; i < n; i += (uint64_t)gridDim.x * RB) p[i] = v;
}

// Roots (parent == self) start the tight-graph BFS at level 0.
__global__ __launch_bounds__(RB) void k_lvl_init(uint32_t nslots, uint32_t kshift, const uint64_t *st, uint32_t *lvl,
                                                 uint32_t *best, uint32_t *stamp, uint32_t epoch, uint32_t *out_q,
                                                 uint32_t *out_cnt, uint32_t cap, Counters *c)
{
    const uint32_t s = blockIdx.x * RB + threadIdx.x;
    bool root = false;
    if (s < nslots) {
        const uint64_t v = st[s];
        root = v != ST_INF && st_parent(v) == (s >> kshift);
        lvl[s] = root ? 0u : LVL_INF;
        best[s] = 0xFFFFFFFFu;
        if (root) stamp[s] = epoch;
    }
    wave_append(root, s, out_q, out_cnt, cap, c);
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

__global__ void k_phase_init(Counters *c, uint32_t ring_base, uint32_t q_base, uint32_t epoch_base)
{
    if (threadIdx.x == 0) {
        c->iter = 0;
        c->blocks_done = 0;
        c->ring_base = ring_base;
        c->q_base = q_base;
        c->epoch_base = epoch_base;
    }
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + RB - 1) / RB); }

}  // namespace

// ---------------------------------------------------------------------------
// Host launchers.
// ---------------------------------------------------------------------------

void launch_frontier(int mode, const FArgs &a, int grid, hipStream_t s)
{
    switch (mode) {
    case M_RELAX: hipLaunchKernelGGL(k_frontier<M_RELAX>, dim3(grid), dim3(FB), 0, s, a); break;
    case M_INVAL: hipLaunchKernelGGL(k_frontier<M_INVAL>, dim3(grid), dim3(FB), 0, s, a); break;
    case M_SEED: hipLaunchKernelGGL(k_frontier<M_SEED>, dim3(grid), dim3(FB), 0, s, a); break;
    case M_TBFS: hipLaunchKernelGGL(k_frontier<M_TBFS>, dim3(grid), dim3(FB), 0, s, a); break;
    default: hipLaunchKernelGGL(k_frontier<M_PICK>, dim3(grid), dim3(FB), 0, s, a); break;
    }
}

void launch_reset_counters(Counters *c, hipStream_t s) { hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, s, c); }

void launch_reset_graphctl(GraphCtl *gc, hipStream_t s) { hipLaunchKernelGGL(k_reset_graphctl, dim3(1), dim3(64), 0, s, gc); }

void launch_ins_count(const int4 *rec, int64_t k, uint32_t *tcnt, uint64_t *slot, uint32_t *touched, GraphCtl *gc,
                      hipStream_t s)
{
    if (k > 0) hipLaunchKernelGGL(k_ins_count, dim3(blocks_for(k)), dim3(RB), 0, s, rec, k, tcnt, slot, touched, gc);
}

void launch_ins_alloc(const uint32_t *touched, uint32_t max_touched, uint64_t *dbl, uint32_t *tcnt, uint64_t *pool,
                      uint64_t pool_cap, GraphCtl *gc, hipStream_t s)
{
    if (max_touched > 0)
        hipLaunchKernelGGL(k_ins_alloc, dim3(blocks_for(max_touched)), dim3(RB), 0, s, touched, max_touched, dbl, tcnt,
                           pool, pool_cap, gc);
}

void launch_ins_scatter(const int4 *rec, int64_t k, const uint64_t *slot, const uint64_t *dbl, const uint32_t *tcnt,
                        uint64_t *pool, uint64_t *pos, hipStream_t s)
{
    if (k > 0) hipLaunchKernelGGL(k_ins_scatter, dim3(blocks_for(k)), dim3(RB), 0, s, rec, k, slot, dbl, tcnt, pool, pos);
}

void launch_ins_reset(const uint32_t *touched, uint32_t max_touched, uint32_t *tcnt, GraphCtl *gc, hipStream_t s)
{
    if (max_touched > 0)
        hipLaunchKernelGGL(k_ins_reset, dim3(blocks_for(max_touched)), dim3(RB), 0, s, touched, max_touched, tcnt, gc);
}

void launch_classify(const int4 *rec, int64_t k, uint32_t kshift, GView g, uint64_t *st, uint32_t *out_q,
                     uint32_t *out_cnt, uint32_t *invl, uint32_t *inv_total, uint32_t cap, Counters *c, hipStream_t s)
{
    if (k > 0)
        hipLaunchKernelGGL(k_classify, dim3(blocks_for(k << kshift)), dim3(RB), 0, s, rec, k, kshift, g, st, out_q,
                           out_cnt, invl, inv_total, cap, c);
}

void launch_ins_offer(const int4 *rec, int64_t k, uint32_t kshift, const uint64_t *pos, const uint64_t *pool,
                      uint64_t *st, uint32_t *stamp, uint32_t epoch, uint32_t *out_q, uint32_t *out_cnt, uint32_t cap,
                      Counters *c, hipStream_t s)
{
    if (k > 0)
        hipLaunchKernelGGL(k_ins_offer, dim3(blocks_for(k << kshift)), dim3(RB), 0, s, rec, k, kshift, pos, pool, st,
                           stamp, epoch, out_q, out_cnt, cap, c);
}

void launch_set_sources(uint64_t *st, const uint32_t *src, uint32_t count, uint32_t kshift, bool per_tree,
                        uint32_t *stamp, uint32_t epoch, uint32_t *q, uint32_t *cnt, uint32_t cap, Counters *c,
                        hipStream_t s)
{
    if (count > 0)
        hipLaunchKernelGGL(k_set_sources, dim3(blocks_for(count)), dim3(RB), 0, s, st, src, count, kshift,
                           per_tree ? 1u : 0u, stamp, epoch, q, cnt, cap, c);
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

void launch_lvl_init(uint32_t nslots, uint32_t kshift, const uint64_t *st, uint32_t *lvl, uint32_t *best,
                     uint32_t *stamp, uint32_t epoch, uint32_t *out_q, uint32_t *out_cnt, uint32_t cap, Counters *c,
                     hipStream_t s)
{
    if (nslots > 0)
        hipLaunchKernelGGL(k_lvl_init, dim3(blocks_for(nslots)), dim3(RB), 0, s, nslots, kshift, st, lvl, best, stamp,
                           epoch, out_q, out_cnt, cap, c);
}

void launch_phase_init(Counters *c, uint32_t ring_base, uint32_t q_base, uint32_t epoch_base, hipStream_t s)
{
    hipLaunchKernelGGL(k_phase_init, dim3(1), dim3(64), 0, s, c, ring_base, q_base, epoch_base);
}

void launch_apply_best(uint32_t nslots, uint32_t kshift, uint64_t *st, const uint32_t *best, Counters *c,
                       hipStream_t s)
{
    if (nslots > 0) hipLaunchKernelGGL(k_apply_best, dim3(blocks_for(nslots)), dim3(RB), 0, s, nslots, kshift, st, best, c);
}

}  // namespace dsssp
