/*
 * oracle.c -- CPU oracle for the dynamic SSSP batch update.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it is compiled into
 * oracle/liboracle.so and loaded by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg.  The product library (dynamicssspv2_amd) never
 * links, loads or calls it.
 *
 * Contents (each cites the reference file:line it restates):
 *   orc_build_csr          adjacency in file order     all_structure_undir.cuh:178-198,
 *                                                      SSSPmainForUndirected.cu:66-73
 *   orc_build_csr_any      the same rows in an unspecified order, built by all
 *                          threads (for the order-independent checks at 10^8+
 *                          edges)
 *   orc_djk                the reference's sequential  seqSSSPwithDist_undir.cpp:66-128
 *                          Dijkstra (label-correcting
 *                          with the inQ flag)
 *   orc_dijkstra_lazy      textbook lazy-deletion Dijkstra (context timing,
 *                          cross-check)
 *   orc_bellman_ford       brute force for tiny graphs
 *   orc_apply_changes      graph after a change batch  all_structure_undir.cuh:89-126,
 *                                                      gpuFunctions_undir.cuh:200-213
 *   orc_ref_update         sequential schedule of the  gpuFunctions_undir.cuh:125-433,
 *                          reference's update kernels  SSSPmainForUndirected.cu:196-281
 *   orc_canonical_parents  canonical shortest-path tree (smallest-id tight
 *                          neighbour, or BFS level in the tight subgraph)
 *   orc_certificate        O(n+m) exactness certificate for (dist, parent)
 *
 * How each function is pinned (tests/test_oracle.py):
 *   orc_build_csr_any: row multisets equal to orc_build_csr's on random
 *     multigraphs; Dijkstra, canonical parents and the certificate (planted
 *     mistake included) give the same answers on both.
 *   orc_djk: reference fixture SSSP.txt (= op_ssspSeq on graph.txt, parents and
 *     distances), oracle/_ref/op_ssspSeq built from the reference sources on
 *     random graphs, and Bellman-Ford on random tiny graphs.
 *   orc_dijkstra_lazy, orc_bellman_ford: against each other and orc_djk; the
 *     closed-form path/star/grid distances.
 *   orc_apply_changes: the toy batch cE.txt (hand-derived result),
 *     brute-force multiset arithmetic on random batches, and the exact output
 *     order against a direct restatement on a 200,000-edge multigraph (its
 *     filter runs on all threads and compacts in order).
 *   orc_ref_update: the toy known answer [0,4,6,6,4,9] and Dijkstra on the
 *     updated graph for random batches with distances below 999999.
 *   orc_canonical_parents: brute-force enumeration on tiny graphs.
 *   orc_certificate: planted mistakes (distance +-1, wrong parent, parent cycle,
 *     unreachable marked reachable) must each be rejected.
 *
 * Distances are int64; ORC_INF (INT64_MAX) marks an unreachable vertex and its
 * parent is -1.  A source's parent is itself.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_INF INT64_MAX

/* ------------------------------------------------------------------------ */
/* CSR in file order.                                                        */
/* ------------------------------------------------------------------------ */

/* Adjacency rows in file order: line i = (a,b,w) appends (col=a,w) to row b and
 * (col=b,w) to row a, the same per-row order as read_graphEdges
 * (all_structure_undir.cuh:186-194) and readin_graphU2
 * (seqSSSPwithDist_undir.cpp:49-58).  rp has n+1 entries; col/w hold 2m.
 * Returns 0, or -1 if an endpoint is outside [0,n). */
int orc_build_csr(int64_t n, int64_t m, const int32_t *a, const int32_t *b,
                  const int32_t *w, int64_t *rp, int32_t *col, int32_t *wt)
{
    int64_t i;
    int64_t *fill;
    for (i = 0; i <= n; i++) rp[i] = 0;
    for (i = 0; i < m; i++) {
        if (a[i] < 0 || a[i] >= n || b[i] < 0 || b[i] >= n) return -1;
        rp[a[i] + 1]++;
        rp[b[i] + 1]++;
    }
    for (i = 0; i < n; i++) rp[i + 1] += rp[i];
    fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    if (!fill) return -2;
    for (i = 0; i < n; i++) fill[i] = rp[i];
    for (i = 0; i < m; i++) {
        /* readin_graphU2 pushes into row n1 first, then row n2. */
        int64_t p = fill[a[i]]++;
        col[p] = b[i];
        wt[p] = w[i];
        p = fill[b[i]]++;
        col[p] = a[i];
        wt[p] = w[i];
    }
    free(fill);
    return 0;
}

/* The same adjacency with each row in an unspecified order, built by all
 * threads (atomic row cursors).  For the order-independent checks (Dijkstra
 * distances, canonical parents, the certificate) on graphs of 10^8+ edges.
 * Returns 0, -1 if an endpoint is outside [0,n), -2 on allocation failure. */
int orc_build_csr_any(int64_t n, int64_t m, const int32_t *a, const int32_t *b,
                      const int32_t *w, int64_t *rp, int32_t *col, int32_t *wt)
{
    int64_t i, bad = 0;
    int64_t *fill;
#pragma omp parallel for schedule(static)
    for (i = 0; i <= n; i++) rp[i] = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (i = 0; i < m; i++) {
        if (a[i] < 0 || a[i] >= n || b[i] < 0 || b[i] >= n) {
            bad++;
            continue;
        }
        __atomic_fetch_add(&rp[a[i] + 1], 1, __ATOMIC_RELAXED);
        __atomic_fetch_add(&rp[b[i] + 1], 1, __ATOMIC_RELAXED);
    }
    if (bad) return -1;
    for (i = 0; i < n; i++) rp[i + 1] += rp[i];
    fill = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
    if (!fill) return -2;
#pragma omp parallel for schedule(static)
    for (i = 0; i < n; i++) fill[i] = rp[i];
#pragma omp parallel for schedule(static)
    for (i = 0; i < m; i++) {
        int64_t p = __atomic_fetch_add(&fill[a[i]], 1, __ATOMIC_RELAXED);
        col[p] = b[i];
        wt[p] = w[i];
        p = __atomic_fetch_add(&fill[b[i]], 1, __ATOMIC_RELAXED);
        col[p] = a[i];
        wt[p] = w[i];
    }
    free(fill);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Binary min-heap of (key, vertex) pairs, ordered lexicographically.        */
/* ------------------------------------------------------------------------ */

typedef struct {
    int64_t key;
    int32_t v;
} heap_item;

typedef struct {
    heap_item *a;
    int64_t size, cap;
} heap;

static int heap_less(heap_item x, heap_item y)
{
    return x.key < y.key || (x.key == y.key && x.v < y.v);
}

static int heap_push(heap *h, int64_t key, int32_t v)
{
    int64_t i;
    if (h->size == h->cap) {
        int64_t nc = h->cap ? h->cap * 2 : 1024;
        heap_item *na = (heap_item *)realloc(h->a, sizeof(heap_item) * (size_t)nc);
        if (!na) return -1;
        h->a = na;
        h->cap = nc;
    }
    i = h->size++;
    h->a[i].key = key;
    h->a[i].v = v;
    while (i > 0) {
        int64_t p = (i - 1) / 2;
        if (!heap_less(h->a[i], h->a[p])) break;
        heap_item t = h->a[i];
        h->a[i] = h->a[p];
        h->a[p] = t;
        i = p;
    }
    return 0;
}

static heap_item heap_pop(heap *h)
{
    heap_item top = h->a[0];
    int64_t i = 0;
    h->a[0] = h->a[--h->size];
    for (;;) {
        int64_t l = 2 * i + 1, r = l + 1, s = i;
        if (l < h->size && heap_less(h->a[l], h->a[s])) s = l;
        if (r < h->size && heap_less(h->a[r], h->a[s])) s = r;
        if (s == i) break;
        heap_item t = h->a[i];
        h->a[i] = h->a[s];
        h->a[s] = t;
        i = s;
    }
    return top;
}

/* ------------------------------------------------------------------------ */
/* The reference's sequential Dijkstra.                                      */
/* ------------------------------------------------------------------------ */

/* Restates djk (seqSSSPwithDist_undir.cpp:66-128):
 *   dist[] = cap (the reference uses 1000*n, :71-74), parent[] = -1,
 *   dist[src] = 0, parent[src] = src (:76-81);
 *   a min-priority queue of (dist, vertex) pairs (:88-90);
 *   pop u, clear inQ[u] (:99-101); scan u's row in file order (:105-108);
 *   on dist[v] > dist[u]+w set dist[v], parent[v]=u and push (dist[v], v)
 *   only when v is not already queued (:112-121).
 * A queued vertex keeps its stale key when improved again, so this is a
 * label-correcting search; final distances are exact below the cap.
 * Every queued vertex is queued once, so keys are distinct and the pop order
 * does not depend on the heap implementation.
 * cap <= 0 selects the reference's cap 1000*n.  Vertices left at the cap are
 * reported as unreachable (ORC_INF, parent -1), which is how the reference's
 * SSSP reader treats them (all_structure_undir.cuh:153-156).
 * Returns the number of pops, or -1 on allocation failure. */
int64_t orc_djk(int64_t n, const int64_t *rp, const int32_t *col, const int32_t *wt,
                int32_t src, int64_t cap, int32_t *parent, int64_t *dist)
{
    heap h = {0, 0, 0};
    char *inq;
    int64_t i, pops = 0;
    if (cap <= 0) cap = 1000 * n;
    inq = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    if (!inq) return -1;
    for (i = 0; i < n; i++) {
        dist[i] = cap;
        parent[i] = -1;
    }
    dist[src] = 0;
    parent[src] = src;
    if (heap_push(&h, 0, src)) { free(inq); return -1; }
    inq[src] = 1;
    while (h.size > 0) {
        heap_item it = heap_pop(&h);
        int32_t u = it.v;
        int64_t j, du;
        pops++;
        inq[u] = 0;
        du = dist[u];
        for (j = rp[u]; j < rp[u + 1]; j++) {
            int32_t v = col[j];
            int64_t nd = du + wt[j];
            if (dist[v] > nd) {
                dist[v] = nd;
                parent[v] = u;
                if (!inq[v]) {
                    if (heap_push(&h, nd, v)) { free(h.a); free(inq); return -1; }
                    inq[v] = 1;
                }
            }
        }
    }
    for (i = 0; i < n; i++) {
        if (dist[i] >= cap) {
            dist[i] = ORC_INF;
            parent[i] = -1;
        }
    }
    free(h.a);
    free(inq);
    return pops;
}

/* Textbook Dijkstra with lazy deletion (stale heap entries skipped).
 * Parent = the first vertex that set the final distance.  Multiple sources
 * (each with distance 0, parent itself) are allowed.  Returns pops or -1. */
int64_t orc_dijkstra_lazy(int64_t n, const int64_t *rp, const int32_t *col,
                          const int32_t *wt, const int32_t *srcs, int32_t nsrc,
                          int32_t *parent, int64_t *dist)
{
    heap h = {0, 0, 0};
    int64_t i, pops = 0;
    for (i = 0; i < n; i++) {
        dist[i] = ORC_INF;
        parent[i] = -1;
    }
    for (i = 0; i < nsrc; i++) {
        dist[srcs[i]] = 0;
        parent[srcs[i]] = srcs[i];
        if (heap_push(&h, 0, srcs[i])) return -1;
    }
    while (h.size > 0) {
        heap_item it = heap_pop(&h);
        int32_t u = it.v;
        int64_t j;
        if (it.key != dist[u]) continue;
        pops++;
        for (j = rp[u]; j < rp[u + 1]; j++) {
            int32_t v = col[j];
            int64_t nd = it.key + wt[j];
            if (nd < dist[v]) {
                dist[v] = nd;
                parent[v] = u;
                if (heap_push(&h, nd, v)) { free(h.a); return -1; }
            }
        }
    }
    free(h.a);
    return pops;
}

/* Bellman-Ford over an edge list (undirected), for tiny graphs.
 * Returns the number of passes, or -1 if it did not settle in n passes. */
int64_t orc_bellman_ford(int64_t n, int64_t m, const int32_t *a, const int32_t *b,
                         const int32_t *w, const int32_t *srcs, int32_t nsrc, int64_t *dist)
{
    int64_t i, pass;
    for (i = 0; i < n; i++) dist[i] = ORC_INF;
    for (i = 0; i < nsrc; i++) dist[srcs[i]] = 0;
    for (pass = 0; pass <= n; pass++) {
        int changed = 0;
        for (i = 0; i < m; i++) {
            int32_t x = a[i], y = b[i];
            if (dist[x] != ORC_INF && dist[x] + w[i] < dist[y]) { dist[y] = dist[x] + w[i]; changed = 1; }
            if (dist[y] != ORC_INF && dist[y] + w[i] < dist[x]) { dist[x] = dist[y] + w[i]; changed = 1; }
        }
        if (!changed) return pass + 1;
    }
    return -1;
}

/* ------------------------------------------------------------------------ */
/* Graph after a change batch.                                               */
/* ------------------------------------------------------------------------ */

typedef struct {
    int32_t lo, hi, w;
} del_key;

static int del_key_cmp(const void *pa, const void *pb)
{
    const del_key *x = (const del_key *)pa, *y = (const del_key *)pb;
    if (x->lo != y->lo) return x->lo < y->lo ? -1 : 1;
    if (x->hi != y->hi) return x->hi < y->hi ? -1 : 1;
    if (x->w != y->w) return x->w < y->w ? -1 : 1;
    return 0;
}

/* The live edge multiset after one batch, in the order the reference's
 * adjacency holds it.  Insertions (op 1) are appended to the graph in batch
 * order before anything else happens (readin_changes,
 * all_structure_undir.cuh:106-117).  A deletion (op 0) of (x,y,w) then removes
 * EVERY edge between x and y whose weight equals w, in either orientation,
 * including parallel copies and copies inserted by the same batch
 * (deleteEdge's two row scans, gpuFunctions_undir.cuh:200-213).  Any other op
 * value is ignored (readin_changes only acts on 0 and 1, :106,119).
 * chg is k records of (a,b,w,op) int32.  Outputs have capacity m+k.
 * Returns the number of surviving edges, or -1 on allocation failure. */
int64_t orc_apply_changes(int64_t n, int64_t m, const int32_t *a, const int32_t *b,
                          const int32_t *w, int64_t k, const int32_t *chg,
                          int32_t *oa, int32_t *ob, int32_t *ow)
{
    int64_t i, cnt = 0, nd = 0, total = m + k;
    int64_t *part;
    del_key *dk;
    char *keep;
    int nt = 1, used = 1;
    (void)n;
    dk = (del_key *)malloc(sizeof(del_key) * (size_t)(k > 0 ? k : 1));
    keep = (char *)malloc((size_t)(total > 0 ? total : 1));
    if (!dk || !keep) { free(dk); free(keep); return -1; }
    for (i = 0; i < k; i++) {
        if (chg[4 * i + 3] == 0) {
            int32_t x = chg[4 * i], y = chg[4 * i + 1];
            dk[nd].lo = x < y ? x : y;
            dk[nd].hi = x < y ? y : x;
            dk[nd].w = chg[4 * i + 2];
            nd++;
        }
    }
    qsort(dk, (size_t)nd, sizeof(del_key), del_key_cmp);
    /* Survival of each edge is independent of the others: decided by all
     * threads, then compacted in the original order (the result is the same
     * as one sequential pass). */
#pragma omp parallel for schedule(static)
    for (i = 0; i < total; i++) {
        int32_t x, y, ww;
        del_key key;
        if (i < m) {
            x = a[i]; y = b[i]; ww = w[i];
        } else {
            const int32_t *c = chg + 4 * (i - m);
            keep[i] = 0;
            if (c[3] != 1) continue;
            x = c[0]; y = c[1]; ww = c[2];
        }
        key.lo = x < y ? x : y;
        key.hi = x < y ? y : x;
        key.w = ww;
        keep[i] = !(nd > 0 && bsearch(&key, dk, (size_t)nd, sizeof(del_key), del_key_cmp));
    }
#ifdef _OPENMP
    nt = omp_get_max_threads();
#endif
    part = (int64_t *)calloc((size_t)nt + 1, sizeof(int64_t));
    if (!part) { free(dk); free(keep); return -1; }
#pragma omp parallel num_threads(nt)
    {
        int t = 0, j, nth = 1;
        int64_t lo, hi, c = 0, p;
#ifdef _OPENMP
        t = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
#pragma omp single
        used = nth;
        lo = total * t / nth;
        hi = total * (t + 1) / nth;
        for (p = lo; p < hi; p++) c += keep[p];
        part[t + 1] = c;
#pragma omp barrier
        c = 0;
        for (j = 0; j <= t; j++) c += part[j];
        for (p = lo; p < hi; p++) {
            if (!keep[p]) continue;
            if (p < m) {
                oa[c] = a[p]; ob[c] = b[p]; ow[c] = w[p];
            } else {
                const int32_t *q = chg + 4 * (p - m);
                oa[c] = q[0]; ob[c] = q[1]; ow[c] = q[2];
            }
            c++;
        }
    }
    for (i = 1; i <= used; i++) cnt += part[i];
    free(part);
    free(keep);
    free(dk);
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* Sequential schedule of the reference's update kernels.                    */
/* ------------------------------------------------------------------------ */

#define REF_INVALID 999999 /* gpuFunctions_undir.cuh:190,195,238,323-324 */

/* Restates the live update path of the reference, with every kernel's threads
 * executed one after another in index order (one legal schedule of the CUDA
 * program), on int32 RT_Vertex state (all_structure_undir.cuh:66-75):
 *   STEP1  deleteEdge  gpuFunctions_undir.cuh:179-215 (tree-edge test, then
 *                      tombstone wt=-1 of every matching (col,wt) in both rows)
 *          insertEdge  gpuFunctions_undir.cuh:217-251 (orient by Dist, two
 *                      loop guards, one relaxation)
 *   STEP2A filterAffectedNodes / updateNeighbors_del loop
 *                      gpuFunctions_undir.cuh:125-144,312-349;
 *                      SSSPmainForUndirected.cu:219-239
 *   STEP2B filterAllAffected / updateNeighbors loop
 *                      gpuFunctions_undir.cuh:162-174,370-433;
 *                      SSSPmainForUndirected.cu:263-281
 * Inputs: the adjacency (rp/col/wt) must already contain the batch's
 * insertions appended per row (all_structure_undir.cuh:106-117); wt is
 * modified in place (tombstones).  par/dst/upd are the RT_Vertex fields,
 * read from the SSSP file as read_SSSP does (all_structure_undir.cuh:133-171:
 * unreachable = parent -1, Dist 9999999).  ins/del are (a,b,w) triples.
 * max_rounds bounds each loop.  Returns the number of STEP2B rounds, or -1 if
 * a loop hit max_rounds, or -2 on allocation failure. */
int64_t orc_ref_update(int64_t n, const int64_t *rp, const int32_t *col, int32_t *wt,
                       int32_t *par, int32_t *dst, int32_t *upd,
                       int64_t ni, const int32_t *ins, int64_t ndel, const int32_t *del,
                       int64_t max_rounds)
{
    int64_t i, j, rounds = 0;
    int32_t *list = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int64_t counter;
    int change;
    if (!list) return -2;

    /* STEP1: deleteEdge, one thread per deletion. */
    for (i = 0; i < ndel; i++) {
        int32_t n1 = del[3 * i], n2 = del[3 * i + 1], ew = del[3 * i + 2];
        if (par[n2] == n1) {
            dst[n2] = REF_INVALID;
            upd[n2] = 1;
        } else if (par[n1] == n2) {
            dst[n1] = REF_INVALID;
            upd[n1] = 1;
        }
        for (j = rp[n2]; j < rp[n2 + 1]; j++)
            if (col[j] == n1 && wt[j] == ew) wt[j] = -1;
        for (j = rp[n1]; j < rp[n1 + 1]; j++)
            if (col[j] == n2 && wt[j] == ew) wt[j] = -1;
    }
    /* STEP1: insertEdge, one thread per insertion. */
    for (i = 0; i < ni; i++) {
        int32_t a = ins[3 * i], b = ins[3 * i + 1], ew = ins[3 * i + 2];
        int32_t x, y;
        int flag = 1;
        if (dst[b] > dst[a]) { x = a; y = b; } else { x = b; y = a; }
        if (par[x] == y) flag = 0;
        if (dst[y] == REF_INVALID && upd[y] == 1) flag = 0;
        if (dst[y] > dst[x] + ew && flag) {
            par[y] = x;
            dst[y] = dst[x] + ew;
            upd[y] = 2;
        }
    }

    /* STEP2A: invalidate the subtrees below deleted tree edges. */
    if (ndel > 0) {
        int64_t guard = 0;
        counter = 0;
        for (i = 0; i < n; i++)
            if (upd[i] == 1) list[counter++] = (int32_t)i;
        change = 1;
        while (change > 0) {
            if (++guard > max_rounds) { free(list); return -1; }
            change = 0;
            for (i = 0; i < counter; i++) {
                int32_t u = list[i];
                upd[u] = 2;
                for (j = rp[u]; j < rp[u + 1]; j++) {
                    int32_t x = col[j];
                    if (wt[j] < 0) continue;
                    if (par[x] == u && dst[x] != REF_INVALID) {
                        dst[x] = REF_INVALID;
                        upd[x] = 1;
                        change = 1;
                    }
                }
            }
            counter = 0;
            for (i = 0; i < n; i++)
                if (upd[i] == 1) list[counter++] = (int32_t)i;
        }
    }

    /* STEP2B: pull + push relaxation over the affected set until quiet. */
    counter = 0;
    for (i = 0; i < n; i++)
        if (upd[i] == 1 || upd[i] == 2) list[counter++] = (int32_t)i;
    change = 1;
    while (change == 1) {
        if (++rounds > max_rounds) { free(list); return -1; }
        change = 0;
        for (i = 0; i < counter; i++) {
            int32_t u = list[i];
            upd[u] = 0;
            for (j = rp[u]; j < rp[u + 1]; j++) {
                int32_t v = col[j], ew = wt[j];
                if (ew < 0) continue;
                if (dst[u] > dst[v] + ew) {
                    if (par[v] != u) {
                        dst[u] = dst[v] + ew;
                        upd[u] = 2;
                        par[u] = v;
                        change = 1;
                        continue;
                    }
                }
                if (dst[v] > dst[u] + ew) {
                    if (par[u] != v) {
                        dst[v] = dst[u] + ew;
                        upd[v] = 2;
                        par[v] = u;
                        change = 1;
                    }
                }
            }
        }
        counter = 0;
        for (i = 0; i < n; i++)
            if (upd[i] == 1 || upd[i] == 2) list[counter++] = (int32_t)i;
    }
    free(list);
    return rounds;
}

/* ------------------------------------------------------------------------ */
/* Canonical shortest-path tree and the exactness certificate.               */
/* ------------------------------------------------------------------------ */

/* Canonical parents for exact distances `dist` on the live CSR (rp/col/wt,
 * tombstoned entries have wt < 0 and are skipped).  Sources keep themselves;
 * unreachable vertices get -1.
 *   rule 0: parent[v] = smallest u with a live edge (u,v,w), d[u]+w == d[v].
 *           Acyclic whenever every weight is positive.
 *   rule 1: let lvl be the BFS depth from the sources in the "tight" graph
 *           (u->v when d[u]+w == d[v]); parent[v] = smallest tight u with
 *           lvl[u] == lvl[v]-1.  Acyclic also with zero weights.
 * Returns the number of reachable non-source vertices that have no tight
 * in-edge (0 for exact distances), or -1 on allocation failure. */
int64_t orc_canonical_parents(int64_t n, const int64_t *rp, const int32_t *col,
                              const int32_t *wt, const int64_t *dist,
                              const int32_t *srcs, int32_t nsrc, int rule,
                              int32_t *parent)
{
    int64_t i, j, bad = 0;
    char *is_src = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    if (!is_src) return -1;
    for (i = 0; i < nsrc; i++) is_src[srcs[i]] = 1;
    if (rule == 0) {
#pragma omp parallel for schedule(dynamic, 4096) private(j) reduction(+ : bad)
        for (i = 0; i < n; i++) {
            int32_t best = -1;
            if (is_src[i]) { parent[i] = (int32_t)i; continue; }
            if (dist[i] == ORC_INF) { parent[i] = -1; continue; }
            for (j = rp[i]; j < rp[i + 1]; j++) {
                int32_t u = col[j];
                if (wt[j] < 0 || dist[u] == ORC_INF) continue;
                if (dist[u] + wt[j] == dist[i] && (best < 0 || u < best)) best = u;
            }
            parent[i] = best;
            if (best < 0) bad++;
        }
    } else {
        int64_t *lvl = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
        int32_t *q = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
        int64_t qh = 0, qt = 0;
        if (!lvl || !q) { free(lvl); free(q); free(is_src); return -1; }
        for (i = 0; i < n; i++) lvl[i] = -1;
        for (i = 0; i < nsrc; i++) {
            if (lvl[srcs[i]] < 0) { lvl[srcs[i]] = 0; q[qt++] = srcs[i]; }
        }
        while (qh < qt) {
            int32_t u = q[qh++];
            for (j = rp[u]; j < rp[u + 1]; j++) {
                int32_t v = col[j];
                if (wt[j] < 0 || lvl[v] >= 0) continue;
                if (dist[u] != ORC_INF && dist[u] + wt[j] == dist[v]) {
                    lvl[v] = lvl[u] + 1;
                    q[qt++] = v;
                }
            }
        }
#pragma omp parallel for schedule(dynamic, 4096) private(j) reduction(+ : bad)
        for (i = 0; i < n; i++) {
            int32_t best = -1;
            if (is_src[i]) { parent[i] = (int32_t)i; continue; }
            if (dist[i] == ORC_INF) { parent[i] = -1; continue; }
            for (j = rp[i]; j < rp[i + 1]; j++) {
                int32_t u = col[j];
                if (wt[j] < 0 || dist[u] == ORC_INF || lvl[u] < 0) continue;
                if (dist[u] + wt[j] == dist[i] && lvl[u] == lvl[i] - 1 && (best < 0 || u < best)) best = u;
            }
            parent[i] = best;
            if (best < 0) bad++;
        }
        free(lvl);
        free(q);
    }
    free(is_src);
    return bad;
}

/* Exactness certificate for (dist, parent) on the live CSR (wt < 0 = dead):
 *   C1 every source has dist 0 and parent itself;
 *   C2 every live edge (u,v,w) with finite d[u] has d[v] <= d[u]+w;
 *   C3 every finite non-source v has a live edge (p,v,w), p = parent[v],
 *      with d[p]+w == d[v];
 *   C4 following parents from any finite vertex reaches a source (no cycle);
 *   C5 every infinite vertex has parent -1.
 * C1-C4 hold exactly when d is the shortest-path distance from the sources and
 * parent is a shortest-path tree (C3+C4 make d an attained path length, C2
 * makes it a lower bound along every shortest path).
 * Returns 0 if all hold, else the failing condition number (1-5) and writes
 * the first offending vertex to *where.  -1 on allocation failure. */
int orc_certificate(int64_t n, const int64_t *rp, const int32_t *col, const int32_t *wt,
                    const int64_t *dist, const int32_t *parent,
                    const int32_t *srcs, int32_t nsrc, int64_t *where)
{
    int64_t i, j;
    char *state;
    char *is_src = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    int32_t *stack;
    *where = -1;
    if (!is_src) return -1;
    for (i = 0; i < nsrc; i++) {
        int32_t s = srcs[i];
        is_src[s] = 1;
        if (dist[s] != 0 || parent[s] != s) { *where = s; free(is_src); return 1; }
    }
    /* C2 and C3/C5 are checked by all threads; the failure reported is the
     * one a sequential pass in vertex order meets first. */
    {
        int64_t first = n;
#pragma omp parallel for schedule(dynamic, 4096) private(j) reduction(min : first)
        for (i = 0; i < n; i++) {
            if (dist[i] == ORC_INF || i >= first) continue;
            for (j = rp[i]; j < rp[i + 1]; j++) {
                if (wt[j] < 0) continue;
                if (dist[col[j]] > dist[i] + wt[j]) { first = i; break; }
            }
        }
        if (first < n) {
            for (j = rp[first]; j < rp[first + 1]; j++) {
                int32_t v = col[j];
                if (wt[j] >= 0 && dist[v] > dist[first] + wt[j]) { *where = v; break; }
            }
            free(is_src);
            return 2;
        }
    }
    {
        int64_t first = n;
#pragma omp parallel for schedule(dynamic, 4096) private(j) reduction(min : first)
        for (i = 0; i < n; i++) {
            int32_t p;
            int ok = 0;
            if (i >= first) continue;
            if (dist[i] == ORC_INF) {
                if (parent[i] != -1) first = i;
                continue;
            }
            if (is_src[i]) continue;
            p = parent[i];
            if (p < 0 || p >= n || p == i || dist[p] == ORC_INF) { first = i; continue; }
            for (j = rp[i]; j < rp[i + 1]; j++) {
                if (col[j] == p && wt[j] >= 0 && dist[p] + wt[j] == dist[i]) { ok = 1; break; }
            }
            if (!ok) first = i;
        }
        if (first < n) {
            *where = first;
            free(is_src);
            return dist[first] == ORC_INF ? 5 : 3;
        }
    }
    /* C4: iterative walk with three colours (0 new, 1 on stack, 2 done). */
    state = (char *)calloc((size_t)(n > 0 ? n : 1), 1);
    stack = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!state || !stack) { free(state); free(stack); free(is_src); return -1; }
    for (i = 0; i < n; i++) {
        int64_t top = 0;
        int32_t v = (int32_t)i;
        if (state[i] || dist[i] == ORC_INF) continue;
        while (!state[v] && !is_src[v]) {
            state[v] = 1;
            stack[top++] = v;
            v = parent[v];
        }
        if (state[v] == 1) { *where = v; free(state); free(stack); free(is_src); return 4; }
        while (top > 0) state[stack[--top]] = 2;
        state[v] = 2;
    }
    free(state);
    free(stack);
    free(is_src);
    return 0;
}
