/*
 * dsssp.h -- C ABI of the MI355X (gfx950) dynamic SSSP update library
 * (dynamicssspv2_amd/libdsssp.so).
 *
 * The reference (ArindamK92/DynamicSSSPv2) has no library interface: its
 * boundary is the process CLI
 *     op_main <graph> <n> <m> <sssp> <changes>      (SSSPmainForUndirected.cu:17-32)
 * plus three text formats (README.md:22-37).  Internally it hands raw AoS
 * device pointers to its kernels (gpuFunctions_undir.cuh:125-433).  This header
 * is the boundary those pieces map onto; each entry point names the reference
 * code it replaces.  Every entry point returns an int status (DSSSP_OK = 0);
 * dsssp_last_error() gives the message of the calling thread's last failure.
 * No C++ types, exceptions or torch types cross this ABI.
 *
 * Semantics kept from the reference:
 *   - vertices are 0-based int32 ids below n; weights are int32 >= 0;
 *   - the graph is undirected; an edge "a b w" is one line of the graph file;
 *   - a change record (a, b, w, op) has op 0 = delete, 1 = insert
 *     (inst_status, README.md:33-37); other op values are ignored
 *     (all_structure_undir.cuh:106,119);
 *   - insertions are applied to the graph before deletions, and a deletion of
 *     (a,b,w) removes every a-b edge of weight w, parallel copies and copies
 *     inserted by the same batch included (all_structure_undir.cuh:106-117,
 *     gpuFunctions_undir.cuh:200-213);
 *   - a tree root is a vertex whose parent is itself; an unreachable vertex has
 *     parent -1 (all_structure_undir.cuh:153-156, seqSSSPwithDist_undir.cpp:81).
 * Differences (documented fixes of reference defects): distances are exact
 * int64 on output (held as uint32 on the device; a distance that would reach
 * 2^32-1 fails with DSSSP_ERR_OVERFLOW), vertices that end unreachable get
 * parent -1 instead of a stale parent, and updates are race-free (packed
 * 64-bit atomicMin of (distance << 32 | parent)).
 *
 * Threading: a graph handle and everything attached to it (trees, batches)
 * must be used from one host thread at a time.  Each graph owns one HIP stream
 * of default priority; calls are synchronous unless stated.
 */
#ifndef DSSSP_H
#define DSSSP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------- */
#define DSSSP_OK 0
#define DSSSP_ERR_ARG 1       /* null pointer, negative size, bad flag      */
#define DSSSP_ERR_RANGE 2     /* vertex id outside [0, n)                    */
#define DSSSP_ERR_WEIGHT 3    /* negative weight                             */
#define DSSSP_ERR_OVERFLOW 4  /* a distance reached 2^32-1                   */
#define DSSSP_ERR_NOMEM 5     /* host or device allocation failed            */
#define DSSSP_ERR_HIP 6       /* HIP runtime error                           */
#define DSSSP_ERR_NODEV 7     /* no usable gfx950 device                     */
#define DSSSP_ERR_IO 8        /* file could not be opened / written          */
#define DSSSP_ERR_FORMAT 9    /* malformed line in a text file               */
#define DSSSP_ERR_TREE 10     /* input tree inconsistent with the graph      */
#define DSSSP_ERR_BUDGET 11   /* served form: memory budget exceeded         */
#define DSSSP_ERR_STATE 12    /* call not valid in the object's state        */

/* Distance reported for an unreachable vertex (its parent is -1). */
#define DSSSP_INF INT64_MAX

/* One change record; same fields and meaning as a change-file line
 * "a b w inst" (README.md:33-37, all_structure_undir.cuh:52-57). */
typedef struct {
    int32_t a, b, w, op;
} dsssp_change;

typedef struct dsssp_graph dsssp_graph;  /* device-resident mutable CSR          */
typedef struct dsssp_tree dsssp_tree;    /* device-resident SSSP tree of a graph  */
typedef struct dsssp_batch dsssp_batch;  /* change batch resident in HBM          */

/* Per-update statistics.  Times are measured with HIP events on the graph's
 * stream.  t_update_ms spans "batch resident in HBM" to "converged": upkeep of
 * the resident graph, insertion into the resident CSR, deletion classification
 * and tombstoning, subtree invalidation, seeding and relaxation.  The times of
 * the phases (t_upkeep_ms, t_insert_ms .. t_relax_ms) are measured only with
 * the phase_timing knob set, and read 0 otherwise.  t_h2d_ms is the
 * host-to-device copy of a batch given as host records (dsssp_update),
 * outside t_update_ms. */
typedef struct {
    int64_t n_insert, n_delete;     /* records with op 1 / op 0                */
    int64_t n_invalidated;          /* vertices reset by subtree invalidation  */
    int64_t inv_rounds;             /* invalidation rounds launched            */
    int64_t relax_rounds;           /* relaxation rounds launched              */
    int64_t relax_launches;         /* relaxation kernel launches (incl. empty)*/
    int64_t frontier_vertices;      /* frontier vertices processed (relax)     */
    int64_t relax_edges;            /* adjacency entries scanned (relax)       */
    int64_t relax_improved;         /* successful packed atomicMin (relax)     */
    int64_t relax_enqueued;         /* vertices enqueued (relax)               */
    int64_t other_edges;            /* entries scanned by invalidation + seed  */
    int64_t probes;                 /* entries read by deletion searches       */
    double t_h2d_ms;                /* batch host->device copy (dsssp_update)  */
    double t_update_ms;             /* total device time of the update         */
    double t_insert_ms, t_classify_ms, t_invalidate_ms, t_seed_ms, t_relax_ms;
    double t_relax_kernel_ms;       /* sum of relaxation kernel durations      */
    double alg_bytes;               /* algorithmic HBM bytes, whole update     */
    double alg_bytes_relax;         /* algorithmic HBM bytes, relaxation only  */
    int64_t relax_sweeps;           /* relaxation rounds run as full sweeps    */
    double t_sweep_ms;              /* dsssp_tree_relax_full: its edge pass    */
    double alg_bytes_sweep;         /* and that pass's algorithmic HBM bytes   */
    double t_upkeep_ms;             /* part of t_update_ms: room in the delta
                                       pool, edge map load and item lists made
                                       before the batch's first kernel         */
    int64_t n_rescued;              /* invalidation candidates kept under another
                                       tight parent (their subtrees untouched)  */
} dsssp_stats;

/* Tuning knobs (zero fields keep the current value). */
typedef struct {
    int32_t block;            /* threads per workgroup of frontier kernels: 128, 256
                                 (default) or 512                                    */
    int32_t grid;             /* least workgroups per frontier launch (more when the
                                 bitmap needs them to be read in one pass)          */
    int32_t heavy_degree;     /* a tile (the set slots of one 64-bit frontier word)
                                 with more row entries than this is split           */
    int32_t chunk;            /* entries per split piece (multiple of 64)              */
    int32_t round_batch;      /* rounds per group between convergence checks; a
                                 loop's first group follows the depth of the last
                                 loop of its kind, later groups double up to 64      */
    int32_t kernel_timing;    /* 1: time every relaxation launch with events           */
    int32_t use_graph;        /* 1: launch round groups as captured hipGraphs          */
    int32_t edges_per_thread; /* adjacency entries in flight per lane: 1, 2 or 4       */
    int32_t tile_entries;     /* row entries a wave takes per tile of up to 64 frontier
                                 slots; slots past it are deferred as pieces         */
    int32_t local_tiles;      /* tiles of slots a wave has just marked that it visits
                                 itself in the same launch (-1: none)                 */
    int32_t delta;            /* distance-ordered relaxation: a round visits only
                                 slots below a threshold that advances in steps of
                                 delta once nothing below it is left.  -1: off (every
                                 marked slot is visited in the next round);
                                 DSSSP_DELTA_AUTO (default): twice the graph's mean
                                 weight, for trees whose distances span more than 64
                                 such steps (high-diameter graphs), else off         */
    int32_t sweep;            /* push-vs-full-sweep threshold, per mille of the slots:
                                 a relaxation round after one that marked at least
                                 this share relaxes every slot, rows read in vertex
                                 order, instead of the marked ones.  -1: never.
                                 Only without distance ordering and below 64 trees */
    int32_t give_up_ms;       /* a loop of rounds still running after this long is
                                 abandoned: the call scrubs the tree and returns
                                 DSSSP_ERR_STATE (default 120000)                     */
    int32_t rescue_degree;    /* invalidation rescue: a vertex about to be reset
                                 whose row has at least this many entries first
                                 looks for another live neighbour that is tight for
                                 its distance and still valid; if there is one it
                                 keeps its distance under the smallest-id such
                                 neighbour and its subtree is not invalidated.
                                 -1: off.  DSSSP_RESCUE_AUTO (default): rows of at
                                 least 1024 entries, for batches with at least n/8
                                 deletions (each stage costs a few launches, which
                                 smaller batches do not win back).  Not with zero
                                 weights or 64 trees                                */
    int32_t rescue_levels;    /* rescue stages per update: the deleted tree edges'
                                 children, then the children found by the first
                                 rescue_levels - 1 invalidation rounds (default 16) */
    int32_t phase_timing;     /* 1: events between an update's phases fill
                                 t_upkeep_ms and t_insert_ms .. t_relax_ms (each
                                 event holds the stream for a few microseconds,
                                 about 2 % of a 1 ms update); 0 (default): only
                                 t_update_ms is measured, the phase times read 0  */
} dsssp_tuning;

#define DSSSP_DELTA_AUTO (-2)
#define DSSSP_RESCUE_AUTO (-2)

/* ---- library ---------------------------------------------------------- */
const char *dsssp_last_error(void);
const char *dsssp_version(void);
/* Select the HIP device for objects created afterwards by this thread. */
int dsssp_set_device(int device);
int dsssp_get_tuning(dsssp_tuning *t);
int dsssp_set_tuning(const dsssp_tuning *t);
/* Frees memory the library returned (file readers, generators). */
void dsssp_free(void *p);
/* Launches of known HBM footprint, for calibrating hardware byte counters
 * against the update's access patterns (no reference counterpart): kind 0
 * copies a buffer of `bytes` (reads and writes `bytes`); kind 1 reads `count`
 * 8-byte words at hashed positions of a buffer of `bytes`.  Each is launched
 * `reps` times after one warm-up; *ms = device time per launch. */
int dsssp_probe(int32_t kind, uint64_t bytes, uint64_t count, int32_t reps, double *ms);

/* ---- text formats (host only) --------------------------------------------
 * Replace read_graphEdges / readin_changes / read_SSSP
 * (all_structure_undir.cuh:178-198, 89-126, 133-171) and add the writer the
 * reference lacks (it only prints 40 distances, SSSPmainForUndirected.cu:298-312).
 * Readers accept any run of blanks between fields and a missing final newline,
 * and reject malformed lines, ids outside [0,n) (n > 0) and negative weights. */
int dsssp_read_graph(const char *path, int64_t n, int64_t *m,
                     int32_t **a, int32_t **b, int32_t **w);
int dsssp_read_changes(const char *path, int64_t n, int64_t *k, dsssp_change **chg);
/* parent/dist have n entries.  Missing vertices and parent -1 are unreachable
 * (all_structure_undir.cuh:144-156); the distance token of an unreachable
 * vertex may be anything (the reference's Dijkstra prints it as a lossy
 * double, seqSSSPwithDist_undir.cpp:172). */
int dsssp_read_sssp(const char *path, int64_t n, int32_t *parent, int64_t *dist);
/* "v parent dist" per line, exact integers; unreachable -> "v -1 -1". */
int dsssp_write_sssp(const char *path, int64_t n, const int32_t *parent, const int64_t *dist);

/* ---- synthetic inputs (host, deterministic for a seed, multi-threaded) ---- */
/* Random recursive spanning tree (vertex i>0 joined to a uniform j<i) followed
 * by uniform distinct pairs until m undirected edges; weights U[wlo,whi].
 * The first n-1 edges are the spanning tree.  a/b/w have m entries. */
int dsssp_gen_tree_plus_random(int64_t n, int64_t m, int32_t wlo, int32_t whi, uint64_t seed,
                               int32_t *a, int32_t *b, int32_t *w);
/* R-MAT (a,b,c, d = 1-a-b-c), 2^scale vertices, edge_factor*2^scale draws,
 * deduplicated, symmetrised, self-loops dropped; weights U[wlo,whi] per pair.
 * Returns library-allocated arrays (dsssp_free) with *m undirected edges. */
int dsssp_gen_rmat(int32_t scale, int32_t edge_factor, double pa, double pb, double pc,
                   int32_t wlo, int32_t whi, uint64_t seed, int64_t *m,
                   int32_t **a, int32_t **b, int32_t **w);
/* side x side 4-neighbour grid, a fraction drop of the edges removed at random,
 * restricted to the component of vertex 0 (v = row*side + col); weights U[wlo,whi]. */
int dsssp_gen_grid(int32_t side, double drop, int32_t wlo, int32_t whi, uint64_t seed, int64_t *m,
                   int32_t **a, int32_t **b, int32_t **w);
/* k change records: round(k*ins_frac) insertions of distinct absent pairs
 * (no self-loops) with weights U[wlo,whi], the rest deletions sampled without
 * replacement from edges [first_deletable, m); records in random order. */
int dsssp_gen_changes(int64_t n, int64_t m, const int32_t *a, const int32_t *b, const int32_t *w,
                      int64_t k, double ins_frac, int32_t wlo, int32_t whi,
                      int64_t first_deletable, uint64_t seed, dsssp_change *out);

/* Grid workload batch: round(k*del_frac) deletions, a fraction tree_frac of
 * them drawn from edges of the current tree (parent[] of the tree before the
 * batch), the rest from other edges; the remaining records insert distinct
 * absent pairs at Manhattan distance 1..max_manhattan, weights U[wlo,whi]. */
int dsssp_gen_changes_grid(int32_t side, int64_t m, const int32_t *a, const int32_t *b, const int32_t *w,
                           const int32_t *parent, int64_t k, double del_frac, double tree_frac,
                           int32_t max_manhattan, int32_t wlo, int32_t whi, uint64_t seed, dsssp_change *out);

/* ---- device graph ----------------------------------------------------------
 * Replaces the host CSR flatten + UVM upload (SSSPmainForUndirected.cu:66-91):
 * the CSR is built on the host (rows sorted by (neighbour, weight)) and copied
 * once into HBM, where later batches edit it in place. */
#define DSSSP_GRAPH_DEFAULT 0u
int dsssp_graph_create(int64_t n, int64_t m, const int32_t *a, const int32_t *b, const int32_t *w,
                       uint32_t flags, dsssp_graph **out);
int dsssp_graph_destroy(dsssp_graph *g);
/* n, live undirected edges, HBM bytes held by the graph. */
int dsssp_graph_info(const dsssp_graph *g, int64_t *n, int64_t *m_live, uint64_t *device_bytes);
/* Live undirected edges (each once, a <= b).  With a == NULL only *m is set. */
int dsssp_graph_export(dsssp_graph *g, int64_t *m, int32_t *a, int32_t *b, int32_t *w);
/* Upkeep state of the resident CSR.  base_entries + delta_entries is what
 * scans of every row read (live and tombstoned entries); dead_entries counts
 * the tombstoned ones since the last repack (a copy that duplicate deletion
 * records of one batch kill together counts twice).  A repack (on the device, inside
 * the next update's time, t_upkeep_ms) drops every tombstoned entry and
 * compacts the delta pool; it runs before a batch whose insertions would
 * outgrow the pool, and when tombstones exceed 1/8 of the scanned entries.
 * No reference counterpart: the reference rebuilds its adjacency on the host
 * for every run (all_structure_undir.cuh:106-117, SSSPmainForUndirected.cu:66-72). */
typedef struct {
    int64_t n, base_entries, delta_entries, dead_entries;
    uint64_t pool_top, pool_cap, emap_slots;
    int64_t repacks;
} dsssp_graph_upkeep;
int dsssp_graph_stats(const dsssp_graph *g, dsssp_graph_upkeep *out);
/* A repack now (synchronous); the graph's trees are unaffected. */
int dsssp_graph_repack(dsssp_graph *g);
/* Device-to-device copy of the whole graph state; dst must have been created
 * from a graph of the same n and base size (dsssp_graph_clone makes one). */
int dsssp_graph_clone(dsssp_graph *src, dsssp_graph **out);
int dsssp_graph_copy(dsssp_graph *dst, dsssp_graph *src);

/* ---- device tree -----------------------------------------------------------
 * Replaces read_SSSP into UVM (SSSPmainForUndirected.cu:146-165).
 * parent == NULL: compute the tree from src on the GPU (the op_ssspSeq role,
 * seqSSSPwithDist_undir.cpp:66-128), giving every vertex the smallest-id
 * neighbour on a shortest path as parent.
 * Otherwise parent/dist (n entries) is the given tree; DSSSP_TREE_CANONICAL
 * re-derives parents by that smallest-id rule (distances must be exact). */
#define DSSSP_TREE_DEFAULT 0u
#define DSSSP_TREE_CANONICAL 1u
int dsssp_tree_create(dsssp_graph *g, int32_t src, const int32_t *parent, const int64_t *dist,
                      uint32_t flags, dsssp_tree **out);
int dsssp_tree_destroy(dsssp_tree *t);
int dsssp_tree_read(dsssp_tree *t, int32_t *parent, int64_t *dist);
int dsssp_tree_copy(dsssp_tree *dst, dsssp_tree *src);
/* One full relaxation pass, then relaxation rounds until nothing changes (the
 * relaxation of updateNeighbors, gpuFunctions_undir.cuh:370-433, started from
 * every vertex instead of the affected ones; the reference has no such entry
 * point).  The pass relaxes every live edge in both directions once, from the
 * row of its higher-id endpoint (for 64 trees per vertex: every entry of every
 * reachable slot, pushed).  Distances that are lengths of real paths converge
 * to the shortest ones; on a converged tree nothing changes and, with
 * positive weights, stats->relax_improved is 0: a device-side check that no
 * live edge shortens any distance (with zero weights, equal-distance parents
 * may move to smaller ids, and the parent repair restores the canonical ones).
 * stats: the relaxation fields and times, and the pass's own device time and
 * algorithmic bytes (t_sweep_ms, alg_bytes_sweep; zero for 64 trees per
 * vertex, whose first round is timed with the others). */
int dsssp_tree_relax_full(dsssp_tree *t, dsssp_stats *stats);
/* One shortest-path forest grown from several roots at once (each root has
 * distance 0 and parent itself), computed on the GPU.  Used for many graphs
 * held in one device graph, one root per component.  No reference
 * counterpart: the reference handles one source per run
 * (seqSSSPwithDist_undir.cpp:196). */
int dsssp_tree_create_roots(dsssp_graph *g, int32_t nroots, const int32_t *roots, uint32_t flags,
                            dsssp_tree **out);

/* ---- batched form: K trees (sources) on one graph ---------------------------
 * ntrees independent SSSP trees, tree j from srcs[j], computed on the GPU and
 * updated together by dsssp_update / dsssp_update_batch: the batch is applied
 * to the graph once and every tree is brought to the exact tree of the updated
 * graph in the same launches (state of vertex v, tree j at slot v*K' + j with
 * K' = ntrees rounded up to a power of two, so one vertex's trees share cache
 * lines).  Each tree equals what the plain path gives for its source.  The
 * reference runs one source per process (SSSPmainForUndirected.cu:146-165). */
int dsssp_forest_create(dsssp_graph *g, int32_t ntrees, const int32_t *srcs, uint32_t flags,
                        dsssp_tree **out);
int dsssp_tree_count(const dsssp_tree *t, int32_t *ntrees);
int dsssp_forest_read(dsssp_tree *t, int32_t tree, int32_t *parent, int64_t *dist);

/* ---- the update --------------------------------------------------------------
 * Replaces STEP1 + STEP2 of op_main (SSSPmainForUndirected.cu:196-292):
 * deleteEdge/insertEdge (gpuFunctions_undir.cuh:179-251), the invalidation loop
 * (updateNeighbors_del + filterAffectedNodes, :125-144, :312-349) and the
 * relaxation loop (updateNeighbors + filterAllAffected, :162-174, :370-433).
 * Applies the batch to the graph and brings the tree to the exact SSSP tree of
 * the updated graph.  Changes are validated on the host (ids in [0,n),
 * weights >= 0) before anything is launched; stats may be NULL. */
int dsssp_update(dsssp_graph *g, dsssp_tree *t, const dsssp_change *chg, int64_t k,
                 dsssp_stats *stats);
/* Upload a batch once (validated) so that updates start with it in HBM. */
int dsssp_batch_upload(dsssp_graph *g, const dsssp_change *chg, int64_t k, dsssp_batch **out);
int dsssp_batch_destroy(dsssp_batch *b);
int dsssp_update_batch(dsssp_graph *g, dsssp_tree *t, dsssp_batch *b, dsssp_stats *stats);
/* Allocation-free variant of dsssp_update_batch.  No reference counterpart:
 * it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  d_chg points to k records
 * (dsssp_change layout) already in device memory of the graph's device; they
 * are validated on the device.  The call allocates nothing: its scratch (the
 * delta-block position of each record copy, validation counters) is a static
 * device array of 16 MiB + 256 B per device, shared by every call on that
 * device -- calls that overlap on one device race.  Limits: k <= 2097152
 * records (DSSSP_ERR_ARG beyond); the batch must fit the graph's delta pool,
 * edge map and the tree's item lists as they are (DSSSP_ERR_STATE when they
 * would have to grow, or when a graph with zero weights has not yet allocated
 * its parent-repair buffers: apply one update with dsssp_update_batch first).
 * No per-launch timing events or captured round graphs are used. */
int dsssp_update_noalloc(dsssp_graph *g, dsssp_tree *t, const dsssp_change *d_chg, int64_t k,
                         dsssp_stats *stats);
/* Autotuner (no reference counterpart; the reference has fixed launch
 * shapes, SSSPmainForUndirected.cu:200-203,229-236,271-277): coordinate descent
 * over block, grid, edges_per_thread, heavy_degree, tile_entries, local_tiles,
 * round_batch, delta and sweep, each trial applying the nb resident batches to a
 * snapshot of g and t (restored afterwards, so both end as they started) and
 * timed on the device; a trial whose tree differs from the first one's is an
 * error.  Sets the fastest tuning globally and returns it with its device time
 * per batch.  Run it while no other update is in flight. */
int dsssp_autotune(dsssp_graph *g, dsssp_tree *t, dsssp_batch *const *batches, int32_t nb, int32_t passes,
                   dsssp_tuning *best, double *best_ms);

/* ---- served form -------------------------------------------------------------
 * Resident instances (each a graph with its own source) kept in HBM under a
 * memory budget, and request loops that merge batches arriving for different
 * instances -- and consecutive batches of one instance -- into shared launches
 * (continuous batching).  The instances are spread over `lanes`, each one
 * device graph with its own stream and loop thread, so the merged updates of
 * different lanes overlap on the GPU.  Consecutive requests of one instance are
 * merged only while no request inserts an edge (a, b, w) that an earlier
 * request of the same merge deletes, the one case where one batch would leave
 * a different graph.  Each instance's result is bit-identical to applying its
 * requests alone, one update each.  No reference counterpart: the reference
 * applies one batch to one graph per process (SSSPmainForUndirected.cu:27-335). */
typedef struct dsssp_server dsssp_server;
typedef struct {
    uint64_t mem_budget_bytes;  /* device bytes all admitted instances may use      */
    int32_t max_merge;          /* most requests merged into one launch (0: no cap) */
    int32_t lanes;              /* concurrent lanes (0: 1); more when the instances
                                   do not fit one device graph's 32-bit pool offsets */
    int64_t max_records;        /* most records per merged launch (0: no cap)       */
} dsssp_server_config;
typedef struct {
    int32_t instance, merged;   /* merged: requests that shared the launch          */
    int64_t records;
    double queue_ms;            /* submit -> picked by the loop                     */
    double latency_ms;          /* submit -> result in HBM                          */
    double update_ms;           /* device time of the merged update                 */
} dsssp_request_stats;
typedef struct {
    int32_t instances, pad0;
    uint64_t budget_bytes, used_bytes;
    int64_t launches, records, requests, pending;
    double busy_ms;
} dsssp_server_totals;

int dsssp_server_create(const dsssp_server_config *cfg, dsssp_server **out);
/* Admits an instance; DSSSP_ERR_BUDGET if it does not fit the budget left.
 * Before dsssp_server_start it is placed with the others at start; while the
 * server runs it gets a lane of its own, built in the calling thread while the
 * other lanes go on serving, and takes submissions when this call returns. */
int dsssp_server_add_instance(dsssp_server *s, int64_t n, int64_t m, const int32_t *a, const int32_t *b,
                              const int32_t *w, int32_t src, int32_t *id);
/* Stops submissions to an instance, waits until its submitted requests are
 * applied, and retires it.  Its lane's device memory, and the lane's share of
 * the budget, is released once every instance of that lane is retired (an
 * instance admitted while the server runs has a lane of its own, so its share
 * comes back at once).  Its id is not reused. */
int dsssp_server_remove_instance(dsssp_server *s, int32_t id);
/* Uploads every admitted instance, computes their trees, starts the loop. */
int dsssp_server_start(dsssp_server *s);
/* Queues a batch for one instance (ids local to the instance); returns at once. */
int dsssp_server_submit(dsssp_server *s, int32_t id, const dsssp_change *chg, int64_t k, uint64_t *ticket);
/* Blocks until the request is applied; returns its status. */
int dsssp_server_wait(dsssp_server *s, uint64_t ticket, dsssp_request_stats *st);
/* Tree of one instance after all of its submitted requests. */
int dsssp_server_read(dsssp_server *s, int32_t id, int32_t *parent, int64_t *dist);
int dsssp_server_info(dsssp_server *s, dsssp_server_totals *tot);
int dsssp_server_destroy(dsssp_server *s);

#ifdef __cplusplus
}
#endif

#endif /* DSSSP_H */
