// op_main.cpp -- drop-in replacement for the reference CLI
//     op_main <graph> <n> <m> <sssp> <changes> [--out <sssp_out>] [--json]
// (SSSPmainForUndirected.cu:17-32, README.md:13).  Same positional arguments,
// same file formats and the same progress/timer lines on stdout
// (SSSPmainForUndirected.cu:47-64,140-143,206-292,298-312); the update itself
// runs through libdsssp.so on one gfx950 GPU.  Additions: --out writes the
// updated tree as "v parent dist" (the reference never writes it), --json
// prints one machine-readable line of statistics.  Errors go to stderr and the
// exit code is non-zero (the reference always exits 0).
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/dsssp.h"

using clk = std::chrono::high_resolution_clock;

static int die(const char *what, int code)
{
    fprintf(stderr, "op_main: %s: %s\n", what, dsssp_last_error());
    return code ? code : 1;
}

static long long us_since(clk::time_point t0)
{
    return (long long)std::chrono::duration_cast<std::chrono::microseconds>(clk::now() - t0).count();
}

int main(int argc, char **argv)
{
    std::vector<const char *> pos;
    const char *out_path = nullptr;
    bool json = false;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--out") && i + 1 < argc)
            out_path = argv[++i];
        else if (!strcmp(argv[i], "--json"))
            json = true;
        else
            pos.push_back(argv[i]);
    }
    if (pos.size() != 5) {
        fprintf(stderr,
                "usage: op_main <graph> <number_of_nodes> <number_of_edges> <input_SSSP> <change_edges> "
                "[--out <sssp>] [--json]\n");
        return 2;
    }
    const int64_t n = atoll(pos[1]);
    const int64_t m_arg = atoll(pos[2]);
    if (n <= 0) {
        fprintf(stderr, "op_main: number_of_nodes must be positive\n");
        return 2;
    }

    printf("Reading input graph...\n");
    auto t0 = clk::now();
    int64_t m = 0;
    int32_t *a = nullptr, *b = nullptr, *w = nullptr;
    if (int s = dsssp_read_graph(pos[0], n, &m, &a, &b, &w)) return die("reading the graph", s);
    printf("Reading input graph completed\n");
    printf("Time taken to read input graph: %lld microseconds\n", us_since(t0));
    if (m != m_arg)
        fprintf(stderr, "op_main: warning: graph file holds %" PRId64 " edges, argument says %" PRId64 "\n", m, m_arg);

    printf("Reading input changed edges data...\n");
    t0 = clk::now();
    int64_t k = 0;
    dsssp_change *chg = nullptr;
    if (int s = dsssp_read_changes(pos[4], n, &k, &chg)) return die("reading the change edges", s);
    int64_t total_ins = 0;
    for (int64_t i = 0; i < k; ++i) total_ins += chg[i].op == 1;
    printf("Reading input changed edges data completed. totalInsertion:%" PRId64 "\n", total_ins);
    printf("Time taken to read input changed edges: %lld microseconds\n", us_since(t0));

    printf("Transferring graph data from CPU to GPU\n");
    t0 = clk::now();
    dsssp_graph *g = nullptr;
    if (int s = dsssp_graph_create(n, m, a, b, w, DSSSP_GRAPH_DEFAULT, &g)) return die("building the graph", s);
    printf("**Time taken to transfer graph data from CPU to GPU: %.3f milliseconds**\n", us_since(t0) / 1000.0);

    printf("Reading input SSSP tree data...\n");
    t0 = clk::now();
    std::vector<int32_t> parent((size_t)n);
    std::vector<int64_t> dist((size_t)n);
    if (int s = dsssp_read_sssp(pos[3], n, parent.data(), dist.data())) return die("reading the SSSP tree", s);
    printf("Reading input SSSP tree data completed\n");
    printf("Time taken to read input input SSSP tree: %lld microseconds\n", us_since(t0));
    dsssp_tree *t = nullptr;
    if (int s = dsssp_tree_create(g, -1, parent.data(), dist.data(), DSSSP_TREE_DEFAULT, &t))
        return die("loading the SSSP tree", s);
    printf("reading input data completed\n");

    dsssp_batch *batch = nullptr;
    if (int s = dsssp_batch_upload(g, chg, k, &batch)) return die("uploading the change batch", s);
    dsssp_stats st{};
    // The stage timers below need the events between the update's phases.
    dsssp_tuning tu{};
    if (int s = dsssp_get_tuning(&tu)) return die("reading the tuning", s);
    tu.phase_timing = 1;
    if (int s = dsssp_set_tuning(&tu)) return die("setting the tuning", s);
    if (int s = dsssp_update_batch(g, t, batch, &st)) return die("updating", s);
    const double step1 = st.t_insert_ms + st.t_classify_ms;
    printf("**Time taken for SSSP update STEP1: %.3f milliseconds**\n", step1);
    printf("**Time taken for updateNeighbors_del: %.3f milliseconds**\n", st.t_invalidate_ms);
    printf("**Time taken for updateNeighbors: %.3f milliseconds**\n", st.t_seed_ms + st.t_relax_ms);
    printf("**Time taken for SSSP update STEP2: %.3f milliseconds**\n", st.t_invalidate_ms + st.t_seed_ms + st.t_relax_ms);
    printf("****Total Time taken for SSSP update: %.3f milliseconds****\n", st.t_update_ms);
    printf("Total affected nodes by Delete edge only: %" PRId64 "\n", st.n_invalidated);

    if (int s = dsssp_tree_read(t, parent.data(), dist.data())) return die("reading the tree back", s);
    const int64_t x = n < 40 ? n : 40;
    printf("from GPU: \n[");
    for (int64_t i = 0; i < x; ++i) printf("%" PRId64 ":%" PRId64 ":%d ", i, dist[i] == DSSSP_INF ? -1 : dist[i], parent[i]);
    printf("]\n");
    if (out_path) {
        if (int s = dsssp_write_sssp(out_path, n, parent.data(), dist.data())) return die("writing the SSSP tree", s);
    }
    if (json) {
        printf("{\"n\": %" PRId64 ", \"m\": %" PRId64 ", \"k\": %" PRId64 ", \"insert\": %" PRId64
               ", \"delete\": %" PRId64 ", \"invalidated\": %" PRId64 ", \"relax_rounds\": %" PRId64
               ", \"t_update_ms\": %.4f, \"changed_edges_per_s\": %.1f, \"alg_GBps\": %.2f}\n",
               n, m, k, st.n_insert, st.n_delete, st.n_invalidated, st.relax_rounds, st.t_update_ms,
               st.t_update_ms > 0 ? (double)k / (st.t_update_ms / 1e3) : 0.0,
               st.t_update_ms > 0 ? st.alg_bytes / (st.t_update_ms / 1e3) / 1e9 : 0.0);
    }
    dsssp_batch_destroy(batch);
    dsssp_tree_destroy(t);
    dsssp_graph_destroy(g);
    dsssp_free(a);
    dsssp_free(b);
    dsssp_free(w);
    dsssp_free(chg);
    return 0;
}
