// op_tools.cpp -- the reference's side tools on top of libdsssp.so.
//
//   op_tools sssp  <graph> <n> [src]                   -> "v parent dist" on stdout
//       role of op_ssspSeq (seqSSSPwithDist_undir.cpp:183-198, README.md:39-48):
//       the tree is computed on the GPU, distances printed as exact integers
//       (the reference prints doubles with 6 significant digits), the source
//       is selectable (the reference hard-codes 0, :196).
//   op_tools ce    <graph> <n> <k> <percent_insert> [seed] [first_deletable]
//       role of op_cE (createChangedEdges.cpp:49-125, README.md:50-59): same
//       arguments and "a b w inst" lines; seeded, deletions without
//       replacement, insertions of distinct absent pairs with weights 1..99.
//   op_tools apply <graph> <n> <changes>                -> graph file on stdout
//       applies a change file to a graph file with the update's semantics
//       (insertions appended, every copy of a deleted (a,b,w) removed); the
//       reference has no such tool.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/dsssp.h"

static int die(const char *what, int code)
{
    fprintf(stderr, "op_tools: %s: %s\n", what, dsssp_last_error());
    return code ? code : 1;
}

static int usage()
{
    fprintf(stderr,
            "usage: op_tools sssp  <graph> <n> [src]\n"
            "       op_tools ce    <graph> <n> <k> <percent_insert> [seed] [first_deletable]\n"
            "       op_tools apply <graph> <n> <changes>\n");
    return 2;
}

static void put(std::string &buf, const char *fmt, int64_t x, int64_t y, int64_t z, int64_t q, int nf)
{
    char line[96];
    int len = nf == 3 ? snprintf(line, sizeof line, fmt, x, y, z) : snprintf(line, sizeof line, fmt, x, y, z, q);
    buf.append(line, (size_t)len);
    if (buf.size() > (1u << 20)) {
        fwrite(buf.data(), 1, buf.size(), stdout);
        buf.clear();
    }
}

int main(int argc, char **argv)
{
    if (argc < 4) return usage();
    const std::string cmd = argv[1];
    const int64_t n = atoll(argv[3]);
    if (n <= 0) return usage();
    int64_t m = 0;
    int32_t *a = nullptr, *b = nullptr, *w = nullptr;
    if (int s = dsssp_read_graph(argv[2], n, &m, &a, &b, &w)) return die("reading the graph", s);
    std::string out;
    if (cmd == "sssp") {
        const int32_t src = argc > 4 ? atoi(argv[4]) : 0;
        dsssp_graph *g = nullptr;
        dsssp_tree *t = nullptr;
        if (int s = dsssp_graph_create(n, m, a, b, w, DSSSP_GRAPH_DEFAULT, &g)) return die("building the graph", s);
        if (int s = dsssp_tree_create(g, src, nullptr, nullptr, DSSSP_TREE_DEFAULT, &t)) return die("computing", s);
        std::vector<int32_t> par((size_t)n);
        std::vector<int64_t> dist((size_t)n);
        if (int s = dsssp_tree_read(t, par.data(), dist.data())) return die("reading the tree", s);
        for (int64_t v = 0; v < n; ++v) {
            if (par[v] < 0)
                put(out, "%" PRId64 " %" PRId64 " %" PRId64 "\n", v, -1, -1, 0, 3);
            else
                put(out, "%" PRId64 " %" PRId64 " %" PRId64 "\n", v, par[v], dist[v], 0, 3);
        }
        dsssp_tree_destroy(t);
        dsssp_graph_destroy(g);
    } else if (cmd == "ce") {
        if (argc < 6) return usage();
        const int64_t k = atoll(argv[4]);
        const double pct = atof(argv[5]) / 100.0;
        const uint64_t seed = argc > 6 ? strtoull(argv[6], nullptr, 10) : 1;
        const int64_t first = argc > 7 ? atoll(argv[7]) : 0;
        std::vector<dsssp_change> c((size_t)k);
        if (int s = dsssp_gen_changes(n, m, a, b, w, k, pct, 1, 99, first, seed, c.data()))
            return die("generating changes", s);
        for (const auto &x : c) put(out, "%" PRId64 " %" PRId64 " %" PRId64 " %" PRId64 "\n", x.a, x.b, x.w, x.op, 4);
    } else if (cmd == "apply") {
        if (argc < 5) return usage();
        int64_t k = 0;
        dsssp_change *c = nullptr;
        if (int s = dsssp_read_changes(argv[4], n, &k, &c)) return die("reading the changes", s);
        std::set<std::tuple<int32_t, int32_t, int32_t>> dead;
        for (int64_t i = 0; i < k; ++i)
            if (c[i].op == 0) dead.insert({std::min(c[i].a, c[i].b), std::max(c[i].a, c[i].b), c[i].w});
        auto emit = [&](int32_t x, int32_t y, int32_t z) {
            if (!dead.count({std::min(x, y), std::max(x, y), z})) put(out, "%" PRId64 " %" PRId64 " %" PRId64 "\n", x, y, z, 0, 3);
        };
        for (int64_t i = 0; i < m; ++i) emit(a[i], b[i], w[i]);
        for (int64_t i = 0; i < k; ++i)
            if (c[i].op == 1) emit(c[i].a, c[i].b, c[i].w);
        dsssp_free(c);
    } else {
        return usage();
    }
    fwrite(out.data(), 1, out.size(), stdout);
    dsssp_free(a);
    dsssp_free(b);
    dsssp_free(w);
    return 0;
}
