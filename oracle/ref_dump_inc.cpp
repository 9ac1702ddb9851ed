// Dump harness around the reference's incremental parsers
// (random_walk_incremental/functions.cpp: read_graph, read_targets).
// Test infrastructure: pins the oracle's parse restatement.
//   ref_dump_inc graph   <adj_sparse.mtx> <out.txt>
//   ref_dump_inc targets <ext.mtx>        <out.txt>
#include <cstdio>
#include <cstring>

#include "functions.h"

int main(int argc, char** argv) {
    if (argc != 4) return 2;
    FILE* out = std::fopen(argv[3], "w");
    if (!out) return 3;
    if (std::strcmp(argv[1], "graph") == 0) {
        CSRGraph g = read_graph(argv[2], true, 0, 1);
        std::fprintf(out, "%zu %zu\n", g.vert_count, g.mtx.size());
        for (size_t v = 0; v < g.vert_count; ++v) {
            std::fprintf(out, "%u", g.degree[v]);
            for (size_t p = g.begin[v]; p < g.begin[v + 1]; ++p)
                std::fprintf(out, " %u:%.9g", g.adj[p], (double)g.weight[p]);
            std::fprintf(out, "\n");
        }
    } else {
        Targets t = read_targets(argv[2], -1, 1);
        std::fprintf(out, "%zu\n", t.target_count);
        for (size_t i = 0; i < 2 * t.target_count; ++i) std::fprintf(out, "%u\n", t.nodes[i]);
        for (size_t i = 0; i < t.target_count; ++i) std::fprintf(out, "%.9g\n", (double)t.weights[i]);
    }
    std::fclose(out);
    return 0;
}
