// Dump harness around the reference's decremental target reader
// (random_walk_decremental/functions.cpp: CSRGraph::read_targets).  The
// decremental graph constructor is not exercised: with V = v_max it writes
// past its degree array (defect D2), which is undefined behaviour.
// Test infrastructure: pins the oracle's del.mtx batch layout.
//   ref_dump_dec targets <del.mtx> <out.txt>
#include <cstdio>
#include <cstring>

#include "functions.h"

int main(int argc, char** argv) {
    if (argc != 4 || std::strcmp(argv[1], "targets") != 0) return 2;
    FILE* out = std::fopen(argv[3], "w");
    if (!out) return 3;
    CSRGraph g;
    g.read_targets(argv[2], -1, 0, 10);
    std::fprintf(out, "%zu\n", g.sample_count);
    for (size_t i = 0; i < 2 * g.sample_count; ++i) std::fprintf(out, "%u\n", g.sample_nodes[i]);
    std::fclose(out);
    return 0;
}
