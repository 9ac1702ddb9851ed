// decremental_update -- drop-in for random_walk_decremental/main.cu:318-468.
//
//   ./decremental_update <graph name>
//
// Reads ../dataset/<graph>/updated_adj_sparse.mtx, updated_dense.mtx and
// del.mtx (all 1-based), processes the deletions in snapshot batches of
// ceil(M/10) (512 walkers, 100 steps, 16 recovery walkers, as the reference
// hard-codes at main.cu:332-335) and writes added_edges.mtx (1-based pairs;
// the reference writes them in hash order, this writes them sorted).
// Extensions (environment): DYGRASS_SEED (default 0), DYGRASS_DEVICE (0),
// DYGRASS_QUIET=1 suppresses the per-deletion "not found" lines.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/dygrass.h"

static void fail(const std::string& what, dyg_ctx ctx) {
    std::cout << what << ": " << dyg_last_error(ctx) << std::endl;
    std::exit(EXIT_FAILURE);
}

static uint64_t env_u64(const char* name, uint64_t dflt) {
    const char* s = std::getenv(name);
    return s ? std::strtoull(s, nullptr, 0) : dflt;
}

int main(int argc, char* argv[]) {
    if (argc != 2) {
        std::cout << "Input: ./decremental_update <1. graph name>\n";
        std::exit(0);
    }
    const std::string name = argv[1];
    const std::string sparse_file = "../dataset/" + name + "/updated_adj_sparse.mtx";
    const std::string dense_file = "../dataset/" + name + "/updated_dense.mtx";
    const std::string del_file = "../dataset/" + name + "/del.mtx";
    const std::string output_file = "../dataset/" + name + "/added_edges.mtx";
    const uint32_t n_steps = 100, n_walkers = 512, divisor = 10, n_heuristic_walkers = 16;
    const bool quiet = env_u64("DYGRASS_QUIET", 0) != 0;

    std::cout << "Read graph and sparsifier" << std::endl;
    dyg_mtx dense = nullptr, sparse = nullptr, del = nullptr;
    dyg_status st = dyg_mtx_read_graph(dense_file.c_str(), 1, 1, &dense);
    if (st == DYG_ERR_IO) {
        std::perror("Error opening file");
        std::exit(-1);
    }
    if (st != DYG_OK) fail("Error reading " + dense_file, nullptr);
    st = dyg_mtx_read_edges(sparse_file.c_str(), 1, &sparse);
    if (st == DYG_ERR_IO) {
        std::perror("Error opening file");
        std::exit(-1);
    }
    if (st != DYG_OK) fail("Error reading " + sparse_file, nullptr);
    std::cout << "Read targets" << std::endl;
    st = dyg_mtx_read_edges(del_file.c_str(), 0, &del);
    if (st == DYG_ERR_IO) {
        std::perror("File open error");
        std::exit(-1);
    }
    if (st != DYG_OK) fail("Error reading " + del_file, nullptr);
    const uint32_t V = dyg_mtx_nv(dense);
    if (dyg_mtx_nv(sparse) != V) {
        // GPU_Dual_Graph asserts equal vertex counts (gpu_graph.cuh:185)
        std::cout << "sparsifier and graph differ in vertex count" << std::endl;
        std::exit(EXIT_FAILURE);
    }

    dyg_ctx ctx = nullptr;
    if (dyg_ctx_create((int)env_u64("DYGRASS_DEVICE", 0), &ctx) != DYG_OK)
        fail("HIP context", nullptr);
    dyg_dual d = nullptr;
    if (dyg_dual_create(ctx, V, dyg_mtx_row_ptr(dense), dyg_mtx_col(dense), dyg_mtx_w(dense),
                        dyg_mtx_line_u(sparse), dyg_mtx_line_v(sparse), dyg_mtx_nlines(sparse),
                        &d) != DYG_OK)
        fail("graph upload", ctx);
    const uint64_t old_edge_num = dyg_dual_sparse_size(d);
    const long mult = (long)dyg_dual_key_multiplier(d);

    dyg_walk_params p{};
    p.walkers = n_walkers;
    p.max_steps = n_steps;
    p.threshold = 0.0f;
    p.flags = 0;
    p.seed = env_u64("DYGRASS_SEED", 0);

    const uint64_t M = dyg_mtx_nlines(del);
    const uint32_t* du = dyg_mtx_line_u(del);
    const uint32_t* dv = dyg_mtx_line_v(del);
    for (uint64_t i = 0; i < M; ++i)
        if (du[i] >= V || dv[i] >= V) {
            std::cout << "del.mtx names a vertex outside the graph" << std::endl;
            std::exit(EXIT_FAILURE);
        }
    const uint64_t I = (M + divisor - 1) / divisor;
    std::vector<int8_t> status;
    std::vector<uint64_t> ids;
    auto start = std::chrono::high_resolution_clock::now();
    uint64_t b0 = 0;
    bool first = true;
    do {
        const uint64_t n = std::min<uint64_t>(I, M - b0);
        status.assign(n ? n : 1, 0);
        ids.resize(n ? n : 1);
        for (uint64_t i = 0; i < n; ++i) ids[i] = b0 + i;
        uint32_t queued = 0;
        if (dyg_dual_remove(d, du + b0, dv + b0, ids.data(), (uint32_t)n, status.data(),
                            &queued) != DYG_OK)
            fail("removeAndUpdateDelEdges", ctx);
        if (!quiet) {
            for (uint64_t i = 0; i < n; ++i) {
                long a = du[b0 + i], b = dv[b0 + i];
                if (a > b) std::swap(a, b);
                if (status[i] == 0)
                    std::cout << "dense_map not found: a: " << a << " b: " << b
                              << " key: " << a * mult + b << std::endl;
                else if (status[i] == 1)
                    std::cout << "sparse_map not found: a: " << a << " b: " << b
                              << " key: " << a * mult + b << std::endl;
            }
        }
        if (first) {
            std::cout << "Graph copied to GPU" << std::endl;
            start = std::chrono::high_resolution_clock::now();
            first = false;
        }
        uint32_t no_path = 0;
        uint64_t added = 0;
        if (dyg_dual_walk(d, &p, &no_path, &added, nullptr) != DYG_OK) fail("random walk", ctx);
        std::cout << " result copied" << std::endl;
        std::cout << "No path count: " << no_path << std::endl;
        std::cout << "Added edges: " << added << std::endl;
        if (no_path > 0) {
            if (dyg_dual_heuristic(d, n_heuristic_walkers, p.seed) != DYG_OK)
                fail("heuristic recovery", ctx);
            std::cout << " result copied" << std::endl;
        }
        b0 += n;
    } while (b0 < M);
    const auto end = std::chrono::high_resolution_clock::now();
    const auto duration = std::chrono::duration_cast<std::chrono::milliseconds>(end - start);
    std::cout << "Time taken: " << duration.count() << " milliseconds" << std::endl;

    const size_t new_edge_num = dyg_dual_sparse_size(d);
    std::cout << "Increased edges number: " << (size_t)(new_edge_num - old_edge_num) << std::endl;
    std::cout << "Density: " << (float)(size_t)(new_edge_num - V) / V * 100 << " %" << std::endl;
    const uint64_t na = dyg_dual_added_count(d);
    std::cout << "Size of added edges: " << na << std::endl;
    std::vector<uint32_t> au(na ? na : 1), av(na ? na : 1);
    dyg_dual_added_edges(d, au.data(), av.data());
    {
        FILE* f = std::fopen(output_file.c_str(), "wb");
        if (f) {
            for (uint64_t i = 0; i < na; ++i) std::fprintf(f, "%u %u\n", au[i] + 1, av[i] + 1);
            std::fclose(f);
        }
    }
    std::cout << "Added edges saved to: " << output_file << std::endl;
    std::cout << "Random walk data deleted" << std::endl;
    dyg_dual_free(d);
    dyg_ctx_destroy(ctx);
    std::cout << "Decremental graph sparsification completed successfully!" << std::endl;
    dyg_mtx_free(del);
    dyg_mtx_free(sparse);
    std::cout << "CSRGraph deleted" << std::endl;
    dyg_mtx_free(dense);
    std::cout << "CSRGraph deleted" << std::endl;
    return 0;
}
