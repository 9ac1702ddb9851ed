// incremental_update -- drop-in for random_walk_incremental/main.cu:240-362.
//
//   ./incremental_update <graph name> <distortion threshold> <# of walkers>
//
// Reads ../dataset/<graph>/adj_sparse.mtx (0-based) and ext.mtx (1-based),
// runs the snapshot-batched insertion walks on the GPU and writes
// ../dataset/<graph>/updated_adj_sparse.mtx (original lines, then the added
// edges in decision order).  Standard output carries the reference's lines.
// Extensions (environment, defaults equal the reference's behaviour):
//   DYGRASS_SEED    Philox key of the draw streams (default 0)
//   DYGRASS_DEVICE  HIP device index (default 0)
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
    if (argc != 4) {
        std::cout << "Input: ./incremental_update <1.graph name> <2.distortion_threshold> <3.# of "
                     "walkers>\n";
        std::exit(0);
    }
    const std::string name = argv[1];
    const std::string graph_file = "../dataset/" + name + "/adj_sparse.mtx";
    const std::string target_file = "../dataset/" + name + "/ext.mtx";
    const std::string output_file = "../dataset/" + name + "/updated_adj_sparse.mtx";
    const float distortion = (float)std::atof(argv[2]);
    const int n_walkers = std::atoi(argv[3]);

    dyg_mtx g = nullptr, t = nullptr;
    dyg_status st = dyg_mtx_read_graph(graph_file.c_str(), 0, 1, &g);
    if (st == DYG_ERR_IO) {
        std::perror("Error opening file");
        std::exit(-1);
    }
    if (st != DYG_OK) fail("Error reading " + graph_file, nullptr);
    st = dyg_mtx_read_edges(target_file.c_str(), 1, &t);
    if (st == DYG_ERR_IO) {
        std::perror("File open error");
        std::exit(-1);
    }
    if (st != DYG_OK) fail("Error reading " + target_file, nullptr);
    std::cout << "after read_targets" << std::endl;
    std::cout << "CSRGraph deleted" << std::endl;

    const uint32_t V = dyg_mtx_nv(g);
    const uint64_t M = dyg_mtx_nlines(t);
    const uint32_t* src = dyg_mtx_line_u(t);
    const uint32_t* dst = dyg_mtx_line_v(t);
    const float* wts = dyg_mtx_line_w(t);
    for (uint64_t i = 0; i < M; ++i)
        if (src[i] >= V || dst[i] >= V) {
            std::cout << "ext.mtx names a vertex outside the sparsifier" << std::endl;
            std::exit(EXIT_FAILURE);
        }
    if (n_walkers < 1 || n_walkers > 1024 || (n_walkers & (n_walkers - 1)) != 0) {
        std::cout << "# of walkers must be a power of two in [1,1024]" << std::endl;
        std::exit(EXIT_FAILURE);
    }

    dyg_ctx ctx = nullptr;
    if (dyg_ctx_create((int)env_u64("DYGRASS_DEVICE", 0), &ctx) != DYG_OK)
        fail("HIP context", nullptr);
    dyg_graph H = nullptr;
    if (dyg_graph_upload(ctx, V, dyg_mtx_row_ptr(g), dyg_mtx_col(g), dyg_mtx_w(g), 2, 0, &H) !=
        DYG_OK)
        fail("graph upload", ctx);
    std::cout << "Graph copied to GPU" << std::endl;

    dyg_walk_params p{};
    p.walkers = (uint32_t)n_walkers;
    p.max_steps = 100;
    p.threshold = distortion;
    // only found / not found is read below (as the reference's host reads only
    // flag == -1, random_walk_incremental/gpu_graph.cuh:302): an edge's walkers
    // may stop once one of them found the target
    p.flags = DYG_WALK_DECISION_ONLY;
    p.seed = env_u64("DYGRASS_SEED", 0);

    std::vector<int32_t> found(M ? M : 1);
    const auto start = std::chrono::high_resolution_clock::now();
    dyg_run_stats rs{};
    if (dyg_incremental_update(H, src, dst, wts, (uint32_t)M, 10, &p, found.data(), &rs) != DYG_OK)
        fail("incremental update", ctx);
    const auto end = std::chrono::high_resolution_clock::now();

    // per-batch lines: cumulative no-path count after each snapshot batch
    const uint64_t I = (M + 9) / 10;
    uint64_t no_path = 0, lines_out = 0;
    std::vector<uint32_t> au, av;
    std::vector<float> aw;
    for (uint64_t b0 = 0; b0 < M; b0 += I) {
        const uint64_t b1 = std::min(M, b0 + I);
        for (uint64_t i = b0; i < b1; ++i)
            if (found[i] == -1) {
                ++no_path;
                au.push_back(src[i]);
                av.push_back(dst[i]);
                aw.push_back(wts[i]);
            }
        std::cout << "no path count: " << no_path << std::endl;
        ++lines_out;
    }
    // the reference's loop launches one empty batch when M <= 1
    // (random_walk_incremental/main.cu:314-329, gpu_graph.cuh:399-405)
    for (; lines_out < 2 && M <= 1; ++lines_out) std::cout << "no path count: " << no_path << std::endl;

    const auto duration = std::chrono::duration_cast<std::chrono::milliseconds>(end - start);
    std::cout << "Time taken: " << duration.count() << " milliseconds" << std::endl;
    std::cout << "No path %: " << (float)no_path / M * 100 << std::endl;
    std::cout << "Edges added: " << no_path << " out of " << M << " samples" << std::endl;
    std::cout << "saving mtx to " << output_file << std::endl;
    // original lines in file order, then the added edges in decision order
    if (dyg_mtx_write_weighted(output_file.c_str(), dyg_mtx_line_u(g), dyg_mtx_line_v(g),
                               dyg_mtx_line_w(g), dyg_mtx_nlines(g), 0) != DYG_OK ||
        dyg_mtx_write_weighted(output_file.c_str(), au.data(), av.data(), aw.data(), au.size(),
                               1) != DYG_OK) {
        std::cerr << "Unable to open file" << std::endl;
    }
    std::cout << "Updated graph saved to " << output_file << std::endl;
    std::cout << "Random walk data deleted" << std::endl;
    dyg_graph_free(H);
    dyg_ctx_destroy(ctx);
    std::cout << "Incremental graph sparsification completed successfully!" << std::endl;
    std::cout << "gpu_graph" << std::endl << "gpu_graph deleted" << std::endl;
    dyg_mtx_free(t);
    std::cout << "Targets deleted" << std::endl;
    dyg_mtx_free(g);
    std::cout << "CSRGraph deleted" << std::endl;
    return 0;
}
