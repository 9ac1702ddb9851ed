// philox.h -- the counter-based draws shared by the walk kernels and the
// one-workgroup stream kernels (device code only).
//
// Draws come from Philox4x32-10 (Salmon et al., "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011) keyed by the 64-bit seed; the counter of block
// `blk` of a walker's stream is (blk, walker, id low word, tag << 24 | id
// bits 32..55), so every (tag, edge id, walker) owns its own stream.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"  // stream tags

namespace dyg {

constexpr uint32_t kM0 = 0xD2511F53u, kM1 = 0xCD9E8D57u;
constexpr uint32_t kW0 = 0x9E3779B9u, kW1 = 0xBB67AE85u;

// Philox4x32-10 block `blk` of the stream (c1, c2, c3) under key (k0, k1):
// counter (blk, c1, c2, c3).
__device__ __forceinline__ uint4 philox_block(uint32_t blk, uint32_t c1, uint32_t c2,
                                              uint32_t c3, uint32_t k0, uint32_t k1) {
    uint32_t x0 = blk, x1 = c1, x2 = c2, x3 = c3, q0 = k0, q1 = k1;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        // one 32x32->64 multiply gives both halves (a single v_mad_u64_u32)
        const uint64_t p0 = (uint64_t)kM0 * x0, p1 = (uint64_t)kM1 * x2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        x0 = hi1 ^ x1 ^ q0;
        x1 = lo1;
        x2 = hi0 ^ x3 ^ q1;
        x3 = lo0;
        q0 += kW0;
        q1 += kW1;
    }
    return make_uint4(x0, x1, x2, x3);
}

__device__ __forceinline__ uint32_t stream_c3(uint32_t tag, uint64_t id) {
    return (tag << 24) | ((uint32_t)(id >> 32) & 0xFFFFFFu);
}

// Heuristic two-hop walk from u (random_walk_decremental/main.cu:243-290):
// first hop uniform over u's row, second hop over n1's row, non-backtracking
// (a draw of u is redirected to that row's last slot).  Returns (u, n1, n2);
// (u, u, u) for an isolated u.  Uses block 0 of the walker's stream.  nodesc:
// n1's row comes from rows[] (the graph's slots carry no descriptors).
__device__ __forceinline__ uint3 heuristic_walk(const uint2* __restrict__ rows,
                                                const uint4* __restrict__ adj, uint32_t u,
                                                uint64_t key, uint32_t walker, uint64_t seed,
                                                bool nodesc) {
    const uint4 d = philox_block(0u, walker, (uint32_t)key, stream_c3(kTagHeuristic, key),
                                 (uint32_t)seed, (uint32_t)(seed >> 32));
    uint32_t n1 = u, n2 = u;
    const uint2 ru = rows[u];
    if (ru.y != 0) {
        const uint4 s1 = adj[ru.x + d.x % ru.y];
        n1 = s1.x;
        const uint2 r1 = nodesc ? rows[n1] : make_uint2(s1.z, s1.w);  // r1.y >= 1
        n2 = adj[r1.x + d.y % r1.y].x;
        if (n2 == u) n2 = adj[r1.x + r1.y - 1].x;
    }
    return make_uint3(u, n1, n2);
}

}  // namespace dyg
