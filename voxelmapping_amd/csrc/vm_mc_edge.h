// Edge vertices of a marching cube without a per-edge branch
// (make_vertex, marching_cubes.cu:145-204).
//
// Edge e runs from corner a(e) to corner b(e) (Bourke numbering); the vertex
// is p + (ox, oy, oz) with each offset one of 0, 1, t or 1 - t, where
// t = interpolateLinear(sdf[a], sdf[b]) (:106-115). The same float
// operations as the per-edge cases, in the same order, so results are
// bit-identical; lanes holding different edges no longer diverge.
#pragma once
#include <cmath>
#include <cstdint>

namespace vmamd {

// interpolateLinear, marching_cubes.cu:106-115 (the 1e-6 guards compare in double)
__host__ __device__ inline float interp_lin(float v1, float v2) {
    if ((double)fabsf(0.0f - v1) < 1e-6) return 0.f;
    if ((double)fabsf(0.0f - v2) < 1e-6) return 1.f;
    if ((double)fabsf(v1 - v2) < 1e-6) return 0.f;
    return (0.0f - v1) / (v2 - v1);
}

// descriptor: corner a (3 bits) | corner b << 3 | x, y, z offset codes
// (2 bits each, 0: 0, 1: 1, 2: t, 3: 1 - t) << 6, 8, 10
constexpr uint64_t mc_edge_desc(int a, int b, int xc, int yc, int zc) {
    return (uint64_t)(a | b << 3 | xc << 6 | yc << 8 | zc << 10);
}
// four 16-bit descriptors per word: edges 0-3, 4-7, 8-11
constexpr uint64_t kMcEdge0 = mc_edge_desc(0, 1, 2, 0, 0) | mc_edge_desc(1, 2, 1, 2, 0) << 16 |
                              mc_edge_desc(2, 3, 3, 1, 0) << 32 | mc_edge_desc(3, 0, 0, 3, 0) << 48;
constexpr uint64_t kMcEdge1 = mc_edge_desc(4, 5, 2, 0, 1) | mc_edge_desc(5, 6, 1, 2, 1) << 16 |
                              mc_edge_desc(6, 7, 3, 1, 1) << 32 | mc_edge_desc(7, 4, 0, 3, 1) << 48;
constexpr uint64_t kMcEdge2 = mc_edge_desc(0, 4, 0, 0, 2) | mc_edge_desc(1, 5, 1, 0, 2) << 16 |
                              mc_edge_desc(2, 6, 1, 1, 2) << 32 | mc_edge_desc(3, 7, 0, 1, 2) << 48;

// offset of corner c in a 9x9x9 halo (x fastest): c = 0..7 is
// (0,0,0) (1,0,0) (1,1,0) (0,1,0) (0,0,1) (1,0,1) (1,1,1) (0,1,1)
__host__ __device__ inline int mc_corner_offset(int c) {
    return ((c ^ (c >> 1)) & 1) + 9 * ((c >> 1) & 1) + 81 * (c >> 2);
}

__host__ __device__ inline float mc_edge_offset(int code, float t) {
    return code == 2 ? t : code == 3 ? 1.f - t : (float)code;
}

// vertex of edge e of the cube at p, whose corner 0 is hs[0] in the halo
__host__ __device__ inline float3 mc_edge_vertex(int e, float3 p, const float* hs) {
    const uint64_t w = e < 4 ? kMcEdge0 : e < 8 ? kMcEdge1 : kMcEdge2;
    const int d = (int)((w >> (16 * (e & 3))) & 0xffffu);
    const float t = interp_lin(hs[mc_corner_offset(d & 7)], hs[mc_corner_offset((d >> 3) & 7)]);
    return make_float3(p.x + mc_edge_offset((d >> 6) & 3, t), p.y + mc_edge_offset((d >> 8) & 3, t),
                       p.z + mc_edge_offset((d >> 10) & 3, t));
}

}  // namespace vmamd
