// Marching cubes (BuildVertexArray / Polygonize, marching_cubes.cu:9-305),
// triangle centroids (:314-344), vertex-map normals and Phong shading
// (image_proc.cu:8-104) for gfx950.
//
// Marching cubes: one 64-lane workgroup per allocated block. The block's
// 8x8x8 voxels plus the +x/+y/+z face layers of its 7 neighbours (a 9^3
// halo) are read once into LDS, with one hash lookup per neighbour instead of
// 8 lookups per voxel; the case tables are LDS-resident too. Each lane owns an (x, y) column of 8 cubes, counts its
// triangles, a wave prefix sum gives each lane its output range, and one
// atomic per block reserves the block's range in the triangle buffer. Output
// order differs from the reference (which is already nondeterministic);
// triangle contents, winding and normals follow :209-242 exactly.
#include "vm_common.h"
#include "vm_dev_ray.h"
#include "vm_launch.h"
#include "vm_mc_table.h"

namespace vmamd {

namespace {

__constant__ McTables c_mc = kMcTables;

// interpolateLinear, marching_cubes.cu:106-115 (the 1e-6 guards compare in double)
__device__ inline float interp_lin(float v1, float v2) {
    if ((double)fabsf(0.0f - v1) < 1e-6) return 0.f;
    if ((double)fabsf(0.0f - v2) < 1e-6) return 1.f;
    if ((double)fabsf(v1 - v2) < 1e-6) return 0.f;
    return (0.0f - v1) / (v2 - v1);
}

// make_vertex, marching_cubes.cu:145-204: vertex of edge e of the cube at pos
__device__ inline float3 edge_vertex(int e, float3 p, const float* s) {
    float t;
    switch (e) {
        case 0: t = interp_lin(s[0], s[1]); return make_float3(p.x + t, p.y + 0.f, p.z + 0.f);
        case 1: t = interp_lin(s[1], s[2]); return make_float3(p.x + 1.f, p.y + t, p.z + 0.f);
        case 2: t = interp_lin(s[2], s[3]); return make_float3(p.x + (1.f - t), p.y + 1.f, p.z + 0.f);
        case 3: t = interp_lin(s[3], s[0]); return make_float3(p.x + 0.f, p.y + (1.f - t), p.z + 0.f);
        case 4: t = interp_lin(s[4], s[5]); return make_float3(p.x + t, p.y + 0.f, p.z + 1.f);
        case 5: t = interp_lin(s[5], s[6]); return make_float3(p.x + 1.f, p.y + t, p.z + 1.f);
        case 6: t = interp_lin(s[6], s[7]); return make_float3(p.x + (1.f - t), p.y + 1.f, p.z + 1.f);
        case 7: t = interp_lin(s[7], s[4]); return make_float3(p.x + 0.f, p.y + (1.f - t), p.z + 1.f);
        case 8: t = interp_lin(s[0], s[4]); return make_float3(p.x + 0.f, p.y + 0.f, p.z + t);
        case 9: t = interp_lin(s[1], s[5]); return make_float3(p.x + 1.f, p.y + 0.f, p.z + t);
        case 10: t = interp_lin(s[2], s[6]); return make_float3(p.x + 1.f, p.y + 1.f, p.z + t);
        default: t = interp_lin(s[3], s[7]); return make_float3(p.x + 0.f, p.y + 1.f, p.z + t);
    }
}

constexpr int kHalo = 9 * 9 * 9;

template <int FMT>
__global__ __launch_bounds__(64) void k_polygonize(MapView m, float* verts, float* norms, int64_t cap) {
    __shared__ float s_sdf[kHalo];
    __shared__ uint8_t s_ok[kHalo];
    __shared__ int s_nb[8];
    // the case tables, read with per-lane case indices, live in LDS for the
    // workgroup's whole life (2.5 KB, loaded once)
    __shared__ uint64_t s_tri[256];
    __shared__ uint16_t s_edges[256];
    const int lane = threadIdx.x;
    for (int i = lane; i < 256; i += 64) {
        s_tri[i] = c_mc.tri[i];
        s_edges[i] = c_mc.edges[i];
    }
    __syncthreads();
    const int A = allocated_count(m.ctr, m.nBlock);
    const float vs = m.voxelSize;
    for (int b = blockIdx.x; b < A; b += gridDim.x) {
        const int4 key = m.blockPos[b];
        if (key.x == kEmptyKey) continue;
        if (lane < 8) {
            const int dx = lane & 1, dy = (lane >> 1) & 1, dz = (lane >> 2) & 1;
            s_nb[lane] = lane == 0 ? b * kBlockVox : find_block(m, key.x + dx, key.y + dy, key.z + dz);
        }
        __syncthreads();
        for (int i = lane; i < kHalo; i += 64) {
            const int hx = i % 9, hy = (i / 9) % 9, hz = i / 81;
            const int nb = (hx >> 3) | ((hy >> 3) << 1) | ((hz >> 3) << 2);
            const int p = s_nb[nb];
            float s = 1.f;
            bool ok = false;
            if (p >= 0) Vox<FMT>::read(m.voxels, (int64_t)p + local_index(hx & 7, hy & 7, hz & 7), s, ok);
            s_sdf[i] = s;
            s_ok[i] = ok;
        }
        __syncthreads();
        const int lx = lane & 7, ly = lane >> 3;
        // corner offsets in Bourke order (read_sdf_list, marching_cubes.cu:68-104)
        const int cofs[8] = {0, 1, 1 + 9, 9, 81, 1 + 81, 1 + 9 + 81, 9 + 81};
        int cubes[8];
        int ntri = 0;
#pragma unroll
        for (int z = 0; z < 8; ++z) {
            const int h0 = lx + 9 * ly + 81 * z;
            int cube = 0;
            bool all = true;
#pragma unroll
            for (int c = 0; c < 8; ++c) {
                all = all && s_ok[h0 + cofs[c]];
                if (s_sdf[h0 + cofs[c]] < 0.f) cube |= 1 << c;
            }
            if (!all || s_edges[cube] == 0) cube = 0;
            cubes[z] = cube;
            ntri += (int)(s_tri[cube] & 15u);
        }
        // wave-inclusive scan of the per-lane triangle counts
        int incl = ntri;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        const int total = __shfl(incl, 63, 64);
        int base = 0;
        if (total > 0) {
            if (lane == 0) base = atomicAdd((unsigned int*)&m.ctr[kTriangleCount], (unsigned)total);
            base = __shfl(base, 0, 64);
        }
        int64_t id = (int64_t)(unsigned)base + (incl - ntri);
        if (ntri > 0) {
            for (int z = 0; z < 8; ++z) {
                const int cube = cubes[z];
                const uint64_t tri = s_tri[cube];
                const int nt = (int)(tri & 15u);
                if (nt == 0) continue;
                const int h0 = lx + 9 * ly + 81 * z;
                float s[8];
#pragma unroll
                for (int c = 0; c < 8; ++c) s[c] = s_sdf[h0 + cofs[c]];
                const float3 pos = make_float3((float)(key.x * 8 + lx), (float)(key.y * 8 + ly), (float)(key.z * 8 + z));
                for (int k = 0; k < nt; ++k, ++id) {
                    if (id >= cap) continue;
                    const int ea = (int)((tri >> (4 + 12 * k)) & 15u);
                    const int eb = (int)((tri >> (8 + 12 * k)) & 15u);
                    const int ec = (int)((tri >> (12 + 12 * k)) & 15u);
                    const float3 va = edge_vertex(ea, pos, s), vb = edge_vertex(eb, pos, s),
                                 vc = edge_vertex(ec, pos, s);
                    // stored reversed: (c, b, a), scaled to metres
                    const float3 p0 = make_float3(vc.x * vs, vc.y * vs, vc.z * vs);
                    const float3 p1 = make_float3(vb.x * vs, vb.y * vs, vb.z * vs);
                    const float3 p2 = make_float3(va.x * vs, va.y * vs, va.z * vs);
                    const float3 n = normalized3(cross3(make_float3(p1.x - p0.x, p1.y - p0.y, p1.z - p0.z),
                                                        make_float3(p2.x - p0.x, p2.y - p0.y, p2.z - p0.z)));
                    float* o = verts + id * 9;
                    o[0] = p0.x; o[1] = p0.y; o[2] = p0.z;
                    o[3] = p1.x; o[4] = p1.y; o[5] = p1.z;
                    o[6] = p2.x; o[7] = p2.y; o[8] = p2.z;
                    float* q = norms + id * 9;
                    q[0] = n.x; q[1] = n.y; q[2] = n.z;
                    q[3] = n.x; q[4] = n.y; q[5] = n.z;
                    q[6] = n.x; q[7] = n.y; q[8] = n.z;
                }
            }
        }
        __syncthreads();
    }
}

__global__ void k_surface_points(const float* v, float* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* a = v + i * 9;
#pragma unroll
    for (int j = 0; j < 3; ++j) p[i * 3 + j] = ((a[j] + a[3 + j]) + a[6 + j]) / 3.0f;
}

// computeNormalKernel, image_proc.cu:74-93
__global__ void k_normals(const float* vmap, int64_t vpitch, float* nmap, int64_t npitch, int W, int H) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W - 1 || y >= H - 1) return;
    normal_pixel(vmap, vpitch, nmap, npitch, W, H, x, y);
}

__device__ inline float saturate01(float v) { return v != v ? 0.f : fminf(1.f, fmaxf(0.f, v)); }
__device__ inline float dot3(float3 a, float3 b) { return sum3(a.x * b.x, a.y * b.y, a.z * b.z); }

// renderPoint / renderSceneKernel, image_proc.cu:8-61; light at (5, 5, 5)
__global__ void k_shading(const float* vmap, int64_t vpitch, const float* nmap, int64_t npitch, uint8_t* rgba,
                          int64_t rpitch, int W, int H) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= W || y >= H) return;
    const float4 P4 = *((const float4*)((const uint8_t*)vmap + (int64_t)y * vpitch) + x);
    const float4 N4 = *((const float4*)((const uint8_t*)nmap + (int64_t)y * npitch) + x);
    const float3 P = make_float3(P4.x, P4.y, P4.z), N = make_float3(N4.x, N4.y, N4.z);
    float3 col = make_float3(4.f / 255.f, 2.f / 255.f, 2.f / 255.f);
    if (!isnan(P.x)) {
        const float Ka = 0.3f, Kd = 0.5f, Ks = 0.2f, n = 20.f, Ax = 1.f, Dx = 1.f, Sx = 1.f, Lx = 1.f;
        const float3 L = normalized3(make_float3(5.f - P.x, 5.f - P.y, 5.f - P.z));
        const float3 V = normalized3(make_float3(0.f - P.x, 0.f - P.y, 0.f - P.z));
        const float nl = dot3(N, L);
        const float3 R = normalized3(make_float3(2.f * N.x * nl - L.x, 2.f * N.y * nl - L.y, 2.f * N.z * nl - L.z));
        const float I = Ax * Ka * Dx + Lx * Kd * Dx * fmaxf(0.f, nl) + Lx * Ks * Sx * powf(fmaxf(0.f, dot3(R, V)), n);
        col = make_float3(I, I, I);
    }
    uchar4 o;
    o.x = (uint8_t)(saturate01(col.x) * 255.f);
    o.y = (uint8_t)(saturate01(col.y) * 255.f);
    o.z = (uint8_t)(saturate01(col.z) * 255.f);
    o.w = 255;
    *((uchar4*)(rgba + (int64_t)y * rpitch) + x) = o;
}

}  // namespace

hipError_t launch_polygonize(const MapView& m, float* verts, float* norms, int64_t cap, hipStream_t s) {
    hipError_t e = launch_fill_u32(m.ctr + kTriangleCount, 0u, 1, s);
    if (e != hipSuccess) return e;
    int g = num_compute_units() * 16;
    if (g > m.nBlock) g = m.nBlock > 0 ? m.nBlock : 1;
    switch (m.fmt) {
        case kFmtRef: hipLaunchKernelGGL(k_polygonize<kFmtRef>, dim3(g), dim3(64), 0, s, m, verts, norms, cap); break;
        case kFmtRefRGB:
            hipLaunchKernelGGL(k_polygonize<kFmtRefRGB>, dim3(g), dim3(64), 0, s, m, verts, norms, cap);
            break;
        default: hipLaunchKernelGGL(k_polygonize<kFmtF32>, dim3(g), dim3(64), 0, s, m, verts, norms, cap);
    }
    return hipGetLastError();
}

hipError_t launch_surface_points(const float* verts, float* points, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_surface_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, verts, points, n);
    return hipGetLastError();
}

hipError_t launch_normals(const float* vmap, int64_t vpitch, float* nmap, int64_t npitch, int W, int H,
                          hipStream_t s) {
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    hipLaunchKernelGGL(k_normals, grid, dim3(256), 0, s, vmap, vpitch, nmap, npitch, W, H);
    return hipGetLastError();
}

hipError_t launch_shading(const float* vmap, int64_t vpitch, const float* nmap, int64_t npitch, uint8_t* rgba,
                          int64_t rpitch, int W, int H, hipStream_t s) {
    dim3 grid((W + 15) / 16, (H + 15) / 16);
    hipLaunchKernelGGL(k_shading, grid, dim3(256), 0, s, vmap, vpitch, nmap, npitch, rgba, rpitch, W, H);
    return hipGetLastError();
}

}  // namespace vmamd
