// Raycaster (MapRenderingDelegate, ray_tracing.cu:255-401) and the
// rendering-block z-range maps (ray_tracing.cu:13-234) for gfx950.
//
// One lane per pixel, one 64-lane wave per 8x8 pixel tile. Consecutive
// workgroups are spread over the 8 XCDs by the dispatcher, so the tile index
// is remapped to give each XCD a contiguous band of image rows: rays of one
// band traverse the same voxel blocks and share that XCD's L2.
//
// The per-pixel march is raycast_pixel (vm_dev_ray.h).
#include <algorithm>
#include <cmath>

#include "vm_common.h"
#include "vm_dev_ray.h"
#include "vm_launch.h"

namespace vmamd {

namespace {

// the runs the single-map raycast takes (RunKinds, vm_dev_ray.h) and its
// register budget: a 640x480 image is 4800 waves, under five per SIMD on 256
// compute units, so six waves per SIMD keep every wave resident
constexpr int kRaycastRuns = kRunsInvalid;
constexpr int kRaycastWavesPerEu = 6;

template <int FMT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(kRaycastWavesPerEu))) void k_raycast(MapView m, float* vmap, int64_t pitch, int W, int H,
                                                Intrinsics K, Affine pose, Affine Tinv, float vsInv, float rtStep,
                                                int gate, int tilesX, int tilesPerXcd, int nTiles,
                                                RayBounds rb, float margin) {
    const int g = blockIdx.x;
    const int t = (g & 7) * tilesPerXcd + (g >> 3);
    if (t >= nTiles) return;
    const int x = (t % tilesX) * 8 + (threadIdx.x & 7);
    const int y = (t / tilesX) * 8 + (threadIdx.x >> 3);
    if (x >= W || y >= H) return;
    if (gate && m.ctr[kVisibleCount] == 0) return;
    float nearLim = 0.0f, farLim = INFINITY;
    if (rb.keys) ray_bounds_limits(rb, bounds_tile(x, y, W), margin, nearLim, farLim);
    float* out = (float*)((uint8_t*)vmap + (int64_t)y * pitch) + x * 4;
    raycast_pixel<FMT, kRaycastRuns>(m, out, x, y, K, pose, Tinv, vsInv, rtStep, nearLim, farLim);
}

// Ray-distance bounds of every tile (ray_bounds_chunk), kBoundsChunk
// allocated blocks per pass of a 256-thread workgroup.
__global__ __launch_bounds__(256) void k_ray_bounds(MapView m, RayBounds rb, int W, int H, Intrinsics K, Affine Tinv,
                                                    float3 c, float vs, float grow) {
    __shared__ BoundsShared<kBoundsChunk> sh;
    const int A = allocated_count(m.ctr, m.nBlock);
    for (int base = blockIdx.x * kBoundsChunk; base < A; base += gridDim.x * kBoundsChunk)
        ray_bounds_chunk<kBoundsChunk>(m, rb, base, A, W, H, K, Tinv, c, vs, grow, sh);
}

// RenderingBlockDelegate::createRenderingBlock (ray_tracing.cu:30-80) per
// visible entry, then the fill of FillRenderingBlockFunctor (:145-168) done
// directly over the block's rectangle (the 16x16 tiling only partitions it).
// z values are positive, so float min/max order equals int order.
__global__ void k_rendering_blocks(MapView m, float* zx, float* zy, int64_t pitch, int cols8, int rows8,
                                   Intrinsics K, Affine Tinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.ctr[kVisibleCount]) return;
    const HashEntry e = m.visible[i];
    if (e.ptr == -1) return;
    const float scale = m.voxelSize * 8.0f;
    int ulx = cols8, uly = rows8, lrx = -1, lry = -1;
    float z0 = m.depthMax, z1 = m.depthMin;
    for (int c = 0; c < 8; ++c) {
        float3 w = make_float3((float)(e.x + (c & 1)) * scale, (float)(e.y + ((c >> 1) & 1)) * scale,
                               (float)(e.z + ((c >> 2) & 1)) * scale);
        float3 p = xform(Tinv, w);
        const float px = (K.fx * p.x / p.z + K.cx) / 8.0f, py = (K.fy * p.y / p.z + K.cy) / 8.0f;
        if ((float)ulx > floorf(px)) ulx = (short)f2i_sat(floorf(px));
        if ((float)lrx < ceilf(px)) lrx = (short)f2i_sat(ceilf(px));
        if ((float)uly > floorf(py)) uly = (short)f2i_sat(floorf(py));
        if ((float)lry < ceilf(py)) lry = (short)f2i_sat(ceilf(py));
        if (z0 > p.z) z0 = p.z;
        if (z1 < p.z) z1 = p.z;
    }
    if (ulx < 0) ulx = 0;
    if (uly < 0) uly = 0;
    if (lrx >= cols8) lrx = cols8 - 1;
    if (lry >= rows8) lry = rows8 - 1;
    if (ulx > lrx || uly > lry) return;
    if (z0 < m.depthMin) z0 = m.depthMin;
    if (z1 < m.depthMin) return;
    const int nx = (int)ceilf((float)(lrx - ulx + 1) / 16.f);
    const int ny = (int)ceilf((float)(lry - uly + 1) / 16.f);
    const int old = atomicAdd(&m.ctr[kRenderBlockCount], nx * ny);
    if (old + nx * ny >= 1000000) {  // MaxNumRenderingBlock, ray_tracing.cu:8
        atomicSub(&m.ctr[kRenderBlockCount], nx * ny);
        return;
    }
    for (int yy = uly; yy <= lry; ++yy) {
        int* rx = (int*)((uint8_t*)zx + (int64_t)yy * pitch);
        int* ry = (int*)((uint8_t*)zy + (int64_t)yy * pitch);
        for (int xx = ulx; xx <= lrx; ++xx) {
            atomicMin(rx + xx, __float_as_int(z0));
            atomicMax(ry + xx, __float_as_int(z1));
        }
    }
}

__global__ void k_zrange_init(float* zx, float* zy, int64_t pitch, int cols8, int rows8, int* ctr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) ctr[kRenderBlockCount] = 0;
    if (i >= cols8 * rows8) return;
    const int yy = i / cols8, xx = i % cols8;
    ((float*)((uint8_t*)zx + (int64_t)yy * pitch))[xx] = 100.f;  // setTo(100), ray_tracing.cu:188
    ((float*)((uint8_t*)zy + (int64_t)yy * pitch))[xx] = 0.f;
}

}  // namespace

hipError_t launch_ray_bounds(const MapView& m, const RayBounds& rb, int W, int H, const Intrinsics& K,
                             const Affine& pose, const Affine& Tinv, float grow, hipStream_t s) {
    if (!rb.keys) return hipSuccess;  // no bounds: every ray marches in full
    const float vsInv = (float)(1.0 / (double)m.voxelSize);
    const float3 c = make_float3(pose.t0 * vsInv, pose.t1 * vsInv, pose.t2 * vsInv);
    const int need = (m.nBlock + kBoundsChunk - 1) / kBoundsChunk, g = num_compute_units() * 8;
    hipLaunchKernelGGL(k_ray_bounds, dim3(need < g ? need : g), dim3(256), 0, s, m, rb, W, H, K, Tinv, c,
                       m.voxelSize, grow);
    return hipGetLastError();
}

hipError_t launch_raycast(const MapView& m, float* vmap, int64_t pitch, int W, int H, const Intrinsics& K,
                          const Affine& pose, const Affine& Tinv, bool gate, hipStream_t s, const RayBounds* bounds,
                          bool boundsReady) {
    const int tilesX = (W + 7) / 8, tilesY = (H + 7) / 8;
    const int nTiles = tilesX * tilesY;
    const int perXcd = (nTiles + 7) / 8;
    const dim3 grid(perXcd * 8);
    const float vsInv = (float)(1.0 / (double)m.voxelSize);  // ray_tracing.cu:430
    const float rtStep = m.trunc / m.voxelSize;               // :431
    float margin = 0.0f;
    RayBounds rb{nullptr, nTiles, 0u};
    if (bounds) {
        rb = *bounds;
        margin = ray_drift_margin(W, H, K, pose.t0, pose.t1, pose.t2, vsInv);
        if (!boundsReady) {
            hipError_t e = launch_ray_bounds(m, rb, W, H, K, pose, Tinv, margin, s);
            if (e != hipSuccess) return e;
        }
    }
    switch (m.fmt) {
        case kFmtRef:
            hipLaunchKernelGGL(k_raycast<kFmtRef>, grid, dim3(64), 0, s, m, vmap, pitch, W, H, K, pose, Tinv,
                               vsInv, rtStep, (int)gate, tilesX, perXcd, nTiles, rb, margin);
            break;
        case kFmtRefRGB:
            hipLaunchKernelGGL(k_raycast<kFmtRefRGB>, grid, dim3(64), 0, s, m, vmap, pitch, W, H, K, pose, Tinv,
                               vsInv, rtStep, (int)gate, tilesX, perXcd, nTiles, rb, margin);
            break;
        default:
            hipLaunchKernelGGL(k_raycast<kFmtF32>, grid, dim3(64), 0, s, m, vmap, pitch, W, H, K, pose, Tinv,
                               vsInv, rtStep, (int)gate, tilesX, perXcd, nTiles, rb, margin);
    }
    return hipGetLastError();
}

hipError_t launch_rendering_blocks(const MapView& m, float* zx, float* zy, int64_t pitch, int cols8, int rows8,
                                   const Intrinsics& K, const Affine& Tinv, hipStream_t s) {
    const int n = cols8 * rows8 > 1 ? cols8 * rows8 : 1;
    hipLaunchKernelGGL(k_zrange_init, dim3((n + 255) / 256), dim3(256), 0, s, zx, zy, pitch, cols8, rows8, m.ctr);
    const int maxVis = m.nEntry < m.nBlock ? m.nEntry : m.nBlock;
    hipLaunchKernelGGL(k_rendering_blocks, dim3((maxVis + 255) / 256), dim3(256), 0, s, m, zx, zy, pitch, cols8,
                       rows8, K, Tinv);
    return hipGetLastError();
}

}  // namespace vmamd
