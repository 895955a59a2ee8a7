// Raycaster (MapRenderingDelegate, ray_tracing.cu:255-401) and the
// rendering-block z-range maps (ray_tracing.cu:13-234) for gfx950.
//
// One lane per pixel, one 64-lane wave per 8x8 pixel tile. Consecutive
// workgroups are spread over the 8 XCDs by the dispatcher, so the tile index
// is remapped to give each XCD a contiguous band of image rows: rays of one
// band traverse the same voxel blocks and share that XCD's L2.
//
// Every ray runs the reference's march step for step (same sample positions,
// same reads, same float operations). Each lane keeps the last two voxel
// blocks it resolved, so a step that stays inside a block costs no hash
// probe; that changes the number of memory accesses, not any value read.
#include "vm_common.h"
#include "vm_launch.h"

namespace vmamd {

namespace {

// findEntry, map_struct.h:356-374 (read-only walk, bounded)
__device__ inline int find_block(const MapView& m, int bx, int by, int bz) {
    int slot = hash_block(bx, by, bz, m.nBucket, m.bucketMask);
    int guard = m.nEntry - m.nBucket + 1;
    while (true) {
        const HashEntry* e = m.table + slot;
        const int p = e->ptr, off = e->offset, x = e->x, y = e->y, z = e->z;
        if (p >= 0 && x == bx && y == by && z == bz) return p;
        if (off < 0 || --guard < 0) return -1;
        slot = m.nBucket + off - 1;
    }
}

template <int FMT>
struct RayReader {
    const MapView& m;
    int cx, cy, cz, cp;  // most recent block
    int px, py, pz, pp;  // the one before
    __device__ explicit RayReader(const MapView& mv)
        : m(mv), cx(INT_MIN), cy(INT_MIN), cz(INT_MIN), cp(-1), px(INT_MIN), py(INT_MIN), pz(INT_MIN), pp(-1) {}

    __device__ inline int block(int bx, int by, int bz) {
        if (bx == cx && by == cy && bz == cz) return cp;
        if (bx == px && by == py && bz == pz) {
            int t;
            t = cx; cx = px; px = t;
            t = cy; cy = py; py = t;
            t = cz; cz = pz; pz = t;
            t = cp; cp = pp; pp = t;
            return cp;
        }
        const int p = find_block(m, bx, by, bz);
        px = cx; py = cy; pz = cz; pp = cp;
        cx = bx; cy = by; cz = bz; cp = p;
        return p;
    }

    // voxel index of integer voxel (vx, vy, vz), -1 when its block is absent
    __device__ inline int64_t voxel(int vx, int vy, int vz) {
        const int p = block(vox_to_blk(vx), vox_to_blk(vy), vox_to_blk(vz));
        if (p < 0) return -1;
        return (int64_t)p + local_index(vox_to_loc(vx), vox_to_loc(vy), vox_to_loc(vz));
    }

    // read_sdf, ray_tracing.cu:272-286
    __device__ inline float read(float3 q, bool& valid) {
        const int64_t i = voxel(floor_i(q.x), floor_i(q.y), floor_i(q.z));
        float s = 1.0f;
        bool v = false;
        if (i >= 0) Vox<FMT>::read(m.voxels, i, s, v);
        valid = v;
        return v ? s : 1.0f;
    }

    // read_sdf_interped, ray_tracing.cu:288-321. floor(q + e_k) is computed
    // per axis exactly as the reference adds the unit offset before flooring.
    __device__ inline float interp(float3 q, bool& valid) {
        const float fx = q.x - floorf(q.x), fy = q.y - floorf(q.y), fz = q.z - floorf(q.z);
        const int X0 = floor_i(q.x), X1 = floor_i(q.x + 1.f);
        const int Y0 = floor_i(q.y), Y1 = floor_i(q.y + 1.f);
        const int Z0 = floor_i(q.z), Z1 = floor_i(q.z + 1.f);
        int64_t idx[8];
        idx[0] = voxel(X0, Y0, Z0);
        idx[1] = voxel(X1, Y0, Z0);
        idx[2] = voxel(X0, Y1, Z0);
        idx[3] = voxel(X1, Y1, Z0);
        idx[4] = voxel(X0, Y0, Z1);
        idx[5] = voxel(X1, Y0, Z1);
        idx[6] = voxel(X0, Y1, Z1);
        idx[7] = voxel(X1, Y1, Z1);
        float s[8];
        bool all = true;
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            float v = 1.0f;
            bool ok = false;
            if (idx[k] >= 0) Vox<FMT>::read(m.voxels, idx[k], v, ok);
            s[k] = ok ? v : 1.0f;
            all = all && ok;
        }
        float r0 = (1.0f - fx) * s[0] + fx * s[1];
        float r1 = (1.0f - fx) * s[2] + fx * s[3];
        const float r2 = (1.0f - fy) * r0 + fy * r1;
        r0 = (1.0f - fx) * s[4] + fx * s[5];
        r1 = (1.0f - fx) * s[6] + fx * s[7];
        const float r3 = (1.0f - fy) * r0 + fy * r1;
        valid = all;
        return (1.0f - fz) * r2 + fz * r3;
    }
};

__device__ inline float norm3(float3 v) { return sqrtf(sum3(v.x * v.x, v.y * v.y, v.z * v.z)); }

template <int FMT>
__global__ __launch_bounds__(64) void k_raycast(MapView m, float* vmap, int64_t pitch, int W, int H,
                                                Intrinsics K, Affine pose, Affine Tinv, int gate,
                                                int tilesX, int tilesPerXcd, int nTiles) {
    const int g = blockIdx.x;
    const int t = (g & 7) * tilesPerXcd + (g >> 3);
    if (t >= nTiles) return;
    const int x = (t % tilesX) * 8 + (threadIdx.x & 7);
    const int y = (t / tilesX) * 8 + (threadIdx.x >> 3);
    if (x >= W || y >= H) return;
    if (gate && m.ctr[kVisibleCount] == 0) return;
    float* out = (float*)((uint8_t*)vmap + (int64_t)y * pitch) + x * 4;
    out[0] = __int_as_float(0x7fffffff);

    const float vsInv = m.voxelSize > 0.f ? (float)(1.0 / (double)m.voxelSize) : 0.f;  // ray_tracing.cu:430
    const float rtStep = m.trunc / m.voxelSize;                                       // :431
    const float zNear = 0.1f, zFar = 5.0f;                                            // :340-341
    const float ux = K.ifx * ((float)x - K.cx), uy = K.ify * ((float)y - K.cy);
    float3 pt = make_float3(ux * zNear, uy * zNear, zNear);
    float dist_s = norm3(pt) * vsInv;
    float3 bs = xform(pose, pt);
    bs = make_float3(bs.x * vsInv, bs.y * vsInv, bs.z * vsInv);
    pt = make_float3(ux * zFar, uy * zFar, zFar);
    const float distEnd = norm3(pt) * vsInv;
    float3 be = xform(pose, pt);
    be = make_float3(be.x * vsInv, be.y * vsInv, be.z * vsInv);
    float3 dir = make_float3(be.x - bs.x, be.y - bs.y, be.z - bs.z);
    {
        // Eigen normalized(): divide by the norm when the squared norm is > 0
        const float n2 = sum3(dir.x * dir.x, dir.y * dir.y, dir.z * dir.z);
        if (n2 > 0.f) {
            const float n = sqrtf(n2);
            dir = make_float3(dir.x / n, dir.y / n, dir.z / n);
        }
    }
    float3 res = bs;
    RayReader<FMT> rd(m);
    float sdf = 1.0f, last;
    bool valid = false;
    float step;
    while (dist_s < distEnd) {
        last = sdf;
        sdf = rd.read(res, valid);
        if (sdf <= 0.5f && sdf >= -0.5f) sdf = rd.interp(res, valid);
        if (sdf <= 0.0f) break;
        if (sdf >= 0.f && last < 0.f) return;
        step = valid ? fmaxf(sdf * rtStep, 1.0f) : 1.0f;
        res = make_float3(res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z);
        dist_s += step;
    }
    if (sdf <= 0.0f) {
        step = sdf * rtStep;
        res = make_float3(res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z);
        sdf = rd.interp(res, valid);
        step = sdf * rtStep;
        res = make_float3(res.x + step * dir.x, res.y + step * dir.y, res.z + step * dir.z);
        if (valid) {
            const float3 w = make_float3(res.x / vsInv, res.y / vsInv, res.z / vsInv);
            const float3 c = xform(Tinv, w);
            *(float4*)out = make_float4(c.x, c.y, c.z, 1.0f);
        }
    }
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

hipError_t launch_raycast(const MapView& m, float* vmap, int64_t pitch, int W, int H, const Intrinsics& K,
                          const Affine& pose, const Affine& Tinv, bool gate, hipStream_t s) {
    const int tilesX = (W + 7) / 8, tilesY = (H + 7) / 8;
    const int nTiles = tilesX * tilesY;
    const int perXcd = (nTiles + 7) / 8;
    const dim3 grid(perXcd * 8);
    switch (m.fmt) {
        case kFmtRef:
            hipLaunchKernelGGL(k_raycast<kFmtRef>, grid, dim3(64), 0, s, m, vmap, pitch, W, H, K, pose, Tinv,
                               (int)gate, tilesX, perXcd, nTiles);
            break;
        case kFmtRefRGB:
            hipLaunchKernelGGL(k_raycast<kFmtRefRGB>, grid, dim3(64), 0, s, m, vmap, pitch, W, H, K, pose, Tinv,
                               (int)gate, tilesX, perXcd, nTiles);
            break;
        default:
            hipLaunchKernelGGL(k_raycast<kFmtF32>, grid, dim3(64), 0, s, m, vmap, pitch, W, H, K, pose, Tinv,
                               (int)gate, tilesX, perXcd, nTiles);
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
