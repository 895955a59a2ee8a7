// Single-map kernels of reset, block allocation, visibility and TSDF
// integration for gfx950 (device functions in vm_dev_fusion.h).
#include <algorithm>

#include "vm_common.h"
#include "vm_dev_fusion.h"
#include "vm_dev_ray.h"
#include "vm_launch.h"

namespace vmamd {

namespace {

// ---------------------------------------------------------------------------
// reset (map_struct.cu:7-42)

__global__ void k_reset_table(HashEntry* t, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    t[i].ptr = -1;
    t[i].offset = -1;
    t[i].x = kEmptyKey;
    t[i].y = kEmptyKey;
    t[i].z = kEmptyKey;
}

// heap[i] = nBlock - 1 - i and heapPtr = nBlock - 1, so pops return 0, 1, 2...
__global__ void k_reset_heap(int* heap, int4* blockPos, int* ctr, int* mutex, int nBlock, int nBucket) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nBlock) {
        heap[i] = nBlock - 1 - i;
        blockPos[i] = make_int4(kEmptyKey, kEmptyKey, kEmptyKey, -1);
    }
    if (i < nBucket) mutex[i] = 0;
    if (i < kCounterInts) ctr[i] = (i == kHeapPtr) ? nBlock - 1 : 0;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t v, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t n4 = n >> 2;
    uint4 v4 = make_uint4(v, v, v, v);
    for (int64_t j = i; j < n4; j += stride) ((uint4*)p)[j] = v4;
    for (int64_t j = (n4 << 2) + i; j < n; j += stride) p[j] = v;
}

__global__ void k_clear_block_pos(int4* blockPos, int nBlock) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nBlock) blockPos[i] = make_int4(kEmptyKey, kEmptyKey, kEmptyKey, -1);
}

__global__ void k_rebuild_index(MapView m) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m.nEntry) return;
    HashEntry e = m.table[i];
    if (e.ptr >= 0 && e.ptr / kBlockVox < m.nBlock) {
        m.blockPos[e.ptr / kBlockVox] = make_int4(e.x, e.y, e.z, i);
        occupancy_mark(m, e.x, e.y, e.z);
    }
}

// ---------------------------------------------------------------------------
// allocation: one thread per pixel, 16x16 pixel tiles. Thread 0 also zeroes
// the visible count for the visibility pass that follows.
template <int DF>
__global__ __launch_bounds__(256) void k_allocate(MapView m, DepthImage d, Intrinsics K, Affine T) {
    if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0) {
        m.ctr[kVisibleCount] = 0;
        m.ctr[kFramesFused] += 1;
    }
    const int x = blockIdx.x * 16 + (threadIdx.x & 15);
    const int y = blockIdx.y * 16 + (threadIdx.x >> 4);
    allocate_pixel<DF>(m, d, K, T, x, y, x < d.width && y < d.height);
}

// Visibility over the allocated blocks [0, A) instead of the whole hash
// table: the allocated ids are exactly the popped heap entries. Persistent
// grid; the block count stays on the device.
__global__ __launch_bounds__(256) void k_visible(MapView m, Intrinsics K, Affine Tinv, int cols, int rows) {
    __shared__ VisibleShared sh;
    const int A = allocated_count(m.ctr, m.nBlock);
    for (int chunk = blockIdx.x * 256; chunk < A; chunk += gridDim.x * 256)
        visible_chunk(m, K, Tinv, cols, rows, chunk, A, sh);
}

// The same, plus the per-tile ray-distance bounds of a raycast at this pose
// (ray_bounds_chunk, vm_dev_ray.h), which the raycast that follows at the
// same view then reuses: workgroups from the front of the grid take the
// 256-block visibility chunks, workgroups from its back the 32-block bounds
// chunks, so both run side by side in one launch.
__global__ __launch_bounds__(256) void k_visible_bounds(MapView m, Intrinsics K, Affine Tinv, int cols, int rows,
                                                        RayBounds rb, float3 c, float grow) {
    __shared__ VisibleShared sh;
    __shared__ BoundsShared<kBoundsChunk> bsh;
    const int A = allocated_count(m.ctr, m.nBlock);
    for (int chunk = blockIdx.x * 256; chunk < A; chunk += gridDim.x * 256)
        visible_chunk(m, K, Tinv, cols, rows, chunk, A, sh);
    for (int base = (gridDim.x - 1 - blockIdx.x) * kBoundsChunk; base < A; base += gridDim.x * kBoundsChunk)
        ray_bounds_chunk<kBoundsChunk>(m, rb, base, A, cols, rows, K, Tinv, c, m.voxelSize, grow, bsh);
}

// Integration of the compacted visible table, one wave per visible block,
// persistent grid over the device-side count; also accumulates the frame's
// visible count into the running sum.
template <int FMT, int DF, bool COLOR>
__global__ __launch_bounds__(256) void k_integrate(MapView m, DepthImage d, ColorImage c, Intrinsics K, Affine Tinv) {
    const int V = m.ctr[kVisibleCount];
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w == 0 && lane == 0 && V > 0)
        atomicAdd((unsigned long long*)&m.ctr[kSumVisibleLo], (unsigned long long)V);
    for (int i = w; i < V; i += gridDim.x * 4) {
        const HashEntry e = m.visible[i];
        integrate_block<FMT, DF, COLOR>(m, d, c, K, Tinv, e.ptr / kBlockVox, make_int4(e.x, e.y, e.z, 0), lane);
    }
}

int grid_for_chunks(const MapView& m) {
    const int cu = num_compute_units();
    const int maxChunks = (m.nBlock + 255) / 256;
    const int g = cu * 4;
    return g < maxChunks ? g : (maxChunks > 0 ? maxChunks : 1);
}

// four waves per workgroup, one visible block per wave per iteration
int grid_for_visible(const MapView& m) {
    const int cu = num_compute_units();
    const int maxVis = (m.nBlock < m.nEntry ? m.nBlock : m.nEntry);
    const int need = (maxVis + 3) / 4;
    const int g = cu * 8;
    return g < need ? g : (need > 0 ? need : 1);
}

template <int FMT, int DF, bool COLOR>
hipError_t launch_int(const MapView& m, const DepthImage& d, const ColorImage& c, const Intrinsics& K,
                      const Affine& Tinv, hipStream_t s) {
    hipLaunchKernelGGL((k_integrate<FMT, DF, COLOR>), dim3(grid_for_visible(m)), dim3(256), 0, s, m, d, c, K, Tinv);
    return hipGetLastError();
}

template <int FMT, int DF>
hipError_t launch_int_c(const MapView& m, const DepthImage& d, const ColorImage& c, const Intrinsics& K,
                        const Affine& Tinv, hipStream_t s) {
    if (c.data && Vox<FMT>::kHasColor) return launch_int<FMT, DF, Vox<FMT>::kHasColor>(m, d, c, K, Tinv, s);
    return launch_int<FMT, DF, false>(m, d, c, K, Tinv, s);
}

template <int FMT>
hipError_t launch_int_d(const MapView& m, const DepthImage& d, const ColorImage& c, const Intrinsics& K,
                        const Affine& Tinv, hipStream_t s) {
    if (d.format == kDepthU16) return launch_int_c<FMT, kDepthU16>(m, d, c, K, Tinv, s);
    return launch_int_c<FMT, kDepthF32>(m, d, c, K, Tinv, s);
}

}  // namespace

int num_compute_units() {
    static int cu = 0;
    if (cu == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0)
            cu = 256;
    }
    return cu;
}

hipError_t launch_fill_u32(void* dst, uint32_t value, int64_t count, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    int64_t blocks = (count / 4 + 255) / 256;
    int g = (int)(blocks < 4096 ? (blocks > 0 ? blocks : 1) : 4096);
    hipLaunchKernelGGL(k_fill_u32, dim3(g), dim3(256), 0, s, (uint32_t*)dst, value, count);
    return hipGetLastError();
}

hipError_t launch_reset(const MapView& m, hipStream_t s) {
    hipLaunchKernelGGL(k_reset_table, dim3((m.nEntry + 255) / 256), dim3(256), 0, s, m.table, m.nEntry);
    int n = m.nBlock > m.nBucket ? m.nBlock : m.nBucket;
    if (n < kCounterInts) n = kCounterInts;
    hipLaunchKernelGGL(k_reset_heap, dim3((n + 255) / 256), dim3(256), 0, s, m.heap, m.blockPos, m.ctr, m.mutex,
                       m.nBlock, m.nBucket);
    int64_t bytes = (int64_t)m.nBlock * kBlockVox * voxel_bytes(m.fmt);
    hipError_t e = launch_fill_u32(m.voxels, 0u, bytes / 4, s);
    if (e != hipSuccess) return e;
    e = launch_fill_u32(m.occ, 0u, 2 * ((int64_t)m.occMask + 1), s);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

// block positions and occupancy bricks from the hash table alone
hipError_t launch_rebuild_index(const MapView& m, hipStream_t s) {
    hipError_t e = launch_fill_u32(m.occ, 0u, 2 * ((int64_t)m.occMask + 1), s);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_clear_block_pos, dim3((m.nBlock + 255) / 256), dim3(256), 0, s, m.blockPos, m.nBlock);
    hipLaunchKernelGGL(k_rebuild_index, dim3((m.nEntry + 255) / 256), dim3(256), 0, s, m);
    return hipGetLastError();
}

hipError_t launch_allocate(const MapView& m, const DepthImage& d, const Intrinsics& K, const Affine& T,
                           hipStream_t s) {
    dim3 grid((d.width + 15) / 16, (d.height + 15) / 16);
    if (d.format == kDepthU16)
        hipLaunchKernelGGL(k_allocate<kDepthU16>, grid, dim3(256), 0, s, m, d, K, T);
    else
        hipLaunchKernelGGL(k_allocate<kDepthF32>, grid, dim3(256), 0, s, m, d, K, T);
    return hipGetLastError();
}

hipError_t launch_visible_compact(const MapView& m, int cols, int rows, const Intrinsics& K, const Affine& Tinv,
                                  hipStream_t s, const BoundsAtPose* bounds) {
    if (bounds && bounds->rb.keys) {
        const int need = (m.nBlock + kBoundsChunk - 1) / kBoundsChunk, gb = num_compute_units() * 8;
        const int g = std::max(grid_for_chunks(m), std::min(need, gb));
        hipLaunchKernelGGL(k_visible_bounds, dim3(g), dim3(256), 0, s, m, K, Tinv, cols, rows, bounds->rb,
                           bounds->centre, bounds->grow);
    } else {
        hipLaunchKernelGGL(k_visible, dim3(grid_for_chunks(m)), dim3(256), 0, s, m, K, Tinv, cols, rows);
    }
    return hipGetLastError();
}

hipError_t launch_integrate(const MapView& m, const DepthImage& d, const ColorImage& c, const Intrinsics& K,
                            const Affine& Tinv, hipStream_t s) {
    switch (m.fmt) {
        case kFmtRef: return launch_int_d<kFmtRef>(m, d, c, K, Tinv, s);
        case kFmtRefRGB: return launch_int_d<kFmtRefRGB>(m, d, c, K, Tinv, s);
        default: return launch_int_d<kFmtF32>(m, d, c, K, Tinv, s);
    }
}

hipError_t launch_visible_fuse(const MapView& m, const DepthImage& d, const ColorImage& c, const Intrinsics& K,
                               const Affine& Tinv, hipStream_t s, const BoundsAtPose* bounds) {
    hipError_t e = launch_visible_compact(m, d.width, d.height, K, Tinv, s, bounds);
    if (e != hipSuccess) return e;
    return launch_integrate(m, d, c, K, Tinv, s);
}

hipError_t launch_visible(const MapView& m, int cols, int rows, const Intrinsics& K, const Affine& Tinv,
                          hipStream_t s) {
    // a zero-count reset first: the allocation kernel is not run here
    hipError_t e = launch_fill_u32(m.ctr + kVisibleCount, 0u, 1, s);
    if (e != hipSuccess) return e;
    return launch_visible_compact(m, cols, rows, K, Tinv, s);
}

}  // namespace vmamd
