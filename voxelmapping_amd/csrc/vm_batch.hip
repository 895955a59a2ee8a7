// Multi-scene fuse + raycast: many independent maps of uneven image size,
// voxel size and table size advanced in lockstep. One step runs five
// kernels over the union of a group of scenes' work (allocation tiles,
// allocated blocks -- visibility and the raycast's ray-distance bounds --,
// visible blocks, ray tiles, normal tiles), so small scenes
// share the machine instead of each leaving most of it idle; with several
// groups, each group's kernels run on a stream of its own, so one group's
// latency-bound stages overlap another's arithmetic-bound raycast. The step
// can be captured into a hipGraph (the groups become parallel branches) and
// replayed, since every per-frame input (depth, colour, pose) is found on the
// device from a device-side frame counter.
//
// Per scene the arithmetic is exactly the single-map pipeline's
// (vm_dev_fusion.h, vm_dev_ray.h), so batched and plain results are
// identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "vm_common.h"
#include "vm_dev_fusion.h"
#include "vm_dev_ray.h"
#include "vm_launch.h"
#include "vm_batch.h"

namespace vmamd {

namespace {

constexpr int kMaxScenes = 64;
constexpr int kPersistWavesPerEu = 4;

// s_first[k] = sum of count(j) over scenes j < k, s_first[n] = the total
// (n >= 1), by a scan in the workgroup's first wave: lane j reads scene j's
// count, so the scenes' counters are read in one memory round trip
template <class F>
__device__ inline void scene_prefix(int n, int* s_first, F count) {
    if (threadIdx.x < 64) {
        const int lane = threadIdx.x;
        const int v = lane < n ? count(lane) : 0;
        int incl = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const int u = __shfl_up(incl, off, 64);
            if (lane >= off) incl += u;
        }
        if (lane < n) s_first[lane] = incl - v;
        if (lane == n - 1) s_first[n] = incl;
    }
    __syncthreads();
}

// scene of a flat work index: the last scene whose first index is <= i
__device__ inline int scene_of(const int* first, int n, int i) {
    int s = 0;
    for (int k = 1; k < n; ++k) s = (first[k] <= i) ? k : s;
    return s;
}

__global__ __launch_bounds__(256) void kb_allocate(BatchView B) {
    __shared__ int s_first[kMaxScenes];
    for (int k = threadIdx.x; k < B.nScenes; k += blockDim.x) s_first[k] = B.scenes[k].tile16First;
    __syncthreads();
    const int s = scene_of(s_first, B.nScenes, blockIdx.x);
    const BatchScene& S = B.scenes[s];
    const int f = *B.frame;
    if (f >= S.nFrames) return;  // past its last frame a scene is left as it is
    const int lt = blockIdx.x - S.tile16First;
    if (lt == 0 && threadIdx.x == 0) {
        S.m.ctr[kVisibleCount] = 0;
        S.m.ctr[kFramesFused] += 1;
    }
    const int x = (lt % S.tilesX16) * 16 + (threadIdx.x & 15);
    const int y = (lt / S.tilesX16) * 16 + (threadIdx.x >> 4);
    const bool in = x < S.d.width && y < S.d.height;
    DepthImage d = S.d;
    d.data += (int64_t)f * S.depthFrameStride;
    if (d.format == kDepthU16) allocate_pixel<kDepthU16>(S.m, d, S.K, S.poses[f], x, y, in);
    else allocate_pixel<kDepthF32>(S.m, d, S.K, S.poses[f], x, y, in);
}

// chunks of 256 allocated blocks of every scene, concatenated: the visible
// blocks, and the ray-distance bounds of the step's raycast (same pose)
__global__ __launch_bounds__(256) void kb_visible(BatchView B) {
    __shared__ VisibleShared sh;
    __shared__ BoundsShared<256> bsh;
    __shared__ int s_first[kMaxScenes + 1];
    scene_prefix(B.nScenes, s_first,
                 [&](int k) { return (allocated_count(B.scenes[k].m.ctr, B.scenes[k].m.nBlock) + 255) / 256; });
    const int total = s_first[B.nScenes];
    const int f = *B.frame;
    for (int c = blockIdx.x; c < total; c += gridDim.x) {
        const int s = scene_of(s_first, B.nScenes, c);
        const BatchScene& S = B.scenes[s];
        if (f >= S.nFrames) continue;
        const int A = allocated_count(S.m.ctr, S.m.nBlock);
        const int base = (c - s_first[s]) * 256;
        visible_chunk(S.m, S.K, S.posesInv[f], S.d.width, S.d.height, base, A, sh);
        const Affine& T = S.poses[f];
        const float grow = ray_drift_margin(S.d.width, S.d.height, S.K, T.t0, T.t1, T.t2, S.vsInv);
        const float3 cc = make_float3(T.t0 * S.vsInv, T.t1 * S.vsInv, T.t2 * S.vsInv);
        const int i = base + (int)threadIdx.x;
        const int4 b = i < A ? S.m.blockPos[i] : make_int4(kEmptyKey, 0, 0, 0);
        ray_bounds_blocks<256, kBatchBoundsTileShift>(b, RayBounds{S.rayKeys, S.rayTiles, *B.steps + 1u}, S.d.width,
                                                      S.d.height, S.K, S.posesInv[f], cc, S.m.voxelSize, grow, bsh);
    }
}

// visible blocks of every scene, concatenated; one wave per block
template <int FMT, int DF, bool COLOR>
__global__ __launch_bounds__(256) void kb_integrate(BatchView B) {
    __shared__ int s_first[kMaxScenes + 1];
    scene_prefix(B.nScenes, s_first, [&](int k) { return B.scenes[k].m.ctr[kVisibleCount]; });
    const int total = s_first[B.nScenes];
    const int f = *B.frame;
    const int lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (w < B.nScenes && lane == 0 && f < B.scenes[w].nFrames) {
        const int V = s_first[w + 1] - s_first[w];
        if (V > 0) atomicAdd((unsigned long long*)&B.scenes[w].m.ctr[kSumVisibleLo], (unsigned long long)V);
    }
    for (int i = w; i < total; i += gridDim.x * 4) {
        const int s = scene_of(s_first, B.nScenes, i);
        const BatchScene& S = B.scenes[s];
        if (f >= S.nFrames) continue;
        DepthImage d = S.d;
        d.data += (int64_t)f * S.depthFrameStride;
        ColorImage col = S.c;
        if (COLOR) col.data += (int64_t)f * S.rgbFrameStride;
        const HashEntry e = S.m.visible[i - s_first[s]];
        integrate_block<FMT, DF, COLOR>(S.m, d, col, S.K, S.posesInv[f], e.ptr / kBlockVox,
                                        make_int4(e.x, e.y, e.z, 0), lane);
    }
}

// ray tiles of every scene, concatenated; tile shape tw x (64 / tw); with
// xcdBands the flat tile index is remapped so each XCD gets a contiguous range
template <int FMT>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8))) void kb_raycast(BatchView B, int tw, int xcdBands) {
    __shared__ int s_first[kMaxScenes];
    for (int k = threadIdx.x; k < B.nScenes; k += blockDim.x) s_first[k] = B.scenes[k].rayTileFirst;
    __syncthreads();
    const int g = blockIdx.x;
    const int t = xcdBands ? (g & 7) * B.rayTilesPerXcd + (g >> 3) : g;
    if (t >= B.totalRayTiles) return;
    const int s = scene_of(s_first, B.nScenes, t);
    const BatchScene& S = B.scenes[s];
    const int f = *B.frame;
    if (f >= S.nFrames) return;
    const int th = 64 / tw;
    const int tilesX = (S.d.width + tw - 1) / tw;
    const int lt = t - S.rayTileFirst;
    const int x = (lt % tilesX) * tw + (threadIdx.x % tw);
    const int y = (lt / tilesX) * th + (threadIdx.x / tw);
    if (x >= S.d.width || y >= S.d.height) return;
    if (S.m.ctr[kVisibleCount] == 0) return;  // the facade's gate (vmapping.cpp:87-88)
    const Affine& T = S.poses[f];
    const float margin = ray_drift_margin(S.d.width, S.d.height, S.K, T.t0, T.t1, T.t2, S.vsInv);
    float nearLim, farLim;
    ray_bounds_limits(RayBounds{S.rayKeys, S.rayTiles, *B.steps + 1u},
                      bounds_tile<kBatchBoundsTileShift>(x, y, S.d.width), margin, nearLim, farLim);
    float* out = (float*)((uint8_t*)S.vmap + (int64_t)y * S.vpitch) + x * 4;
    raycast_pixel<FMT>(S.m, out, x, y, S.K, T, S.posesInv[f], S.vsInv, S.rtStep, nearLim, farLim);
}

__global__ __launch_bounds__(256) void kb_normals(BatchView B) {
    __shared__ int s_first[kMaxScenes];
    for (int k = threadIdx.x; k < B.nScenes; k += blockDim.x) s_first[k] = B.scenes[k].tile16First;
    __syncthreads();
    const int s = scene_of(s_first, B.nScenes, blockIdx.x);
    const BatchScene& S = B.scenes[s];
    if (*B.frame >= S.nFrames) return;
    const int lt = blockIdx.x - S.tile16First;
    const int x = (lt % S.tilesX16) * 16 + (threadIdx.x & 15);
    const int y = (lt / S.tilesX16) * 16 + (threadIdx.x >> 4);
    const int W = S.d.width, H = S.d.height;
    if (x >= W - 1 || y >= H - 1) return;
    normal_pixel(S.vmap, S.vpitch, S.nmap, S.npitch, W, H, x, y);
}

// ---------------------------------------------------------------------------
// The same step as one launch: the five stages run in turn inside one kernel
// whose workgroups are all resident at once, separated by software grid
// barriers. Each stage spreads its work over every workgroup (allocation and
// normal tiles and visibility chunks per workgroup, visible blocks and ray
// tiles per wave), with the arithmetic of the stage kernels above.

// Grid barrier over the nWg workgroups of a launch. Arrivals are counted in
// one word that is never reset: barrier k (1..4) of the launch with sequence
// number seq waits for the count to reach (4 seq + k) nWg (mod 2^32), so a
// launch needs no clearing and every launch adds exactly 4 nWg. The wait is
// bounded: a workgroup that polls kGridSpinLimit times without the count
// reaching its target sets the status word and goes on, so a launch whose
// workgroups are not all resident ends (its results are then invalid and the
// status says so) instead of hanging.
constexpr int kGridSpinLimit = 1 << 18;

__device__ inline void grid_barrier(const PersistCtl& P, unsigned target) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();  // the workgroup's writes, before its arrival
        atomicAdd(P.arrivals, 1u);
        int n = 0;
        while ((int)(__hip_atomic_load(P.arrivals, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - target) < 0) {
            if (++n > kGridSpinLimit) {
                atomicOr(P.status, 1);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __threadfence();  // every workgroup's writes, before this one reads them
    }
    __syncthreads();
}

// The next ray tile of a wave from a counter (all lanes call), so a wave
// that finishes early takes the next tile: tiles differ in cost by an
// order of magnitude. (One atomic per item costs about 10-15 ns of the
// counter's serial throughput, too much for the 10^5 visible blocks.)
__device__ inline int wave_take(unsigned* q) {
    int v = 0;
    if ((threadIdx.x & 63) == 0) v = (int)atomicAdd(q, 1u);
    return __shfl(v, 0, 64);
}

template <int FMT, int DF, bool COLOR>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(kPersistWavesPerEu))) void kb_persistent(
    BatchView B, PersistCtl P, int tw, int xcdBands) {
    __shared__ VisibleShared vsh;
    __shared__ BoundsShared<256> bsh;
    __shared__ int s_tile[kMaxScenes], s_ray[kMaxScenes], s_first[kMaxScenes + 1];
    for (int k = threadIdx.x; k < B.nScenes; k += blockDim.x) {
        s_tile[k] = B.scenes[k].tile16First;
        s_ray[k] = B.scenes[k].rayTileFirst;
    }
    const int f = *B.frame;
    const unsigned epoch = *B.steps + 1u;
    const unsigned seq = *P.seq;
    const unsigned nWg = gridDim.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int gw = blockIdx.x * 4 + wave;
    __syncthreads();

    // the raycast's tile counters are cleared before the first barrier
    if (blockIdx.x == 0 && threadIdx.x < 8) atomicExch(P.queues + kQueueStride * threadIdx.x, 0u);

    // 1. allocation, one 16x16 tile per workgroup
    for (int t = blockIdx.x; t < B.totalTile16; t += gridDim.x) {
        const int s = scene_of(s_tile, B.nScenes, t);
        const BatchScene& S = B.scenes[s];
        if (f >= S.nFrames) continue;
        const int lt = t - S.tile16First;
        if (lt == 0 && threadIdx.x == 0) {
            S.m.ctr[kVisibleCount] = 0;
            S.m.ctr[kFramesFused] += 1;
        }
        const int x = (lt % S.tilesX16) * 16 + (threadIdx.x & 15);
        const int y = (lt / S.tilesX16) * 16 + (threadIdx.x >> 4);
        const bool in = x < S.d.width && y < S.d.height;
        DepthImage d = S.d;
        d.data += (int64_t)f * S.depthFrameStride;
        allocate_pixel<DF>(S.m, d, S.K, S.poses[f], x, y, in);
    }
    grid_barrier(P, (4u * seq + 1u) * nWg);

    // 2. visibility and the ray-distance bounds, one 256-block chunk per workgroup
    scene_prefix(B.nScenes, s_first,
                 [&](int k) { return (allocated_count(B.scenes[k].m.ctr, B.scenes[k].m.nBlock) + 255) / 256; });
    for (int c = blockIdx.x; c < s_first[B.nScenes]; c += gridDim.x) {
        const int s = scene_of(s_first, B.nScenes, c);
        const BatchScene& S = B.scenes[s];
        if (f >= S.nFrames) continue;
        const int A = allocated_count(S.m.ctr, S.m.nBlock);
        const int base = (c - s_first[s]) * 256;
        visible_chunk(S.m, S.K, S.posesInv[f], S.d.width, S.d.height, base, A, vsh);
        const Affine& T = S.poses[f];
        const float grow = ray_drift_margin(S.d.width, S.d.height, S.K, T.t0, T.t1, T.t2, S.vsInv);
        const float3 cc = make_float3(T.t0 * S.vsInv, T.t1 * S.vsInv, T.t2 * S.vsInv);
        const int i = base + (int)threadIdx.x;
        const int4 b = i < A ? S.m.blockPos[i] : make_int4(kEmptyKey, 0, 0, 0);
        ray_bounds_blocks<256, kBatchBoundsTileShift>(b, RayBounds{S.rayKeys, S.rayTiles, epoch}, S.d.width,
                                                      S.d.height, S.K, S.posesInv[f], cc, S.m.voxelSize, grow, bsh);
    }
    grid_barrier(P, (4u * seq + 2u) * nWg);

    // 3. integration, one visible block per wave
    scene_prefix(B.nScenes, s_first, [&](int k) { return B.scenes[k].m.ctr[kVisibleCount]; });
    if (gw < B.nScenes && lane == 0 && f < B.scenes[gw].nFrames) {
        const int V = s_first[gw + 1] - s_first[gw];
        if (V > 0) atomicAdd((unsigned long long*)&B.scenes[gw].m.ctr[kSumVisibleLo], (unsigned long long)V);
    }
    for (int i = gw; i < s_first[B.nScenes]; i += gridDim.x * 4) {
        const int s = scene_of(s_first, B.nScenes, i);
        const BatchScene& S = B.scenes[s];
        if (f >= S.nFrames) continue;
        DepthImage d = S.d;
        d.data += (int64_t)f * S.depthFrameStride;
        ColorImage col = S.c;
        if (COLOR) col.data += (int64_t)f * S.rgbFrameStride;
        const HashEntry e = S.m.visible[i - s_first[s]];
        integrate_block<FMT, DF, COLOR>(S.m, d, col, S.K, S.posesInv[f], e.ptr / kBlockVox,
                                        make_int4(e.x, e.y, e.z, 0), lane);
    }
    grid_barrier(P, (4u * seq + 3u) * nWg);

    // 4. raycast, one tile of 64 rays per wave; with xcdBands the waves of
    // XCD x (workgroups x, x + 8, ...) take the x-th eighth of the tiles,
    // from a counter of their own
    {
        const int th = 64 / tw;
        int t0 = 0, t1 = B.totalRayTiles;
        unsigned* q = P.queues;
        if (xcdBands) {
            const int x = blockIdx.x & 7;
            t0 = min(x * B.rayTilesPerXcd, B.totalRayTiles);
            t1 = min(t0 + B.rayTilesPerXcd, B.totalRayTiles);
            q += x * kQueueStride;
        }
        for (int t = t0 + wave_take(q); t < t1; t = t0 + wave_take(q)) {
            const int s = scene_of(s_ray, B.nScenes, t);
            const BatchScene& S = B.scenes[s];
            if (f >= S.nFrames) continue;
            const int tilesX = (S.d.width + tw - 1) / tw;
            const int lt = t - S.rayTileFirst;
            const int x = (lt % tilesX) * tw + (lane % tw);
            const int y = (lt / tilesX) * th + (lane / tw);
            if (S.m.ctr[kVisibleCount] == 0) continue;  // the facade's gate (vmapping.cpp:87-88)
            // lanes off the image skip the body (no divergent continue: the
            // whole wave takes the next tile)
            if (x < S.d.width && y < S.d.height) {
                const Affine& T = S.poses[f];
                const float margin = ray_drift_margin(S.d.width, S.d.height, S.K, T.t0, T.t1, T.t2, S.vsInv);
                float nearLim, farLim;
                ray_bounds_limits(RayBounds{S.rayKeys, S.rayTiles, epoch},
                                  bounds_tile<kBatchBoundsTileShift>(x, y, S.d.width), margin, nearLim, farLim);
                float* out = (float*)((uint8_t*)S.vmap + (int64_t)y * S.vpitch) + x * 4;
                raycast_pixel<FMT>(S.m, out, x, y, S.K, T, S.posesInv[f], S.vsInv, S.rtStep, nearLim, farLim);
            }
        }
    }
    grid_barrier(P, (4u * seq + 4u) * nWg);
    // every workgroup has read the frame, step and sequence numbers
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        B.frame[0] = f + 1;
        B.frame[1] = (int)epoch;
        *P.seq = seq + 1u;
    }

    // 5. normals, one 16x16 tile per workgroup
    for (int t = blockIdx.x; t < B.totalTile16; t += gridDim.x) {
        const int s = scene_of(s_tile, B.nScenes, t);
        const BatchScene& S = B.scenes[s];
        if (f >= S.nFrames) continue;
        const int lt = t - S.tile16First;
        const int x = (lt % S.tilesX16) * 16 + (threadIdx.x & 15);
        const int y = (lt / S.tilesX16) * 16 + (threadIdx.x >> 4);
        if (x < S.d.width - 1 && y < S.d.height - 1)
            normal_pixel(S.vmap, S.vpitch, S.nmap, S.npitch, S.d.width, S.d.height, x, y);
    }
}

// frame[0]: the frame index; frame[1]: the step count (the ray-bound epoch)
__global__ void kb_advance(int* frame) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        frame[0] += 1;
        frame[1] += 1;
    }
}

__global__ void kb_set_frame(int* frame, int v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *frame = v;
}

}  // namespace

hipError_t batch_launch_step(const BatchView& B, const BatchTuning& t, int fmt, int df, bool color, hipStream_t s,
                             bool advance) {
    const int cu = num_compute_units();
    hipLaunchKernelGGL(kb_allocate, dim3(B.totalTile16), dim3(256), 0, s, B);
    hipLaunchKernelGGL(kb_visible, dim3(std::max(1, std::min(cu * t.visibleWgPerCu, B.maxChunks))), dim3(256), 0,
                       s, B);
    const dim3 gi(std::max(1, std::min(cu * t.integrateWgPerCu, (B.maxVisible + 3) / 4)));
#define VM_INTEGRATE(F, D, C) hipLaunchKernelGGL((kb_integrate<F, D, C>), gi, dim3(256), 0, s, B)
    if (fmt == kFmtRef) {
        if (df == kDepthU16) VM_INTEGRATE(kFmtRef, kDepthU16, false); else VM_INTEGRATE(kFmtRef, kDepthF32, false);
    } else if (fmt == kFmtRefRGB) {
        if (df == kDepthU16) { if (color) VM_INTEGRATE(kFmtRefRGB, kDepthU16, true); else VM_INTEGRATE(kFmtRefRGB, kDepthU16, false); }
        else { if (color) VM_INTEGRATE(kFmtRefRGB, kDepthF32, true); else VM_INTEGRATE(kFmtRefRGB, kDepthF32, false); }
    } else {
        if (df == kDepthU16) { if (color) VM_INTEGRATE(kFmtF32, kDepthU16, true); else VM_INTEGRATE(kFmtF32, kDepthU16, false); }
        else { if (color) VM_INTEGRATE(kFmtF32, kDepthF32, true); else VM_INTEGRATE(kFmtF32, kDepthF32, false); }
    }
#undef VM_INTEGRATE
    const int tiles = t.xcdBands ? B.rayTilesPerXcd * 8 : B.totalRayTiles;
    if (fmt == kFmtRef) hipLaunchKernelGGL(kb_raycast<kFmtRef>, dim3(tiles), dim3(64), 0, s, B, t.rayTileW, t.xcdBands);
    else if (fmt == kFmtRefRGB) hipLaunchKernelGGL(kb_raycast<kFmtRefRGB>, dim3(tiles), dim3(64), 0, s, B, t.rayTileW, t.xcdBands);
    else hipLaunchKernelGGL(kb_raycast<kFmtF32>, dim3(tiles), dim3(64), 0, s, B, t.rayTileW, t.xcdBands);
    hipLaunchKernelGGL(kb_normals, dim3(B.totalTile16), dim3(256), 0, s, B);
    if (advance) hipLaunchKernelGGL(kb_advance, dim3(1), dim3(64), 0, s, B.frame);
    return hipGetLastError();
}

template <int FMT, int DF, bool COLOR>
hipError_t persist_launch(const BatchView& B, const PersistCtl& P, const BatchTuning& t, hipStream_t s) {
    int perCu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kb_persistent<FMT, DF, COLOR>, 256, 0);
    if (e != hipSuccess) return e;
    if (perCu < 1) return hipErrorInvalidConfiguration;
    const int grid = num_compute_units() * std::min(perCu, t.persistWgPerCu);
    if (t.persistent == 2) {
        BatchView b = B;
        PersistCtl p = P;
        int tw = t.rayTileW, xb = t.xcdBands;
        void* args[] = {&b, &p, &tw, &xb};
        return hipLaunchCooperativeKernel((const void*)kb_persistent<FMT, DF, COLOR>, dim3(grid), dim3(256), args, 0,
                                          s);
    }
    hipLaunchKernelGGL((kb_persistent<FMT, DF, COLOR>), dim3(grid), dim3(256), 0, s, B, P, t.rayTileW, t.xcdBands);
    return hipGetLastError();
}

hipError_t batch_launch_persistent(const BatchView& B, const PersistCtl& P, const BatchTuning& t, int fmt, int df,
                                   bool color, hipStream_t s) {
    if (fmt == kFmtRef) {
        if (df == kDepthU16) return persist_launch<kFmtRef, kDepthU16, false>(B, P, t, s);
        return persist_launch<kFmtRef, kDepthF32, false>(B, P, t, s);
    }
    if (fmt == kFmtRefRGB) {
        if (df == kDepthU16) return color ? persist_launch<kFmtRefRGB, kDepthU16, true>(B, P, t, s)
                                          : persist_launch<kFmtRefRGB, kDepthU16, false>(B, P, t, s);
        return color ? persist_launch<kFmtRefRGB, kDepthF32, true>(B, P, t, s)
                     : persist_launch<kFmtRefRGB, kDepthF32, false>(B, P, t, s);
    }
    if (df == kDepthU16) return color ? persist_launch<kFmtF32, kDepthU16, true>(B, P, t, s)
                                      : persist_launch<kFmtF32, kDepthU16, false>(B, P, t, s);
    return color ? persist_launch<kFmtF32, kDepthF32, true>(B, P, t, s)
                 : persist_launch<kFmtF32, kDepthF32, false>(B, P, t, s);
}

hipError_t batch_advance(int* frame, hipStream_t s) {
    hipLaunchKernelGGL(kb_advance, dim3(1), dim3(64), 0, s, frame);
    return hipGetLastError();
}

hipError_t batch_set_frame(int* frame, int v, hipStream_t s) {
    hipLaunchKernelGGL(kb_set_frame, dim3(1), dim3(64), 0, s, frame, v);
    return hipGetLastError();
}

int batch_max_scenes() { return kMaxScenes; }

}  // namespace vmamd
