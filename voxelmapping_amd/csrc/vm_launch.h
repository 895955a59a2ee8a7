// Host-side launchers of the gfx950 kernels (one translation unit per stage).
#pragma once

#include "vm_common.h"

namespace vmamd {

// map_struct.cu:29-42 equivalent (plus the block-position index)
hipError_t launch_reset(const MapView& m, hipStream_t s);
// rebuild blockPos from the hash table (after state is loaded from outside)
hipError_t launch_rebuild_index(const MapView& m, hipStream_t s);
// CreateBlockLineTracingFunctor, depth_fusion.cu:8-60; also zeroes the visible count
hipError_t launch_allocate(const MapView& m, const DepthImage& d, const Intrinsics& K,
                           const Affine& camToWorld, hipStream_t s);
// The raycast's per-tile ray-distance bounds (RayBounds, vm_dev_ray.h) at a
// pose, computed in the visibility launch of a fusion at that pose: the
// camera centre and the boxes' growth (ray_drift_margin) in voxels.
struct BoundsAtPose {
    RayBounds rb;
    float3 centre;
    float grow;
};
// CheckEntryVisibilityFunctor + DepthFusionFunctor/ImageFusionFunctor fused;
// with `bounds`, the visibility launch also computes the ray bounds of a
// raycast at the same pose, image size and intrinsics
hipError_t launch_visible_fuse(const MapView& m, const DepthImage& d, const ColorImage& c,
                               const Intrinsics& K, const Affine& worldToCam, hipStream_t s,
                               const BoundsAtPose* bounds = nullptr);
// the two halves of launch_visible_fuse: compaction of the visible blocks
// (the count must be zero beforehand) and integration of the compacted list
hipError_t launch_visible_compact(const MapView& m, int cols, int rows, const Intrinsics& K,
                                  const Affine& worldToCam, hipStream_t s, const BoundsAtPose* bounds = nullptr);
hipError_t launch_integrate(const MapView& m, const DepthImage& d, const ColorImage& c, const Intrinsics& K,
                            const Affine& worldToCam, hipStream_t s);
// CheckEntryVisibilityFunctor alone (CheckVisibleBlocks, depth_fusion.cu:494-539)
hipError_t launch_visible(const MapView& m, int width, int height, const Intrinsics& K,
                          const Affine& worldToCam, hipStream_t s);
// MapRenderingDelegate, ray_tracing.cu:255-401. gate: skip when the visible
// count is zero (the facade's gate, vmapping.cpp:87-88 / :101)
// bounds: null, or the per-tile ray-distance bounds (RayBounds, vm_dev_ray.h)
// for this image size and a fresh epoch, computed before the march
hipError_t launch_raycast(const MapView& m, float* vmap, int64_t pitch, int width, int height,
                          const Intrinsics& K, const Affine& camToWorld, const Affine& worldToCam,
                          bool gate, hipStream_t s, const RayBounds* bounds = nullptr, bool boundsReady = false);
// tiles of a width x height image (the bounds hold 2 (tiles + 1) 64-bit keys)
inline int ray_bounds_tiles(int width, int height) {
    const int t = 1 << kBoundsTileShift;
    return ((width + t - 1) / t) * ((height + t - 1) / t);
}
// RenderingBlockDelegate + FillRenderingBlockFunctor, ray_tracing.cu:13-234
hipError_t launch_rendering_blocks(const MapView& m, float* zx, float* zy, int64_t pitch,
                                   int cols8, int rows8, const Intrinsics& K,
                                   const Affine& worldToCam, hipStream_t s);
// computeNormalKernel, image_proc.cu:74-104
hipError_t launch_normals(const float* vmap, int64_t vpitch, float* nmap, int64_t npitch,
                          int width, int height, hipStream_t s);
// renderSceneKernel, image_proc.cu:46-72
hipError_t launch_shading(const float* vmap, int64_t vpitch, const float* nmap, int64_t npitch,
                          uint8_t* rgba, int64_t rpitch, int width, int height, hipStream_t s);
// BuildVertexArray / Polygonize, marching_cubes.cu:9-305; `work` holds
// polygonize_work_ints(capBlocks, polygonize_groups(nBlock)) ints
// (neighbour offsets, per-block triangle counts, workgroup offsets;
// rewritten by every call, nothing in it needs clearing). The passes take
// the first min(allocated, capBlocks) blocks: callers size capBlocks to the
// pool or check the allocated count first.
int polygonize_groups(int nBlock);
size_t polygonize_work_ints(int capBlocks, int groups);
hipError_t launch_polygonize(const MapView& m, int* work, int capBlocks, float* verts, float* norms,
                             int64_t capTriangles, hipStream_t s);
// compute_surface_point_kernel, marching_cubes.cu:314-344
hipError_t launch_surface_points(const float* verts, float* points, int64_t n, hipStream_t s);

// adjoint of the raycast + normals (float voxel format); scratch must hold
// raycast_backward_scratch_bytes(W, H); stats = {hit pixels, overflowed
// pixels} or null; launchBounds: compute rb's bounds first (false when rb's
// epoch already holds them for this pose)
size_t raycast_backward_scratch_bytes(int W, int H);
hipError_t launch_raycast_backward(const MapView& m, const float* vmap, int64_t vpitch, const float* gVertex,
                                   int64_t gvpitch, const float* gNormal, int64_t gnpitch, int W, int H,
                                   const Intrinsics& K, const Affine& camToWorld, const Affine& worldToCam,
                                   float* gTsdf, double* gPose, void* scratch, int* stats, const RayBounds& rb,
                                   bool launchBounds, hipStream_t s);
// the per-tile ray-distance bounds of rb's epoch for a raycast at this pose;
// grow = ray_drift_margin of the pose (vm_dev_ray.h)
hipError_t launch_ray_bounds(const MapView& m, const RayBounds& rb, int W, int H, const Intrinsics& K,
                             const Affine& camToWorld, const Affine& worldToCam, float grow, hipStream_t s);

// device-side fill of 32-bit words (graph-capturable replacement of memset)
hipError_t launch_fill_u32(void* dst, uint32_t value, int64_t count, hipStream_t s);

int num_compute_units();

// batched Gauss-Newton pose refinement (vm_pose.hip)
size_t refine_poses_scratch_bytes(int width, int height, int nHyp, int iters, int stride);
hipError_t launch_refine_poses(const MapView& m, const DepthImage& d, const Intrinsics& K, const double* posesIn,
                               int nHyp, int iters, int stride, float maxRes, double lambda, void* scratch,
                               double** posesOut, float** costOut, double** systemOut, hipStream_t s);

}  // namespace vmamd
