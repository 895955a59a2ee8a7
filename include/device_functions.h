// Drop-in replacement of the reference's stage API (src/device_functions.h:
// FuseImage, FuseDepthAndImage, CheckVisibleBlocks, ProjectRenderingBlocks,
// RenderScene, Polygonize, GetSurfacePoints), header-only over the C ABI.
// Same names, argument order and meaning; images are vmap::GpuMat views of
// device memory, matrices column-major.
#ifndef VMAPPING_DEVICE_FUNCTIONS_H
#define VMAPPING_DEVICE_FUNCTIONS_H

#include "map_struct.h"

namespace vmap {

namespace detail {
inline void k4(const Matrix3f& K, float out[4]) {
    out[0] = K(0, 0);
    out[1] = K(1, 1);
    out[2] = K(0, 2);
    out[3] = K(1, 2);
}
inline int depth_format(const GpuMat& d) {
    if (d.type() == CV_32FC1) return VM_DEPTH_F32;
    if (d.type() == CV_16UC1) return VM_DEPTH_U16;
    throw std::invalid_argument("depth must be CV_32FC1 (metres) or CV_16UC1");
}
}  // namespace detail

// depth_fusion.cu:199-301. Returns the visible-block count. depthScale
// converts CV_16UC1 depth to metres (ignored for CV_32FC1).
template <class TVoxel>
int FuseImage(MapStruct<TVoxel> map, const GpuMat depth, const Matrix4f& camToWorld, const Matrix3f& K,
              float depthScale = 0.001f) {
    float k[4];
    detail::k4(K, k);
    uint32_t n = 0;
    vm_check(vm_fuse_depth(map.map, depth.data, (int64_t)depth.step, depth.cols, depth.rows,
                           detail::depth_format(depth), depthScale, nullptr, 0, k, camToWorld.data(), nullptr, &n));
    return (int)n;
}

// depth_fusion.cu:383-488 (image: CV_8UC3 device image)
template <class TVoxel>
int FuseDepthAndImage(MapStruct<TVoxel> map, const GpuMat image, const GpuMat depth, const Matrix4f& camToWorld,
                      const Matrix3f& K, float depthScale = 0.001f) {
    if (image.type() != CV_8UC3) throw std::invalid_argument("image must be CV_8UC3");
    float k[4];
    detail::k4(K, k);
    uint32_t n = 0;
    vm_check(vm_fuse_depth(map.map, depth.data, (int64_t)depth.step, depth.cols, depth.rows,
                           detail::depth_format(depth), depthScale, image.data, (int64_t)image.step, k,
                           camToWorld.data(), nullptr, &n));
    return (int)n;
}

// depth_fusion.cu:494-539
template <class TVoxel>
int CheckVisibleBlocks(MapStruct<TVoxel> map, const int& width, const int& height, const Matrix4f& camToWorld,
                       const Matrix3f& K) {
    float k[4];
    detail::k4(K, k);
    uint32_t n = 0;
    vm_check(vm_check_visible_blocks(map.map, width, height, k, camToWorld.data(), nullptr, &n));
    return (int)n;
}

// ray_tracing.cu:171-234. The z-range maps are filled directly; the
// rendering-block list argument is accepted for signature compatibility and
// not written (the reference's raycast never reads it either).
template <class TVoxel>
void ProjectRenderingBlocks(MapStruct<TVoxel> map, unsigned int count_visible_block,
                            unsigned int& count_rendering_block, GpuMat& zRangeX, GpuMat& zRangeY,
                            RenderingBlock* /*listRenderingBlock*/, const Matrix4f& worldToCam, const Matrix3f& K) {
    count_rendering_block = 0;
    if (count_visible_block == 0) return;
    if (zRangeX.step != zRangeY.step) throw std::invalid_argument("z-range maps must share a pitch");
    float k[4];
    detail::k4(K, k);
    uint32_t n = 0;
    vm_check(vm_project_rendering_blocks(map.map, (float*)zRangeX.data, (float*)zRangeY.data,
                                         (int64_t)zRangeX.step, zRangeX.cols, zRangeX.rows, k, worldToCam.data(),
                                         nullptr, &n));
    count_rendering_block = n;
}

#ifdef VMAPPING_HAVE_OPENCV
// The reference's signature (src/device_functions.h:35-43): the z-range maps
// are cv::cuda::GpuMat outputs; empty ones are created at 1/8 of the image
// size given by cols8 x rows8 = the reference facade's (width / 8) x
// (height / 8) (vmapping.cpp:62-63).
template <class TVoxel>
void ProjectRenderingBlocks(MapStruct<TVoxel> map, unsigned int count_visible_block,
                            unsigned int& count_rendering_block, cv::cuda::GpuMat& zRangeX, cv::cuda::GpuMat& zRangeY,
                            RenderingBlock* listRenderingBlock, const Matrix4f& worldToCam, const Matrix3f& K,
                            int cols8 = 0, int rows8 = 0) {
    if (zRangeX.empty()) detail::ensure_output(zRangeX, rows8, cols8, CV_32FC1);
    if (zRangeY.empty()) detail::ensure_output(zRangeY, zRangeX.rows, zRangeX.cols, CV_32FC1);
    GpuMat zx(zRangeX), zy(zRangeY);
    ProjectRenderingBlocks(map, count_visible_block, count_rendering_block, zx, zy, listRenderingBlock, worldToCam,
                           K);
}
#endif

// ray_tracing.cu:403-437 (zRange maps unused, as in the reference)
template <class TVoxel>
void RenderScene(MapStruct<TVoxel> map, GpuMat vmap, GpuMat /*zRangeX*/, GpuMat /*zRangeY*/,
                 const Matrix4f& camToWorld, const Matrix3f& K) {
    float k[4];
    detail::k4(K, k);
    vm_check(vm_raycast(map.map, (float*)vmap.data, (int64_t)vmap.step, vmap.cols, vmap.rows, k, camToWorld.data(),
                        nullptr));
}

// marching_cubes.cu:251-305
template <class TVoxel>
void Polygonize(MapStruct<TVoxel> map, unsigned int& block_count, unsigned int& triangle_count, void* vertex_out,
                void* normal_out, size_t bufferSize) {
    uint32_t b = 0, t = 0;
    vm_check(vm_polygonize(map.map, (float*)vertex_out, (float*)normal_out, (int64_t)bufferSize, nullptr, &b, &t));
    block_count = b;
    triangle_count = t;
}

// marching_cubes.cu:330-344
inline void GetSurfacePoints(void* vertex_in, void* vertex_out, unsigned int triangle_count) {
    vm_check(vm_surface_points((const float*)vertex_in, (float*)vertex_out, triangle_count, nullptr));
    vm_check(hipDeviceSynchronize() == hipSuccess ? VM_OK : VM_ERR_HIP);
}

}  // namespace vmap

#endif
