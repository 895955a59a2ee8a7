// Drop-in replacement of the reference's include/vmapping.h:10-27
// (vmap::VoxelMapping), implemented in libvmapping_amd.so over the C ABI.
// Eigen / OpenCV argument types are replaced by the layout-compatible types
// of vm_types.h (implicitly constructible from Eigen / OpenCV when present);
// with OpenCV and Eigen present, the overloads at the end of the class take
// the reference's own types for the arguments that a conversion cannot bind
// (output images, host images, the surface-point vector).
#ifndef VMAPPING_VMAPPING_H
#define VMAPPING_VMAPPING_H

#include <memory>
#include <vector>

#include "vm_types.h"

namespace vmap {

class VoxelMapping {
public:
    VoxelMapping(int w, int h, const Matrix3f& K);
    ~VoxelMapping();

    void reset();
    // nBucket = int(0.9 * numVoxels), truncation = 4 * voxelSize, vmap::Voxel
    // storage (vmapping.cpp:57-75)
    void CreateMap(int numEntries, int numVoxels, float voxelSize);
    // depth: CV_32FC1 metres or CV_16UC1 millimetres
    void FuseDepth(GpuMat depth, const Matrix4f& camToWorld);
    // A stub in the reference (vmapping.cpp:167-170). Here it fuses the depth,
    // and the colour when the map stores colour (CreateMapRGB; a CreateMap
    // map stores none, which is reported once on stderr); rgb is CV_8UC3.
    void FuseDepthAndImage(GpuMat rgb, GpuMat depth, const Matrix4f& camToWorld);
    // The same with host images (the reference's signature takes cv::Mat):
    // uploaded into device buffers owned by this object, then fused.
    void FuseDepthAndImageHost(const void* rgb, size_t rgbStep, int rgbType, const void* depth, size_t depthStep,
                               int depthType, int rows, int cols, const Matrix4f& camToWorld);
    // vmap: CV_32FC4 camera-frame vertex map, created when empty
    void RenderScene(GpuMat& vmap, const Matrix4f& camToWorld);
    // host triangle soup owned by this object, valid until the next call
    int Polygonize(float*& verts_out, float*& norms_out);
    // triangle centroids (vmapping.cpp:134-147)
    std::vector<Vector3f> SurfacePoints();

    // CreateMap with the VoxelRGB format (no counterpart in the reference,
    // whose facade always uses vmap::Voxel)
    void CreateMapRGB(int numEntries, int numVoxels, float voxelSize);

    int width() const;
    int height() const;

#ifdef VMAPPING_HAVE_EIGEN
    // reference include/vmapping.h:22
    std::vector<Eigen::Vector3f> GetSurfacePoints() {
        const std::vector<Vector3f> p = SurfacePoints();
        std::vector<Eigen::Vector3f> out;
        out.reserve(p.size());
        for (const Vector3f& v : p) out.emplace_back(v[0], v[1], v[2]);
        return out;
    }
#else
    std::vector<Vector3f> GetSurfacePoints() { return SurfacePoints(); }
#endif

#ifdef VMAPPING_HAVE_OPENCV
    // reference include/vmapping.h:19 (rgb CV_8UC3, depth CV_32FC1 metres or
    // CV_16UC1 millimetres, host memory)
    void FuseDepthAndImage(cv::Mat rgb, cv::Mat depth, const Matrix4f& camToWorld) {
        FuseDepthAndImageHost(rgb.data, rgb.step[0], rgb.type(), depth.data, depth.step[0], depth.type(), depth.rows,
                              depth.cols, camToWorld);
    }
    // reference include/vmapping.h:20: an empty vmap is created (CV_32FC4,
    // HIP memory released with its last header), as the reference creates it
    void RenderScene(cv::cuda::GpuMat& vmap, const Matrix4f& camToWorld) {
        detail::ensure_output(vmap, height(), width(), CV_32FC4);
        GpuMat v(vmap);
        RenderScene(v, camToWorld);
    }
#endif

protected:
    struct VoxelizationImpl;
    std::unique_ptr<VoxelizationImpl> impl;
};

}  // namespace vmap

#endif
