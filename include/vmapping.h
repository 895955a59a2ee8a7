// Drop-in replacement of the reference's include/vmapping.h:10-27
// (vmap::VoxelMapping), implemented in libvmapping_amd.so over the C ABI.
// Eigen / OpenCV argument types are replaced by the layout-compatible types
// of vm_types.h (implicitly constructible from Eigen / OpenCV when present).
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
    // A stub in the reference (vmapping.cpp:167-170). Here it fuses the
    // colour too when the map stores colour (CreateMapRGB); rgb is CV_8UC3.
    void FuseDepthAndImage(GpuMat rgb, GpuMat depth, const Matrix4f& camToWorld);
    // vmap: CV_32FC4 camera-frame vertex map, created when empty
    void RenderScene(GpuMat& vmap, const Matrix4f& camToWorld);
    // host triangle soup owned by this object, valid until the next call
    int Polygonize(float*& verts_out, float*& norms_out);
    std::vector<Vector3f> GetSurfacePoints();

    // CreateMap with the VoxelRGB format (no counterpart in the reference,
    // whose facade always uses vmap::Voxel)
    void CreateMapRGB(int numEntries, int numVoxels, float voxelSize);

protected:
    struct VoxelizationImpl;
    std::unique_ptr<VoxelizationImpl> impl;
};

}  // namespace vmap

#endif
