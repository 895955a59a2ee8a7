// vmap::VoxelMapping (include/vmapping.h) over the C ABI: the reference's
// VoxelizationImpl (src/vmapping.cpp:10-148) owns the map and the mesh
// buffers and sequences the stages. Differences: no host synchronisation per
// fused frame (the visible count stays on the device), mesh buffers are
// allocated on the first Polygonize instead of in CreateMap, and errors throw
// std::runtime_error instead of exiting the process.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "map_struct.h"
#include "vmapping.h"
#include "vmapping_capi.h"

namespace vmap {

namespace {
constexpr int64_t kMaxTriangles = 20000000;  // MAX_VERTS_BUFFER, vmapping.cpp:6

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw std::runtime_error(std::string("vmapping: ") + what + ": " + hipGetErrorString(e));
}
}  // namespace

struct VoxelMapping::VoxelizationImpl {
    int width, height;
    float K[4];
    vm_map* map = nullptr;
    float* verts_gpu = nullptr;
    float* norms_gpu = nullptr;
    float* points_gpu = nullptr;
    std::vector<float> verts_cpu, norms_cpu;
    int fmt = -1;
    bool warnedNoColour = false;
    GpuMat rgbUp, depthUp;  // device copies of host images (FuseDepthAndImageHost)

    VoxelizationImpl(int w, int h, const Matrix3f& k) : width(w), height(h) {
        K[0] = k(0, 0);
        K[1] = k(1, 1);
        K[2] = k(0, 2);
        K[3] = k(1, 2);
    }
    ~VoxelizationImpl() {
        vm_map_destroy(map);
        if (verts_gpu) (void)hipFree(verts_gpu);
        if (norms_gpu) (void)hipFree(norms_gpu);
        if (points_gpu) (void)hipFree(points_gpu);
    }

    void create_map(int numEntries, int numVoxels, float voxelSize, int fmt) {
        vm_map_destroy(map);
        map = nullptr;
        vm_map_desc d;
        d.n_entry = numEntries;
        d.n_bucket = int(numVoxels * 0.9);  // vmapping.cpp:59
        d.n_block = numVoxels;
        d.voxel_size = voxelSize;
        d.trunc_dist = voxelSize * 4;
        d.depth_min = MIN_DEPTH;
        d.depth_max = MAX_DEPTH;
        d.voxel_format = fmt;
        d.max_weight = 255.f;
        vm_check(vm_map_create(&d, &map));
        this->fmt = fmt;
    }

    vm_map* need_map() {
        if (!map) throw std::runtime_error("vmapping: CreateMap must be called first");
        return map;
    }

    uint32_t create_mesh() {
        if (!verts_gpu) {
            hip_check(hipMalloc((void**)&verts_gpu, sizeof(float) * 9 * kMaxTriangles), "mesh buffer");
            hip_check(hipMalloc((void**)&norms_gpu, sizeof(float) * 9 * kMaxTriangles), "mesh buffer");
        }
        uint32_t blocks = 0, tris = 0;
        vm_check(vm_polygonize(need_map(), verts_gpu, norms_gpu, kMaxTriangles, nullptr, &blocks, &tris));
        return tris;
    }
};

VoxelMapping::VoxelMapping(int w, int h, const Matrix3f& K) : impl(new VoxelizationImpl(w, h, K)) {}
VoxelMapping::~VoxelMapping() = default;

void VoxelMapping::reset() { vm_check(vm_map_reset(impl->need_map(), nullptr)); }

void VoxelMapping::CreateMap(int numEntries, int numVoxels, float voxelSize) {
    impl->create_map(numEntries, numVoxels, voxelSize, VM_VOXEL_I16_U8);
}

void VoxelMapping::CreateMapRGB(int numEntries, int numVoxels, float voxelSize) {
    impl->create_map(numEntries, numVoxels, voxelSize, VM_VOXEL_I16_U8_RGB);
}

void VoxelMapping::FuseDepth(GpuMat depth, const Matrix4f& camToWorld) {
    const int fmt = depth.type() == CV_16UC1 ? VM_DEPTH_U16 : VM_DEPTH_F32;
    if (depth.type() != CV_16UC1 && depth.type() != CV_32FC1)
        throw std::invalid_argument("vmapping: depth must be CV_32FC1 or CV_16UC1");
    vm_check(vm_fuse_depth(impl->need_map(), depth.data, (int64_t)depth.step, depth.cols, depth.rows, fmt, 0.001f,
                           nullptr, 0, impl->K, camToWorld.data(), nullptr, nullptr));
}

void VoxelMapping::FuseDepthAndImage(GpuMat rgb, GpuMat depth, const Matrix4f& camToWorld) {
    const int fmt = depth.type() == CV_16UC1 ? VM_DEPTH_U16 : VM_DEPTH_F32;
    if (depth.type() != CV_16UC1 && depth.type() != CV_32FC1)
        throw std::invalid_argument("vmapping: depth must be CV_32FC1 or CV_16UC1");
    if (rgb.type() != CV_8UC3) throw std::invalid_argument("vmapping: rgb must be CV_8UC3");
    if (rgb.rows != depth.rows || rgb.cols != depth.cols)
        throw std::invalid_argument("vmapping: rgb and depth sizes differ");
    vm_map* m = impl->need_map();
    if (impl->fmt == VM_VOXEL_I16_U8 && !impl->warnedNoColour) {
        std::fprintf(stderr, "vmapping: FuseDepthAndImage on a map without colour (CreateMap stores vmap::Voxel); "
                             "the depth is fused, the colour is not (CreateMapRGB stores colour)\n");
        impl->warnedNoColour = true;
    }
    vm_check(vm_fuse_depth(m, depth.data, (int64_t)depth.step, depth.cols, depth.rows, fmt, 0.001f, rgb.data,
                           (int64_t)rgb.step, impl->K, camToWorld.data(), nullptr, nullptr));
}

void VoxelMapping::FuseDepthAndImageHost(const void* rgb, size_t rgbStep, int rgbType, const void* depth,
                                         size_t depthStep, int depthType, int rows, int cols,
                                         const Matrix4f& camToWorld) {
    if (!rgb || !depth || rows <= 0 || cols <= 0) throw std::invalid_argument("vmapping: empty host image");
    if (rgbType != CV_8UC3) throw std::invalid_argument("vmapping: rgb must be CV_8UC3");
    if (depthType != CV_16UC1 && depthType != CV_32FC1)
        throw std::invalid_argument("vmapping: depth must be CV_32FC1 or CV_16UC1");
    impl->rgbUp.create(rows, cols, CV_8UC3);
    impl->depthUp.create(rows, cols, depthType);
    const size_t dRow = elem_size(depthType) * (size_t)cols;
    hip_check(hipMemcpy2D(impl->rgbUp.data, impl->rgbUp.step, rgb, rgbStep, 3 * (size_t)cols, rows,
                          hipMemcpyHostToDevice),
              "rgb upload");
    hip_check(hipMemcpy2D(impl->depthUp.data, impl->depthUp.step, depth, depthStep, dRow, rows,
                          hipMemcpyHostToDevice),
              "depth upload");
    FuseDepthAndImage(impl->rgbUp, impl->depthUp, camToWorld);
}

int VoxelMapping::width() const { return impl->width; }
int VoxelMapping::height() const { return impl->height; }

// vmapping.cpp:83-104: visibility, then the raycast unless nothing is
// visible; the gate is evaluated on the device (no host round trip).
void VoxelMapping::RenderScene(GpuMat& vmap, const Matrix4f& camToWorld) {
    if (vmap.empty()) vmap.create(impl->height, impl->width, CV_32FC4);
    vm_check(vm_render_scene(impl->need_map(), (float*)vmap.data, (int64_t)vmap.step, vmap.cols, vmap.rows, impl->K,
                             camToWorld.data(), nullptr));
}

// vmapping.cpp:119-132
int VoxelMapping::Polygonize(float*& verts_out, float*& norms_out) {
    const uint32_t n = impl->create_mesh();
    if (n == 0) return 0;
    impl->verts_cpu.resize((size_t)n * 9);
    impl->norms_cpu.resize((size_t)n * 9);
    hip_check(hipMemcpy(impl->verts_cpu.data(), impl->verts_gpu, sizeof(float) * 9 * n, hipMemcpyDeviceToHost),
              "mesh copy");
    hip_check(hipMemcpy(impl->norms_cpu.data(), impl->norms_gpu, sizeof(float) * 9 * n, hipMemcpyDeviceToHost),
              "mesh copy");
    verts_out = impl->verts_cpu.data();
    norms_out = impl->norms_cpu.data();
    return (int)n;
}

// vmapping.cpp:134-147
std::vector<Vector3f> VoxelMapping::SurfacePoints() {
    const uint32_t n = impl->create_mesh();
    if (n == 0) return {};
    if (!impl->points_gpu)
        hip_check(hipMalloc((void**)&impl->points_gpu, sizeof(float) * 3 * kMaxTriangles), "points buffer");
    vm_check(vm_surface_points(impl->verts_gpu, impl->points_gpu, n, nullptr));
    std::vector<Vector3f> out(n);
    hip_check(hipMemcpy(out.data(), impl->points_gpu, sizeof(float) * 3 * n, hipMemcpyDeviceToHost), "points copy");
    return out;
}

}  // namespace vmap
