// Drives the drop-in C++ API with the reference's own argument types
// (include/vmapping.h:13-22, include/image_proc.h:10-11,
// src/device_functions.h:35-43 of the reference): host cv::Mat images,
// cv::cuda::GpuMat outputs that start empty, Eigen matrices and the
// std::vector<Eigen::Vector3f> of GetSurfacePoints. Built against the
// stand-in OpenCV / Eigen headers of tests/cpp/stub (neither library is
// installed here). The scene is a fronto-parallel plane 1.2037 m in front of
// the camera in one colour, so every result has a closed form. Prints
// "REFTYPES OK" on success, exits non-zero otherwise.
#include <opencv2/core/cuda.hpp>
#include <Eigen/Core>
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "device_functions.h"
#include "image_proc.h"
#include "map_struct.h"
#include "vmapping.h"

#if !defined(VMAPPING_HAVE_OPENCV) || !defined(VMAPPING_HAVE_EIGEN)
#error "the stand-in OpenCV / Eigen headers were not picked up"
#endif

static int failures = 0;
#define EXPECT(cond, ...)                      \
    do {                                       \
        if (!(cond)) {                         \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");                 \
            ++failures;                        \
        }                                      \
    } while (0)

template <class T>
static std::vector<T> download(const cv::cuda::GpuMat& g, int channels) {
    std::vector<T> h((size_t)g.rows * g.cols * channels);
    const size_t row = sizeof(T) * channels * (size_t)g.cols;
    hipMemcpy2D(h.data(), row, g.data, g.step, row, g.rows, hipMemcpyDeviceToHost);
    return h;
}

int main() {
    const int W = 160, H = 120;
    const float D = 1.2037f;
    const unsigned char colour[3] = {10, 200, 30};
    Eigen::Matrix3f K = Eigen::Matrix3f::Identity();
    K(0, 0) = 140.f; K(1, 1) = 140.f; K(0, 2) = 79.5f; K(1, 2) = 59.5f;
    const Eigen::Matrix4f pose = Eigen::Matrix4f::Identity();

    std::vector<float> depth((size_t)W * H, D);
    std::vector<unsigned char> rgb((size_t)W * H * 3);
    for (size_t i = 0; i < (size_t)W * H; ++i)
        for (int c = 0; c < 3; ++c) rgb[i * 3 + c] = colour[c];
    cv::Mat depthHost(H, W, CV_32FC1, depth.data(), W * sizeof(float));
    cv::Mat rgbHost(H, W, CV_8UC3, rgb.data(), W * 3);

    // facade: host images in, empty cv::cuda::GpuMat outputs created by the library
    vmap::VoxelMapping vm(W, H, K);
    vm.CreateMapRGB(80000, 65536, 0.01f);
    for (int i = 0; i < 3; ++i) vm.FuseDepthAndImage(rgbHost, depthHost, pose);
    cv::cuda::GpuMat vmapG, nmapG, shadedG;
    vm.RenderScene(vmapG, pose);
    vmap::computeNormal(vmapG, nmapG);
    vmap::renderScene(vmapG, nmapG, shadedG);
    hipDeviceSynchronize();
    EXPECT(!vmapG.empty() && vmapG.rows == H && vmapG.cols == W && vmapG.type() == CV_32FC4, "vmap created");
    EXPECT(!nmapG.empty() && nmapG.type() == CV_32FC4, "nmap created");
    EXPECT(!shadedG.empty() && shadedG.type() == CV_8UC4, "shaded image created");
    EXPECT(vmapG.refcount && *vmapG.refcount == 1, "vmap is reference counted");
    {
        cv::cuda::GpuMat alias = vmapG;  // a copy shares the buffer
        EXPECT(alias.data == vmapG.data && *vmapG.refcount == 2, "copies share the buffer");
    }
    EXPECT(*vmapG.refcount == 1, "released copy");
    const std::vector<float> v = download<float>(vmapG, 4), n = download<float>(nmapG, 4);
    const std::vector<unsigned char> img = download<unsigned char>(shadedG, 4);
    const size_t c = (size_t)(H / 2) * W + W / 2;
    EXPECT(!std::isnan(v[c * 4]) && std::fabs(v[c * 4 + 2] - D) < 1e-4f, "centre depth %f", v[c * 4 + 2]);
    EXPECT(std::fabs(n[c * 4 + 2] - 1.f) < 1e-4f, "centre normal z %f", n[c * 4 + 2]);
    EXPECT(img[c * 4 + 3] == 255 && img[c * 4] > 60, "shaded centre %d", (int)img[c * 4]);

    // a vertex map given by the caller is written in place
    cv::cuda::GpuMat own = vmapG;
    vm.RenderScene(own, pose);
    EXPECT(own.data == vmapG.data, "non-empty output kept");

    float *verts = nullptr, *norms = nullptr;
    const int T = vm.Polygonize(verts, norms);
    EXPECT(T > 100, "triangles %d", T);
    const std::vector<Eigen::Vector3f> pts = vm.GetSurfacePoints();
    EXPECT((int)pts.size() == T, "surface points %zu of %d", pts.size(), T);
    if (T > 0 && !pts.empty()) {
        for (int k = 0; k < 3; ++k) {
            const float cen = ((verts[k] + verts[3 + k]) + verts[6 + k]) / 3.0f;
            EXPECT(pts[0](k) == cen, "centroid %d: %f vs %f", k, pts[0](k), cen);
        }
    }

    // stage API: colour fused from cv::cuda::GpuMat views, z-range maps created empty
    vmap::GpuMat dDev(H, W, CV_32FC1), cDev(H, W, CV_8UC3);
    hipMemcpy2D(dDev.data, dDev.step, depth.data(), W * 4, W * 4, H, hipMemcpyHostToDevice);
    hipMemcpy2D(cDev.data, cDev.step, rgb.data(), W * 3, W * 3, H, hipMemcpyHostToDevice);
    const cv::cuda::GpuMat depthG = dDev, rgbG = cDev;
    vmap::MapStruct<vmap::VoxelRGB> ms;
    ms.create(4096 + 1024, 4096, 2048, 0.01f, 0.04f);
    const int nv = vmap::FuseDepthAndImage(ms, rgbG, depthG, pose, K);
    EXPECT(nv > 0, "visible blocks %d", nv);
    cv::cuda::GpuMat zx, zy;
    unsigned int nrb = 0;
    vmap::ProjectRenderingBlocks(ms, (unsigned)nv, nrb, zx, zy, nullptr, pose, K, W / 8, H / 8);
    hipDeviceSynchronize();
    EXPECT(nrb > 0 && zx.rows == H / 8 && zx.cols == W / 8 && zy.rows == H / 8, "rendering blocks %u", nrb);
    // the voxel (0, 0, 120) of block (0, 0, 15) sits 3.7 mm in front of the plane
    std::vector<vmap::HashEntry> table(ms.nEntry);
    vmap::vm_check(vm_map_download(ms.map, 0, table.data(), (int64_t)sizeof(vmap::HashEntry) * ms.nEntry));
    int ptr = -1;
    for (const vmap::HashEntry& e : table)
        if (e.ptr >= 0 && e.pos[0] == 0 && e.pos[1] == 0 && e.pos[2] == 15) ptr = e.ptr;
    EXPECT(ptr >= 0, "block (0, 0, 15) allocated");
    if (ptr >= 0) {
        vmap::VoxelRGB vox;
        hipMemcpy(&vox, ms.voxelBlock + ptr, sizeof(vox), hipMemcpyDeviceToHost);
        const float sdf = vox.sdf / 32767.f, expect = (D - 1.20f) / 0.04f;
        EXPECT(std::fabs(sdf - expect) < 1e-3f && vox.wt == 1, "voxel sdf %f (%f) weight %d", sdf, expect, vox.wt);
        for (int k = 0; k < 3; ++k)
            EXPECT(vox.rgb[k] == colour[k], "voxel colour %d: %d", k, (int)vox.rgb[k]);
    }
    ms.release();

    if (failures) {
        std::printf("REFTYPES FAILED (%d)\n", failures);
        return 1;
    }
    std::printf("REFTYPES OK triangles=%d\n", T);
    return 0;
}
