// Exercises the drop-in C++ API (vmapping.h, map_struct.h, device_functions.h,
// image_proc.h) the way the reference's example program uses it: create a
// map, fuse depth frames, render, compute normals, extract a mesh. The scene
// is a fronto-parallel plane 1.2037 m in front of the camera, so every result
// has a closed form. Prints "DROPIN OK" on success, exits non-zero otherwise.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <vector>

#include "device_functions.h"
#include "image_proc.h"
#include "map_struct.h"
#include "vmapping.h"

using namespace vmap;

static int failures = 0;
#define EXPECT(cond, ...)                      \
    do {                                       \
        if (!(cond)) {                         \
            std::printf("FAIL: " __VA_ARGS__); \
            std::printf("\n");                 \
            ++failures;                        \
        }                                      \
    } while (0)

static GpuMat upload_depth(const std::vector<float>& h, int W, int Hh) {
    GpuMat g(Hh, W, CV_32FC1);
    hipMemcpy2D(g.data, g.step, h.data(), W * sizeof(float), W * sizeof(float), Hh, hipMemcpyHostToDevice);
    return g;
}

static std::vector<float> download4(const GpuMat& g) {
    std::vector<float> h((size_t)g.rows * g.cols * 4);
    hipMemcpy2D(h.data(), g.cols * 16, g.data, g.step, g.cols * 16, g.rows, hipMemcpyDeviceToHost);
    return h;
}

int main() {
    const int W = 160, H = 120;
    const float D = 1.2037f;
    Matrix3f K;
    K(0, 0) = 140.f; K(1, 1) = 140.f; K(0, 2) = 79.5f; K(1, 2) = 59.5f;
    Matrix4f pose;  // identity camera-to-world
    std::vector<float> depth((size_t)W * H, D);
    GpuMat d = upload_depth(depth, W, H);

    // facade
    VoxelMapping vm(W, H, K);
    vm.CreateMap(80000, 65536, 0.01f);
    for (int i = 0; i < 3; ++i) vm.FuseDepth(d, pose);
    GpuMat vmap, nmap, shaded;
    vm.RenderScene(vmap, pose);
    computeNormal(vmap, nmap);
    renderScene(vmap, nmap, shaded);
    hipDeviceSynchronize();
    std::vector<float> v = download4(vmap), n = download4(nmap);
    const size_t c = ((size_t)(H / 2) * W + W / 2) * 4;
    EXPECT(!std::isnan(v[c]) && std::fabs(v[c + 2] - D) < 1e-4f, "centre depth %f", v[c + 2]);
    EXPECT(std::fabs(n[c + 2] - 1.f) < 1e-4f, "centre normal z %f", n[c + 2]);
    float *verts = nullptr, *norms = nullptr;
    int T = vm.Polygonize(verts, norms);
    EXPECT(T > 100, "triangles %d", T);
    double zmean = 0;
    for (int t = 0; t < T; ++t) zmean += (verts[t * 9 + 2] + verts[t * 9 + 5] + verts[t * 9 + 8]) / 3.0;
    zmean /= (T > 0 ? T : 1);
    EXPECT(std::fabs(zmean - D) < 2e-3, "mean mesh z %f", zmean);
    std::vector<Vector3f> pts = vm.GetSurfacePoints();
    EXPECT((int)pts.size() == T, "surface points %zu", pts.size());

    // stage API on a float-TSDF MapStruct
    MapStruct<VoxelF32> ms;
    ms.create(4096 + 1024, 4096, 2048, 0.01f, 0.04f);
    int nv = FuseImage(ms, d, pose, K);
    EXPECT(nv > 0, "visible blocks %d", nv);
    EXPECT(CheckVisibleBlocks(ms, W, H, pose, K) == nv, "visibility recount");
    GpuMat vm2(H, W, CV_32FC4), zx, zy;
    RenderScene(ms, vm2, zx, zy, pose, K);
    hipDeviceSynchronize();
    std::vector<float> v2 = download4(vm2);
    EXPECT(std::fabs(v2[c + 2] - D) < 1e-4f, "stage raycast depth %f", v2[c + 2]);
    unsigned int blocks = 0, tris = 0;
    float *vg = nullptr, *ng = nullptr;
    hipMalloc((void**)&vg, sizeof(float) * 9 * 200000);
    hipMalloc((void**)&ng, sizeof(float) * 9 * 200000);
    Polygonize(ms, blocks, tris, vg, ng, 200000);
    EXPECT(tris > 100 && blocks > 0, "stage mesh %u triangles %u blocks", tris, blocks);
    GpuMat zrx(H / 8, W / 8, CV_32FC1), zry(H / 8, W / 8, CV_32FC1);
    unsigned int nrb = 0;
    Matrix4f wc;  // identity world-to-camera
    ProjectRenderingBlocks(ms, (unsigned)nv, nrb, zrx, zry, nullptr, wc, K);
    EXPECT(nrb > 0, "rendering blocks %u", nrb);
    unsigned int host = 0;
    ms.getVisibleBlockCount(host);
    EXPECT((int)host == nv, "visible count %u", host);
    ms.release();
    hipFree(vg);
    hipFree(ng);

    if (failures) {
        std::printf("DROPIN FAILED (%d)\n", failures);
        return 1;
    }
    std::printf("DROPIN OK triangles=%d\n", T);
    return 0;
}
