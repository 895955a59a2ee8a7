// Drop-in replacement of the reference's include/image_proc.h:10-11.
#ifndef VMAPPING_IMAGE_PROC_H
#define VMAPPING_IMAGE_PROC_H

#include "map_struct.h"

namespace vmap {

// image_proc.cu:63-72: Phong shading of (vmap, nmap) into a CV_8UC4 image
inline void renderScene(const GpuMat vmap, const GpuMat nmap, GpuMat& image) {
    if (image.empty()) image.create(vmap.rows, vmap.cols, CV_8UC4);
    vm_check(vm_render_shading((const float*)vmap.data, (int64_t)vmap.step, (const float*)nmap.data,
                               (int64_t)nmap.step, (uint8_t*)image.data, (int64_t)image.step, vmap.cols, vmap.rows,
                               nullptr));
}

// image_proc.cu:95-104: vertex-map normals (last row and column not written)
inline void computeNormal(const GpuMat vmap, GpuMat& nmap) {
    if (nmap.empty()) nmap.create(vmap.rows, vmap.cols, vmap.type());
    vm_check(vm_compute_normals((const float*)vmap.data, (int64_t)vmap.step, (float*)nmap.data, (int64_t)nmap.step,
                                vmap.cols, vmap.rows, nullptr));
}

#ifdef VMAPPING_HAVE_OPENCV
// The reference's signatures (include/image_proc.h:10-11): an empty output is
// created as the reference creates it (image_proc.cu:64-65, :96-97), in HIP
// memory released with its last header.
inline void renderScene(const cv::cuda::GpuMat vmap, const cv::cuda::GpuMat nmap, cv::cuda::GpuMat& image) {
    detail::ensure_output(image, vmap.rows, vmap.cols, CV_8UC4);
    GpuMat out(image);
    renderScene(GpuMat(vmap), GpuMat(nmap), out);
}

inline void computeNormal(const cv::cuda::GpuMat vmap, cv::cuda::GpuMat& nmap) {
    detail::ensure_output(nmap, vmap.rows, vmap.cols, vmap.type());
    GpuMat out(nmap);
    computeNormal(GpuMat(vmap), out);
}
#endif

}  // namespace vmap

#endif
