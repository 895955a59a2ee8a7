// Lightweight value types of the drop-in C++ API.
//
// The reference's headers use Eigen (Matrix3f, Matrix4f, Vector3f, Vector3i)
// and OpenCV (cv::cuda::GpuMat, cv::Mat). These types carry the same data
// layout (column-major matrices; pitched images with data/step/rows/cols), so
// the drop-in headers compile without Eigen or OpenCV. When those libraries
// are available the types convert implicitly from the Eigen / OpenCV ones.
#ifndef VMAPPING_VM_TYPES_H
#define VMAPPING_VM_TYPES_H

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

#if defined(__has_include)
#if __has_include(<Eigen/Core>)
#include <Eigen/Core>
#define VMAPPING_HAVE_EIGEN 1
#endif
#if __has_include(<opencv2/core/cuda.hpp>)
#include <opencv2/core/cuda.hpp>
#define VMAPPING_HAVE_OPENCV 1
#endif
#endif

namespace vmap {

struct Vector3f {
    float v[3] = {0, 0, 0};
    Vector3f() = default;
    Vector3f(float x, float y, float z) : v{x, y, z} {}
    float& operator[](int i) { return v[i]; }
    float operator[](int i) const { return v[i]; }
    float& operator()(int i) { return v[i]; }
    float operator()(int i) const { return v[i]; }
    const float* data() const { return v; }
};

struct Vector3i {
    int v[3] = {0, 0, 0};
    int& operator[](int i) { return v[i]; }
    int operator[](int i) const { return v[i]; }
    int& operator()(int i) { return v[i]; }
    int operator()(int i) const { return v[i]; }
};

struct Vector3b {
    unsigned char v[3] = {0, 0, 0};
    unsigned char& operator[](int i) { return v[i]; }
    unsigned char operator[](int i) const { return v[i]; }
};

// column-major N x N float matrix, like Eigen::Matrix<float, N, N>
template <int N>
struct MatrixNf {
    float m[N * N];
    MatrixNf() { setIdentity(); }
    void setIdentity() {
        for (int i = 0; i < N * N; ++i) m[i] = (i % (N + 1) == 0) ? 1.f : 0.f;
    }
    static MatrixNf Identity() { return MatrixNf(); }
    float& operator()(int r, int c) { return m[c * N + r]; }
    float operator()(int r, int c) const { return m[c * N + r]; }
    const float* data() const { return m; }
    float* data() { return m; }
#ifdef VMAPPING_HAVE_EIGEN
    MatrixNf(const Eigen::Matrix<float, N, N>& e) { std::memcpy(m, e.data(), sizeof(m)); }
#endif
};
using Matrix3f = MatrixNf<3>;
using Matrix4f = MatrixNf<4>;

// OpenCV-compatible element type codes used by the drop-in API
enum : int { CV_8UC3 = 16, CV_8UC4 = 24, CV_16UC1 = 2, CV_32FC1 = 5, CV_32FC4 = 29 };

inline size_t elem_size(int type) {
    switch (type) {
        case CV_8UC3: return 3;
        case CV_8UC4: return 4;
        case CV_16UC1: return 2;
        case CV_32FC1: return 4;
        case CV_32FC4: return 16;
        default: throw std::invalid_argument("unsupported image type " + std::to_string(type));
    }
}

// Pitched 2-D image in device memory (the part of cv::cuda::GpuMat the API
// uses). A GpuMat made by create() owns its memory and frees it; a GpuMat
// wrapping external memory does not. Copies share the buffer (like OpenCV's
// reference counting, without the count: the creator must outlive copies).
struct GpuMat {
    void* data = nullptr;
    size_t step = 0;  // bytes per row
    int rows = 0, cols = 0;
    int type_ = CV_32FC1;
    bool owns = false;

    GpuMat() = default;
    GpuMat(int r, int c, int t) { create(r, c, t); }
    GpuMat(int r, int c, int t, void* d, size_t s) : data(d), step(s), rows(r), cols(c), type_(t) {}
    GpuMat(const GpuMat& o) : data(o.data), step(o.step), rows(o.rows), cols(o.cols), type_(o.type_) {}
    GpuMat& operator=(const GpuMat& o) {
        if (this != &o) {
            release();
            data = o.data; step = o.step; rows = o.rows; cols = o.cols; type_ = o.type_;
        }
        return *this;
    }
    ~GpuMat() { release(); }
#ifdef VMAPPING_HAVE_OPENCV
    GpuMat(const cv::cuda::GpuMat& g) : data(g.data), step(g.step), rows(g.rows), cols(g.cols), type_(g.type()) {}
#endif
    void create(int r, int c, int t) {
        if (data && r == rows && c == cols && t == type_) return;
        release();
        step = ((elem_size(t) * (size_t)c + 255) / 256) * 256;
        if (hipMalloc(&data, step * (size_t)r) != hipSuccess) throw std::runtime_error("hipMalloc failed");
        rows = r; cols = c; type_ = t; owns = true;
    }
    void release() {
        if (owns && data) (void)hipFree(data);
        data = nullptr;
        owns = false;
    }
    bool empty() const { return data == nullptr; }
    int type() const { return type_; }
};

}  // namespace vmap

#endif
