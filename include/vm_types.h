// Lightweight value types of the drop-in C++ API.
//
// The reference's headers use Eigen (Matrix3f, Matrix4f, Vector3f, Vector3i)
// and OpenCV (cv::cuda::GpuMat, cv::Mat). These types carry the same data
// layout (column-major matrices; pitched images with data/step/rows/cols), so
// the drop-in headers compile without Eigen or OpenCV. When those libraries
// are available the types convert implicitly from the Eigen / OpenCV ones,
// and the headers add overloads taking the reference's own argument types
// (cv::cuda::GpuMat& outputs, cv::Mat host images, Eigen vectors).
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
#ifdef VMAPPING_HAVE_EIGEN
    Vector3f(const Eigen::Vector3f& e) : v{e(0), e(1), e(2)} {}
    operator Eigen::Vector3f() const { return Eigen::Vector3f(v[0], v[1], v[2]); }
#endif
};
static_assert(sizeof(Vector3f) == 12, "Vector3f is three packed floats, like Eigen::Vector3f");

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

// OpenCV's element type codes used by the drop-in API (CV_MAKETYPE(depth,
// channels) values). OpenCV defines them as macros with these values, so they
// are declared here only without it.
#ifndef VMAPPING_HAVE_OPENCV
enum : int { CV_8UC3 = 16, CV_8UC4 = 24, CV_16UC1 = 2, CV_32FC1 = 5, CV_32FC4 = 29 };
#endif

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
    // a non-owning view of an OpenCV device image, and back
    GpuMat(const cv::cuda::GpuMat& g) : data(g.data), step(g.step), rows(g.rows), cols(g.cols), type_(g.type()) {}
    operator cv::cuda::GpuMat() const { return cv::cuda::GpuMat(rows, cols, type_, data, step); }
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

#ifdef VMAPPING_HAVE_OPENCV
namespace detail {

// cv::cuda::GpuMat::Allocator over hipMalloc / hipFree. Where the reference
// calls GpuMat::create on an empty output (image_proc.cu:64-65, :96-97;
// vmapping.cpp:90-91), the drop-in gives the output a buffer of this
// allocator with OpenCV's reference count set to one: copies of the header
// share it, and the last one to be released returns it through free().
struct HipGpuMatAllocator final : public cv::cuda::GpuMat::Allocator {
    bool allocate(cv::cuda::GpuMat* mat, int rows, int cols, size_t elemSize) override {
        const size_t step = ((elemSize * (size_t)cols + 255) / 256) * 256;
        void* p = nullptr;
        if (hipMalloc(&p, step * (size_t)rows) != hipSuccess) return false;
        mat->data = mat->datastart = static_cast<unsigned char*>(p);
        mat->dataend = mat->datastart + step * (size_t)rows;
        mat->step = step;
        mat->refcount = new int(1);
        return true;
    }
    void free(cv::cuda::GpuMat* mat) override {
        (void)hipFree(mat->datastart);
        delete mat->refcount;
    }
};

inline cv::cuda::GpuMat::Allocator* hip_gpumat_allocator() {
    static HipGpuMatAllocator a;
    return &a;
}

// An empty `out` becomes a new rows x cols image of `type` (HIP memory);
// a non-empty one must already have that size and type.
inline void ensure_output(cv::cuda::GpuMat& out, int rows, int cols, int type) {
    if (out.empty()) {
        cv::cuda::GpuMat::Allocator* a = hip_gpumat_allocator();
        out.allocator = a;
        out.flags = cv::Mat::MAGIC_VAL | (type & cv::Mat::TYPE_MASK);
        out.rows = rows;
        out.cols = cols;
        if (!a->allocate(&out, rows, cols, elem_size(type))) {
            out.rows = out.cols = 0;
            throw std::runtime_error("vmapping: hipMalloc of an output image failed");
        }
        return;
    }
    if (out.rows != rows || out.cols != cols || out.type() != type)
        throw std::invalid_argument("vmapping: output image has the wrong size or type");
}

}  // namespace detail
#endif

}  // namespace vmap

#endif
