// Test stand-in for <opencv2/core/cuda.hpp>: OpenCV is not installed here.
// It declares only the names the drop-in headers use, with OpenCV's public
// signatures and values: the CV_* type macros, cv::Mat (a host image
// header) and cv::cuda::GpuMat (a reference-counted device image header with
// a pluggable Allocator, whose last release calls Allocator::free). It lets
// tests/cpp/dropin_reference_types.cpp compile and run the reference-typed
// overloads without OpenCV.
#ifndef VMAPPING_TEST_STUB_OPENCV_CUDA_HPP
#define VMAPPING_TEST_STUB_OPENCV_CUDA_HPP

#include <cstddef>

#define CV_8U 0
#define CV_16U 2
#define CV_32F 5
#define CV_CN_SHIFT 3
#define CV_MAKETYPE(depth, cn) ((depth) + (((cn) - 1) << CV_CN_SHIFT))
#define CV_8UC3 CV_MAKETYPE(CV_8U, 3)
#define CV_8UC4 CV_MAKETYPE(CV_8U, 4)
#define CV_16UC1 CV_MAKETYPE(CV_16U, 1)
#define CV_32FC1 CV_MAKETYPE(CV_32F, 1)
#define CV_32FC4 CV_MAKETYPE(CV_32F, 4)

typedef unsigned char uchar;

namespace cv {

class Mat {
public:
    enum { MAGIC_VAL = 0x42FF0000, TYPE_MASK = 0x00000FFF };
    struct MatStep {
        size_t p[2] = {0, 0};
        size_t operator[](int i) const { return p[i]; }
        operator size_t() const { return p[0]; }
    };
    Mat() = default;
    Mat(int r, int c, int t, void* d, size_t s) : flags(MAGIC_VAL | (t & TYPE_MASK)), rows(r), cols(c),
                                                  data(static_cast<uchar*>(d)) {
        step.p[0] = s;
    }
    int type() const { return flags & TYPE_MASK; }
    bool empty() const { return data == nullptr; }

    int flags = MAGIC_VAL;
    int rows = 0, cols = 0;
    uchar* data = nullptr;
    MatStep step;
};

namespace cuda {

class GpuMat {
public:
    class Allocator {
    public:
        virtual ~Allocator() {}
        virtual bool allocate(GpuMat* mat, int rows, int cols, size_t elemSize) = 0;
        virtual void free(GpuMat* mat) = 0;
    };

    GpuMat() = default;
    // a header of user-allocated data (never freed by the header)
    GpuMat(int r, int c, int t, void* d, size_t s)
        : flags(Mat::MAGIC_VAL | (t & Mat::TYPE_MASK)), rows(r), cols(c), step(s), data(static_cast<uchar*>(d)),
          datastart(static_cast<uchar*>(d)), dataend(static_cast<uchar*>(d) + s * (size_t)r) {}
    GpuMat(const GpuMat& m)
        : flags(m.flags), rows(m.rows), cols(m.cols), step(m.step), data(m.data), refcount(m.refcount),
          datastart(m.datastart), dataend(m.dataend), allocator(m.allocator) {
        if (refcount) ++*refcount;
    }
    GpuMat& operator=(const GpuMat& m) {
        if (this != &m) {
            GpuMat t(m);
            release();
            flags = t.flags; rows = t.rows; cols = t.cols; step = t.step; data = t.data;
            refcount = t.refcount; datastart = t.datastart; dataend = t.dataend; allocator = t.allocator;
            t.refcount = nullptr;
        }
        return *this;
    }
    ~GpuMat() { release(); }
    void release() {
        if (refcount && --*refcount == 0 && allocator) allocator->free(this);
        data = datastart = nullptr;
        dataend = nullptr;
        refcount = nullptr;
        rows = cols = 0;
        step = 0;
    }
    bool empty() const { return data == nullptr; }
    int type() const { return flags & Mat::TYPE_MASK; }

    int flags = Mat::MAGIC_VAL;
    int rows = 0, cols = 0;
    size_t step = 0;
    uchar* data = nullptr;
    int* refcount = nullptr;
    uchar* datastart = nullptr;
    const uchar* dataend = nullptr;
    Allocator* allocator = nullptr;
};

}  // namespace cuda
}  // namespace cv

#endif
