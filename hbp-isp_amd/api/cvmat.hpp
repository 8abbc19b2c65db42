// cv::Mat for the drop-in classes.  With OpenCV on the include path this is OpenCV itself;
// without it, a minimal owning 2-D matrix with the members the HBP-ISP call surface uses
// (rows, cols, data, type(), channels(), clone(), at<>(), ptr<>(), empty()).
#pragma once

#if __has_include(<opencv2/core.hpp>)
#include <opencv2/core.hpp>
#else
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <vector>

#ifndef CV_8U
#define CV_8U 0
#define CV_32S 4
#define CV_32F 5
#define CV_MAKETYPE(depth, cn) ((depth) + (((cn)-1) << 3))
#define CV_8UC1 CV_MAKETYPE(CV_8U, 1)
#define CV_8UC3 CV_MAKETYPE(CV_8U, 3)
#define CV_32SC1 CV_MAKETYPE(CV_32S, 1)
#define CV_32FC1 CV_MAKETYPE(CV_32F, 1)
#define CV_32FC3 CV_MAKETYPE(CV_32F, 3)
#endif

namespace cv {

typedef unsigned char uchar;

template <typename T, int n>
struct Vec {
    T val[n];
    T &operator[](int i) { return val[i]; }
    const T &operator[](int i) const { return val[i]; }
};
typedef Vec<float, 3> Vec3f;
typedef Vec<uchar, 3> Vec3b;

struct Point {
    int x = 0, y = 0;
    Point() = default;
    Point(int x_, int y_) : x(x_), y(y_) {}
};

class Mat {
public:
    int rows = 0, cols = 0;
    uchar *data = nullptr;

    Mat() = default;
    Mat(int r, int c, int type) { create(r, c, type); }
    void create(int r, int c, int type) {
        rows = r;
        cols = c;
        type_ = type;
        buf_ = std::make_shared<std::vector<uchar>>((size_t)r * c * elemSize());
        data = buf_->data();
    }
    int type() const { return type_; }
    int depth() const { return type_ & 7; }
    int channels() const { return (type_ >> 3) + 1; }
    size_t elemSize1() const { return depth() == CV_8U ? 1 : 4; }
    size_t elemSize() const { return elemSize1() * channels(); }
    size_t total() const { return (size_t)rows * cols; }
    bool empty() const { return data == nullptr || total() == 0; }
    bool isContinuous() const { return true; }
    Mat clone() const {
        Mat m(rows, cols, type_);
        if (data) std::memcpy(m.data, data, total() * elemSize());
        return m;
    }
    static Mat zeros(int r, int c, int type) {
        Mat m(r, c, type);
        return m;
    }
    template <typename T>
    T *ptr(int y = 0) {
        return (T *)(data + (size_t)y * cols * elemSize());
    }
    template <typename T>
    const T *ptr(int y = 0) const {
        return (const T *)(data + (size_t)y * cols * elemSize());
    }
    template <typename T>
    T &at(int y, int x) {
        return ((T *)(data + (size_t)y * cols * elemSize()))[x];
    }
    template <typename T>
    const T &at(int y, int x) const {
        return ((const T *)(data + (size_t)y * cols * elemSize()))[x];
    }
    template <typename T>
    T &at(Point p) {
        return at<T>(p.y, p.x);
    }

private:
    int type_ = CV_8UC1;
    std::shared_ptr<std::vector<uchar>> buf_;
};

}  // namespace cv
#endif
