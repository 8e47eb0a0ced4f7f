// Test double of the part of OpenCV's cv::Mat that the library's optional cv::Mat overloads use
// (include/stereo/*.h under LUMEN_WITH_OPENCV). OpenCV is not installed here; this lets the
// overloads compile and run in tests/cpp/test_cvmat_overloads.cpp. It follows OpenCV's
// documented semantics for what it has: 2-D matrices of 8U/16S/32F elements with 1-4
// channels, row pitch `step` in bytes, create() that reallocates only when the size or type
// changes, clone() that returns a continuous copy, and ROIs (Mat(Mat, Rect)) that share data
// with a larger step. It is not OpenCV.
#pragma once
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#define CV_8U 0
#define CV_16S 3
#define CV_32F 5
#define CV_MAKETYPE(depth, cn) ((depth) + (((cn) - 1) << 3))
#define CV_8UC1 CV_MAKETYPE(CV_8U, 1)
#define CV_8UC3 CV_MAKETYPE(CV_8U, 3)
#define CV_16SC1 CV_MAKETYPE(CV_16S, 1)
#define CV_32FC1 CV_MAKETYPE(CV_32F, 1)
#define CV_32FC2 CV_MAKETYPE(CV_32F, 2)

namespace cv {

struct Size {
    int width = 0, height = 0;
    Size() = default;
    Size(int w, int h) : width(w), height(h) {}
    bool operator==(const Size& o) const { return width == o.width && height == o.height; }
    bool operator!=(const Size& o) const { return !(*this == o); }
};

struct Rect {
    int x = 0, y = 0, width = 0, height = 0;
    Rect(int x_, int y_, int w, int h) : x(x_), y(y_), width(w), height(h) {}
};

class Mat {
public:
    int rows = 0, cols = 0;
    uint8_t* data = nullptr;
    size_t step = 0;  // bytes between rows

    Mat() = default;
    Mat(int r, int c, int type) { create(r, c, type); }
    Mat(Size s, int type) { create(s.height, s.width, type); }
    Mat(const Mat& m, const Rect& roi)
        : rows(roi.height), cols(roi.width), step(m.step), type_(m.type_), buf_(m.buf_) {
        data = m.data + static_cast<size_t>(roi.y) * m.step + static_cast<size_t>(roi.x) * m.elemSize();
    }

    static int depthBytes(int type) {
        switch (type & 7) {
            case CV_8U: return 1;
            case CV_16S: return 2;
            default: return 4;
        }
    }
    int type() const { return type_; }
    int channels() const { return (type_ >> 3) + 1; }
    size_t elemSize() const { return static_cast<size_t>(depthBytes(type_)) * channels(); }
    bool empty() const { return data == nullptr || rows == 0 || cols == 0; }
    Size size() const { return Size(cols, rows); }
    bool isContinuous() const { return rows <= 1 || step == cols * elemSize(); }

    void create(int r, int c, int type) {
        if (data && r == rows && c == cols && type == type_) return;
        rows = r;
        cols = c;
        type_ = type;
        step = static_cast<size_t>(c) * elemSize();
        buf_ = std::make_shared<std::vector<uint8_t>>(step * r);
        data = buf_->data();
    }
    void create(Size s, int type) { create(s.height, s.width, type); }

    Mat clone() const {
        Mat m(rows, cols, type_);
        for (int y = 0; y < rows; ++y)
            std::memcpy(m.data + y * m.step, data + y * step, cols * elemSize());
        return m;
    }

    template <typename T>
    T& at(int y, int x) { return reinterpret_cast<T*>(data + y * step)[x]; }
    template <typename T>
    const T& at(int y, int x) const { return reinterpret_cast<const T*>(data + y * step)[x]; }

private:
    int type_ = CV_8UC1;
    std::shared_ptr<std::vector<uint8_t>> buf_;
};

}  // namespace cv
