#pragma once
// Drop-in for include/stereo/sgbm_gpu.h of the reference (:29-91): class StereoSGBM with the
// same constructor, compute, setParams/getParams and timing getters, implemented over the C ABI
// of liblumen_sgbm.so (lumen_sgbm.h). Link with -llumen_sgbm.
//
// compute() takes host images. Besides the raw-pointer form there is an ImageView form that
// performs the reference's empty/size-mismatch checks, and, when LUMEN_WITH_OPENCV is defined
// before this header is included, the reference's cv::Mat form (CV_8UC1 or CV_8UC3 in,
// CV_16SC1 out).
#include <cstddef>
#include <cstdint>
#include <string>

#include "../lumen_sgbm.h"
#include "common.h"
#include "sgbm_limits.h"
#include "stereo_params.h"

#ifdef LUMEN_WITH_OPENCV
#include <opencv2/core.hpp>
#endif

namespace stereo {

// A host image: rows of `width * channels` bytes, `stride` bytes apart (0 = packed).
struct ImageView {
    const uint8_t* data = nullptr;
    int width = 0;
    int height = 0;
    int channels = 1;
    size_t stride = 0;
    bool empty() const { return data == nullptr || width <= 0 || height <= 0; }
};

class StereoSGBM {
public:
    explicit StereoSGBM(const StereoParams& params = StereoParams()) : params_(params) {
        const lumen_stereo_params c = params_.toC();
        const int rc = lumen_sgbm_create(&c, &h_);
        if (rc != LUMEN_OK) throw GpuException(rc, lumen_last_error());
    }
    ~StereoSGBM() { lumen_sgbm_destroy(h_); }
    StereoSGBM(const StereoSGBM&) = delete;
    StereoSGBM& operator=(const StereoSGBM&) = delete;

    // Grayscale (channels 1) or BGR (channels 3) pair of equal size; disparity receives
    // width*height int16 values in 1/16 px, -1 where invalid.
    void compute(const uint8_t* left, const uint8_t* right, int width, int height,
                 int16_t* disparity, int channels = 1) {
        check(lumen_sgbm_compute(h_, left, 0, right, 0, width, height, channels, disparity, 0));
        readTimings();
    }

    void compute(const ImageView& left, const ImageView& right, int16_t* disparity,
                 size_t disparityStride = 0) {
        if (left.empty() || right.empty()) throw StereoException("Input images are empty");
        if (left.width != right.width || left.height != right.height)
            throw StereoException("Left and right images must have same size");
        check(lumen_sgbm_compute(h_, left.data, left.stride, right.data, right.stride, left.width,
                                 left.height, left.channels, disparity, disparityStride));
        readTimings();
    }

    // Grayscale images and the output in device memory (for example stereo::DeviceBuffer),
    // width*height bytes / int16 values each, packed. Returns when the output is written.
    // No reference counterpart: the reference's compute always takes host images.
    void computeDevice(const uint8_t* d_left, const uint8_t* d_right, int width, int height,
                       int16_t* d_disparity) {
        check(lumen_sgbm_compute_device(h_, d_left, d_right, width, height, d_disparity, 1));
        readTimings();
    }

#ifdef LUMEN_WITH_OPENCV
    void compute(const cv::Mat& left, const cv::Mat& right, cv::Mat& disparity) {
        if (left.empty() || right.empty()) throw StereoException("Input images are empty");
        if (left.size() != right.size())
            throw StereoException("Left and right images must have same size");
        disparity.create(left.rows, left.cols, CV_16SC1);
        check(lumen_sgbm_compute(h_, left.data, left.step, right.data, right.step, left.cols,
                                 left.rows, left.channels(),
                                 reinterpret_cast<int16_t*>(disparity.data), disparity.step));
        readTimings();
    }
#endif

    void setParams(const StereoParams& params) {
        const lumen_stereo_params c = params.toC();
        check(lumen_sgbm_set_params(h_, &c));
        params_ = params;
    }
    const StereoParams& getParams() const { return params_; }

    float getLastComputeTimeMs() const { return t_[0]; }
    float getCostComputeTimeMs() const { return t_[1]; }
    float getAggregationTimeMs() const { return t_[2]; }
    float getPostProcessTimeMs() const { return t_[3]; }

    lumen_sgbm* handle() const { return h_; }

private:
    void check(int rc) {
        if (rc == LUMEN_OK) return;
        const std::string msg = lumen_sgbm_last_error(h_);
        if (rc == LUMEN_EINVAL) throw StereoException(msg);
        throw GpuException(rc, msg);
    }
    void readTimings() { lumen_sgbm_get_timings(h_, t_); }

    StereoParams params_;
    lumen_sgbm* h_ = nullptr;
    float t_[6] = {0, 0, 0, 0, 0, 0};
};

}  // namespace stereo
