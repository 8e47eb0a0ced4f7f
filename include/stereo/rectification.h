#pragma once
// Drop-in for include/stereo/rectification.h of the reference (:25-111): StereoRectifier and
// rectifyImageGPU over the C ABI of liblumen_sgbm.so (lumen_rectify.h). Link with
// -llumen_sgbm. Images are host buffers of width*height*channels bytes (packed rows); with
// LUMEN_WITH_OPENCV defined before this header, the reference's cv::Mat forms are there too.
// benchmarkRectification (a CPU-vs-GPU printout around cv::remap) is not provided.
#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../lumen_rectify.h"
#include "common.h"
#include "stereo_params.h"

#ifdef LUMEN_WITH_OPENCV
#include <opencv2/core.hpp>
#endif

namespace stereo {

class StereoRectifier {
public:
    // src/rectification.cu:262-284: empty map0x/map0y means pre-rectified.
    explicit StereoRectifier(const CameraParams& params) : params_(params) {
        const bool pre = params.map0x.empty() || params.map0y.empty();
        const size_t n = static_cast<size_t>(params.width) * params.height;
        if (!pre && (params.map0x.size() != n || params.map0y.size() != n ||
                     params.map1x.size() != n || params.map1y.size() != n))
            throw StereoException("Rectification maps must hold width*height entries each");
        const int rc = lumen_rectifier_create(pre ? nullptr : params.map0x.data(),
                                              pre ? nullptr : params.map0y.data(),
                                              pre ? nullptr : params.map1x.data(),
                                              pre ? nullptr : params.map1y.data(), params.width,
                                              params.height, &h_);
        if (rc == LUMEN_EINVAL) throw StereoException(lumen_last_error());
        if (rc != LUMEN_OK) throw GpuException(rc, lumen_last_error());
    }
    ~StereoRectifier() { lumen_rectifier_destroy(h_); }
    StereoRectifier(const StereoRectifier&) = delete;
    StereoRectifier& operator=(const StereoRectifier&) = delete;

    // src/rectification.cu:286-352: both images remapped (or copied, when pre-rectified).
    void rectify(const uint8_t* leftRaw, const uint8_t* rightRaw, int width, int height,
                 int channels, uint8_t* leftRect, uint8_t* rightRect) {
        check(lumen_rectifier_rectify(h_, leftRaw, rightRaw, width, height, channels, leftRect,
                                      rightRect));
    }

#ifdef LUMEN_WITH_OPENCV
    void rectify(const cv::Mat& leftRaw, const cv::Mat& rightRaw, cv::Mat& leftRect,
                 cv::Mat& rightRect) {
        if (isPreRectified()) {
            leftRect = leftRaw.clone();
            rightRect = rightRaw.clone();
            return;
        }
        const cv::Mat l = leftRaw.isContinuous() ? leftRaw : leftRaw.clone();
        const cv::Mat r = rightRaw.isContinuous() ? rightRaw : rightRaw.clone();
        leftRect.create(l.rows, l.cols, l.type());
        rightRect.create(r.rows, r.cols, r.type());
        rectify(l.data, r.data, l.cols, l.rows, l.channels(), leftRect.data, rightRect.data);
    }
#endif

    bool isPreRectified() const { return lumen_rectifier_is_pre_rectified(h_) == 1; }
    const CameraParams& getParams() const { return params_; }
    lumen_rectifier* handle() const { return h_; }

private:
    void check(int rc) {
        if (rc == LUMEN_OK) return;
        const std::string msg = lumen_rectifier_last_error(h_);
        if (rc == LUMEN_EINVAL) throw StereoException(msg);
        throw GpuException(rc, msg);
    }

    CameraParams params_;
    lumen_rectifier* h_ = nullptr;
};

// src/rectification.cu:363-420: one image through its maps (width*height floats each).
inline void rectifyImageGPU(const uint8_t* src, const float* mapX, const float* mapY,
                            uint8_t* dst, int width, int height, int channels = 1) {
    const int rc = lumen_remap(src, mapX, mapY, dst, width, height, channels);
    if (rc == LUMEN_EINVAL) throw StereoException(lumen_last_error());
    if (rc != LUMEN_OK) throw GpuException(rc, lumen_last_error());
}

#ifdef LUMEN_WITH_OPENCV
inline void rectifyImageGPU(const cv::Mat& src, const cv::Mat& mapX, const cv::Mat& mapY,
                            cv::Mat& dst) {
    const cv::Mat s = src.isContinuous() ? src : src.clone();
    dst.create(s.rows, s.cols, s.type());
    rectifyImageGPU(s.data, reinterpret_cast<const float*>(mapX.data),
                    reinterpret_cast<const float*>(mapY.data), dst.data, s.cols, s.rows,
                    s.channels());
}
#endif

}  // namespace stereo
