#pragma once
// Drop-in for include/stereo/depth_map.h of the reference (:14-58) over the C ABI of
// liblumen_sgbm.so (lumen_depth.h). Link with -llumen_sgbm.
//
// disparityToDepth keeps the host function's arithmetic (double quotient) and runs it on the
// GPU. The *GPU functions take device buffers (the reference's CudaBuffer arguments) and a
// stream. colorizeDepth writes the colours of the reference's GPU colour map (R, G, B per
// pixel); the reference's host colorizeDepth uses OpenCV's COLORMAP_TURBO (B, G, R), which is
// not reproduced - callers that need those exact colours must keep OpenCV's.
#include <cstdint>

#include "../lumen_depth.h"
#include "common.h"

#ifdef LUMEN_WITH_OPENCV
#include <opencv2/core.hpp>
#endif

namespace stereo {

namespace detail {
inline void depth_check(int rc) {
    if (rc == LUMEN_EINVAL) throw StereoException(lumen_last_error());
    if (rc != LUMEN_OK) throw GpuException(rc, lumen_last_error());
}
}  // namespace detail

// src/depth_map.cu:142-160 on host buffers of width*height values.
inline void disparityToDepth(const int16_t* disparity, float* depth, int width, int height,
                             double focalLength, double baseline, double disparityOffsetPx = 0.0) {
    detail::depth_check(lumen_disparity_to_depth(disparity, depth, width, height, focalLength,
                                                 baseline, disparityOffsetPx));
}

// src/depth_map.cu:177-197 on device buffers; enqueued on `stream` (hipStream_t, or nullptr).
inline void disparityToDepthGPU(const int16_t* d_disparity, float* d_depth, int width, int height,
                                float focalLength, float baseline, float disparityOffsetPx = 0.0f,
                                void* stream = nullptr) {
    detail::depth_check(lumen_disparity_to_depth_device(d_disparity, d_depth, width, height,
                                                        focalLength, baseline, disparityOffsetPx,
                                                        stream));
}

// Host buffers: depth (width*height floats) -> colorized (width*height*3 bytes, R G B).
inline void colorizeDepth(const float* depth, uint8_t* colorized, int width, int height,
                          float minDepth = 0.5f, float maxDepth = 10.0f) {
    detail::depth_check(lumen_colorize_depth(depth, colorized, width, height, minDepth, maxDepth));
}

// src/depth_map.cu:211-221 on device buffers.
inline void colorizeDepthGPU(const float* d_depth, uint8_t* d_colorized, int width, int height,
                             float minDepth, float maxDepth, void* stream = nullptr) {
    detail::depth_check(lumen_colorize_depth_device(d_depth, d_colorized, width, height, minDepth,
                                                    maxDepth, stream));
}

#ifdef LUMEN_WITH_OPENCV
inline void disparityToDepth(const cv::Mat& disparity, cv::Mat& depth, double focalLength,
                             double baseline, double disparityOffsetPx = 0.0) {
    const cv::Mat d = disparity.isContinuous() ? disparity : disparity.clone();
    depth.create(d.size(), CV_32FC1);
    disparityToDepth(reinterpret_cast<const int16_t*>(d.data), reinterpret_cast<float*>(depth.data),
                     d.cols, d.rows, focalLength, baseline, disparityOffsetPx);
}

inline void colorizeDepth(const cv::Mat& depth, cv::Mat& colorized, float minDepth = 0.5f,
                          float maxDepth = 10.0f) {
    const cv::Mat z = depth.isContinuous() ? depth : depth.clone();
    colorized.create(z.size(), CV_8UC3);
    colorizeDepth(reinterpret_cast<const float*>(z.data), colorized.data, z.cols, z.rows, minDepth,
                  maxDepth);
}
#endif

}  // namespace stereo
