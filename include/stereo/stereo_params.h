#pragma once
// Drop-in for include/stereo/stereo_params.h of the reference: StereoParams (:20-104) with the
// same enum, fields, defaults and helpers; CameraParams (:135-186) with row-major double
// matrices in std::vector (empty = not set) in place of cv::Mat; and the parameter-file and
// Middlebury loaders (:187-211) over lumen_config.h. The camera-calibration YAML files
// (loadCameraParams/saveCameraParams) are cv::FileStorage documents and are not provided.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../lumen_config.h"
#include "../lumen_sgbm.h"

namespace stereo {

enum class MatchingCostMode : int {
    SAD = 0,     // mean absolute difference of Sobel-x prefiltered images over the block
    Census = 1,  // 5x5 census (24 bits) + Hamming distance
};

struct StereoParams {
    int minDisparity = 0;
    int maxDisparity = 128;
    int blockSize = 5;
    int P1 = 200;
    int P2 = 800;
    int preFilterCap = 31;
    int uniquenessRatio = 10;
    int disp12MaxDiff = 1;
    int speckleWindowSize = 100;
    int speckleRange = 32;
    bool useGPU = true;           // accepted and ignored, as in the reference
    bool useSharedMemory = true;  // does not change results
    int numDirections = 4;        // 4, or 8 when >= 8
    MatchingCostMode matchingCostMode = MatchingCostMode::SAD;

    void autoComputePenalties(int numChannels = 1) {
        P1 = 8 * numChannels * blockSize * blockSize;
        P2 = 32 * numChannels * blockSize * blockSize;
    }

    // Not called by StereoSGBM::compute, as in the reference.
    bool isValid() const {
        if (blockSize < 1 || blockSize % 2 == 0) return false;
        if (minDisparity < 0) return false;
        if (maxDisparity <= minDisparity) return false;
        if (P2 <= P1) return false;
        if (uniquenessRatio < 0 || uniquenessRatio > 100) return false;
        return true;
    }

    lumen_stereo_params toC() const {
        lumen_stereo_params c;
        c.minDisparity = minDisparity;
        c.maxDisparity = maxDisparity;
        c.blockSize = blockSize;
        c.P1 = P1;
        c.P2 = P2;
        c.preFilterCap = preFilterCap;
        c.uniquenessRatio = uniquenessRatio;
        c.disp12MaxDiff = disp12MaxDiff;
        c.speckleWindowSize = speckleWindowSize;
        c.speckleRange = speckleRange;
        c.useGPU = useGPU ? 1 : 0;
        c.useSharedMemory = useSharedMemory ? 1 : 0;
        c.numDirections = numDirections;
        c.matchingCostMode = static_cast<int>(matchingCostMode);
        return c;
    }

    void fromC(const lumen_stereo_params& c) {
        minDisparity = c.minDisparity;
        maxDisparity = c.maxDisparity;
        blockSize = c.blockSize;
        P1 = c.P1;
        P2 = c.P2;
        preFilterCap = c.preFilterCap;
        uniquenessRatio = c.uniquenessRatio;
        disp12MaxDiff = c.disp12MaxDiff;
        speckleWindowSize = c.speckleWindowSize;
        speckleRange = c.speckleRange;
        useGPU = c.useGPU != 0;
        useSharedMemory = c.useSharedMemory != 0;
        numDirections = c.numDirections;
        matchingCostMode = static_cast<MatchingCostMode>(c.matchingCostMode);
    }
};

// Calibration of a stereo pair (include/stereo/stereo_params.h:135-186 of the reference).
struct CameraParams {
    int width = 0;
    int height = 0;
    std::vector<double> K0, D0, K1, D1;  // 3x3 intrinsics, distortion coefficients
    std::vector<double> R, T;            // 3x3 rotation, 3x1 translation
    double baseline = 0.0;               // metres
    double focalLength = 0.0;            // pixels
    int disparityOffset = 0;
    int numDisparities = 0;
    std::vector<double> R0, R1, P0, P1, Q;        // rectification outputs
    std::vector<float> map0x, map0y, map1x, map1y;  // width*height each, or empty

    void computeDerivedValues() {
        if (!T.empty()) {
            double s = 0;
            for (double v : T) s += v * v;
            baseline = std::sqrt(s);
        }
        if (K0.size() == 9) focalLength = (K0[0] + K0[4]) / 2.0;
    }
    bool isValid() const {
        return !K0.empty() && !K1.empty() && !R.empty() && !T.empty() && width > 0 && height > 0;
    }
};

// src/config.cpp:17-45: true when the file was read and the parameters are valid.
inline bool loadStereoParams(const std::string& path, StereoParams& params) {
    lumen_stereo_params c = params.toC();
    const int rc = lumen_params_load_yaml(path.c_str(), &c);
    params.fromC(c);
    return rc == LUMEN_OK;
}

// src/config.cpp:47-72.
inline bool saveStereoParams(const std::string& path, const StereoParams& params) {
    const lumen_stereo_params c = params.toC();
    return lumen_params_save_yaml(path.c_str(), &c) == LUMEN_OK;
}

// src/config.cpp:149-236.
inline bool loadMiddleburyCalib(const std::string& path, CameraParams& params) {
    lumen_camera_params c;
    std::memset(&c, 0, sizeof c);
    c.width = params.width;
    c.height = params.height;
    const int rc = lumen_middlebury_calib_load(path.c_str(), &c);
    if (rc != LUMEN_OK && c.K0[0] == 0.0) return false;  // not read at all
    params.width = c.width;
    params.height = c.height;
    params.K0.assign(c.K0, c.K0 + 9);
    params.D0.assign(c.D0, c.D0 + 5);
    params.K1.assign(c.K1, c.K1 + 9);
    params.D1.assign(c.D1, c.D1 + 5);
    params.R.assign(c.R, c.R + 9);
    params.T.assign(c.T, c.T + 3);
    params.baseline = c.baseline;
    params.focalLength = c.focalLength;
    params.disparityOffset = c.disparityOffset;
    params.numDisparities = c.numDisparities;
    return rc == LUMEN_OK;
}

}  // namespace stereo
