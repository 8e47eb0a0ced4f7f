#pragma once
// Drop-in for the StereoParams part of include/stereo/stereo_params.h of the reference
// (:20-104): same enum, same fields, same defaults, same helpers. The camera calibration types
// and the YAML loaders of that header belong to modules outside the matcher and are not here.
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
};

}  // namespace stereo
