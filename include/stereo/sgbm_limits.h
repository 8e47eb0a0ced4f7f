#pragma once
// Drop-in for include/stereo/sgbm_limits.h of the reference (:9, :18-22).
#include "../lumen_sgbm.h"

namespace stereo {

// numDisparities = maxDisparity - minDisparity must not exceed this value.
inline constexpr int kSgbmMaxDisparityRange = LUMEN_MAX_DISPARITY_RANGE;

// Middlebury calib ndisp clamped to the matcher's limit (fallback when ndisp is unknown).
inline int middleburyMatcherMaxDisparity(int calibNumDisparities, int fallbackWhenMissing = 96) {
    const int ndisp = (calibNumDisparities > 0) ? calibNumDisparities : fallbackWhenMissing;
    return (ndisp > kSgbmMaxDisparityRange) ? kSgbmMaxDisparityRange : ndisp;
}

}  // namespace stereo
