// sgbm_host.h - host helpers shared by the per-pair, batched and serving executors.
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <string>

#include "lumen_sgbm.h"
#include "sgbm_internal.h"

namespace lumen {

// 4 paths, or 8 when numDirections >= 8 (src/sgbm_gpu.cu:321-421).
inline int num_paths(const lumen_stereo_params& p) { return p.numDirections >= 8 ? 8 : 4; }

// src/sgbm_gpu.cu:312-317: SAD uses blockSize^2, every other mode the census formula.
inline int cost_scale(const lumen_stereo_params& p) {
    if (p.matchingCostMode == 0) return p.blockSize * p.blockSize;
    return std::max(1, (p.blockSize * p.blockSize * 2 * p.preFilterCap) / 255);
}

// Schedules (sgbm_tune.hip): whether a schedule is valid for slice K (the reason in `why` when
// it is not), and the tuned schedule of a pair (the default when the table has no entry).
bool valid_schedule(const Schedule& s, int K, std::string& why);
Schedule tuned_schedule(int W, int Dp, int mode, int npaths);

// Sets the text lumen_last_error() returns on this thread (sgbm_capi.hip).
void set_last_error(const std::string& msg);

// Scalar fields of a PairJob from the parameters and image size (pointers untouched), with
// the tuned schedule.
inline void fill_job(PairJob& j, const lumen_stereo_params& p, int W, int H) {
    j.W = W;
    j.H = H;
    j.m = p.minDisparity;
    j.D = p.maxDisparity - p.minDisparity;
    j.Dp = padded_range(j.D);
    j.K = j.Dp / kScanLanes;
    j.pathStride = (size_t)W * H * j.Dp;
    j.mode = p.matchingCostMode;
    j.blockSize = p.blockSize;
    j.preFilterCap = p.preFilterCap;
    j.P1 = p.P1;
    j.P2 = p.P2;
    j.costScale = cost_scale(p);
    j.npaths = num_paths(p);
    j.uniquenessRatio = p.uniquenessRatio;
    j.doLR = p.disp12MaxDiff >= 0 ? 1 : 0;
    j.lrMaxDiff = p.disp12MaxDiff * 16;
    j.doSpeckle = p.speckleWindowSize > 0 ? 1 : 0;
    j.speckleMin = std::min(p.speckleWindowSize, 25);
    j.speckleDiff = p.speckleRange * 16;
    j.sched = tuned_schedule(W, j.Dp, j.mode, j.npaths);
}

// Device bytes of one pair's working set: images, padded cost volume, per-path uint16
// buffers and the six int16 maps.
inline size_t pair_footprint(const lumen_stereo_params& p, int W, int H) {
    const size_t n = (size_t)W * H;
    const size_t vol = n * padded_range(p.maxDisparity - p.minDisparity);
    return vol + vol * 2 * num_paths(p) + n * 2 + n * 2 * 6;
}

inline size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

// Bytes of one pair's working set laid out as layout_pair() does: the pieces of
// pair_footprint(), each rounded up to 256 bytes.
inline size_t pair_region_bytes(const lumen_stereo_params& p, int W, int H) {
    const size_t n = (size_t)W * H;
    const size_t vol = n * padded_range(p.maxDisparity - p.minDisparity);
    return 2 * align_up(n, 256) + align_up(vol, 256) + align_up(vol * 2 * num_paths(p), 256) +
           6 * align_up(n * 2, 256);
}

// Points a filled PairJob's buffers into a region of pair_region_bytes() at `base` (256-byte
// aligned): left, right, cost, paths, dispL, dispR, dispLR, wtaInt, spk, then the pair's own
// output map, whose address is returned (the job's `out` is set to it too).
inline int16_t* layout_pair(PairJob& j, char* base) {
    const size_t n = (size_t)j.W * j.H;
    size_t used = 0;
    auto take = [&](size_t bytes) {
        char* q = base + used;
        used += align_up(bytes, 256);
        return q;
    };
    j.left = reinterpret_cast<uint8_t*>(take(n));
    j.right = reinterpret_cast<uint8_t*>(take(n));
    j.cost = reinterpret_cast<uint8_t*>(take(j.pathStride));
    j.L = reinterpret_cast<uint16_t*>(take(j.pathStride * 2 * j.npaths));
    j.dispL = reinterpret_cast<int16_t*>(take(n * 2));
    j.dispR = reinterpret_cast<int16_t*>(take(n * 2));
    j.dispLR = reinterpret_cast<int16_t*>(take(n * 2));
    j.wtaInt = reinterpret_cast<int16_t*>(take(n * 2));
    j.spk = reinterpret_cast<int16_t*>(take(n * 2));
    j.out = reinterpret_cast<int16_t*>(take(n * 2));
    return j.out;
}

// Validation of src/sgbm_gpu.cu:148-166 plus the limits of this implementation; returns ""
// when the pair is accepted, otherwise the message the reference (or this implementation) gives.
inline std::string validate_pair(const lumen_stereo_params& p, int W, int H, bool haveImages) {
    if (!haveImages || W <= 0 || H <= 0) return "Input images are empty";
    // 64-bit, so that extreme parameter values cannot overflow the difference.
    const long long D = static_cast<long long>(p.maxDisparity) - p.minDisparity;
    if (D < 1) return "Invalid disparity range: maxDisparity must be greater than minDisparity";
    if (D > LUMEN_MAX_DISPARITY_RANGE)
        return "Disparity range (maxDisparity - minDisparity) exceeds GPU limit (" +
               std::to_string(LUMEN_MAX_DISPARITY_RANGE) + "). Reduce the search range.";
    if (p.minDisparity < 0) return "minDisparity must be >= 0";
    if (p.matchingCostMode != 1 && W < 2) return "Image width must be at least 2 for the SAD prefilter";
    if (W > kMaxWidth) return "Image width exceeds " + std::to_string(kMaxWidth);
    return "";
}

}  // namespace lumen
