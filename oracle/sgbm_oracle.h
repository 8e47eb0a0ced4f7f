/*
 * sgbm_oracle.h - CPU restatement of the reference SGBM matcher (test infrastructure).
 *
 * THIS IS THE ORACLE, NOT THE PRODUCT. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker (or as the CPU baseline
 * that is reported beside the GPU figure). The GPU product never calls into it.
 *
 * Each function restates one reference kernel (src/sgbm_kernels.cu) or host step
 * (src/sgbm_gpu.cu) of Pfactorial01/LumenStereo; the file:line it follows is cited on each
 * definition in sgbm_oracle.cpp. Arrays are row-major; cost volumes and path costs use the
 * reference's layout idx(x,y,d) = (y*W + x)*D + d with D = maxDisparity - minDisparity.
 *
 * Pinning (what checks each function independently of this file):
 *   prefilter        closed-form ramp/clipping known answers + numpy fixtures
 *   census           constant-image and hand-made-patch known answers + numpy fixtures
 *   cost_census      Hamming-to-uint8 table known answers (0,1,12,24) + numpy fixtures
 *   cost_sad         integral-image box sums (a different algorithm) + invariants + fixtures
 *   aggregate_path   brute-force minimum path energy over all labelings (no-wrap regime),
 *                    hand-computed wrap cases, D=1 and constant-cost closed forms + fixtures
 *   wta_left/right   brute force (first-index argmin), hand uniqueness/parabola cases + fixtures
 *   lr_check,median  brute force on small inputs + hand cases + fixtures
 *   speckle          parity unpinned against the reference (the reference kernel races; this
 *                    is a deterministic snapshot restatement), checked by hand cases only
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Field order and meaning follow stereo::StereoParams (include/stereo/stereo_params.h:57-82). */
typedef struct oracle_params {
    int32_t minDisparity;
    int32_t maxDisparity;
    int32_t blockSize;
    int32_t P1;
    int32_t P2;
    int32_t preFilterCap;
    int32_t uniquenessRatio;
    int32_t disp12MaxDiff;
    int32_t speckleWindowSize;
    int32_t speckleRange;
    int32_t useGPU;
    int32_t useSharedMemory;
    int32_t numDirections;
    int32_t matchingCostMode; /* 0 = SAD, 1 = Census */
} oracle_params;

/* Optional stage outputs of oracle_sgbm; any pointer may be NULL.
 * cost: W*H*D u8; paths: npaths*W*H*D u16 (path-major); sum: W*H*D u16;
 * wta_int: best integer d (int16) per pixel; disp_left: left WTA (1/16 px) before LR;
 * disp_right: right WTA; after_lr: after LR check; after_speckle: after speckle. */
typedef struct oracle_dumps {
    uint8_t* cost;
    uint16_t* paths;
    uint16_t* sum;
    int16_t* wta_int;
    int16_t* disp_left;
    int16_t* disp_right;
    int16_t* after_lr;
    int16_t* after_speckle;
} oracle_dumps;

void oracle_prefilter(const uint8_t* img, int W, int H, int preFilterCap, uint8_t* out);
void oracle_census(const uint8_t* img, int W, int H, uint32_t* out);
void oracle_cost_sad(const uint8_t* fl, const uint8_t* fr, int W, int H, int D, int m,
                     int blockSize, uint8_t* cost);
void oracle_cost_census(const uint32_t* cl, const uint32_t* cr, int W, int H, int D, int m,
                        uint8_t* cost);
int oracle_cost_scale(int matchingCostMode, int blockSize, int preFilterCap);
int oracle_num_paths(int numDirections);
/* path: 0 LR, 1 RL, 2 TB, 3 BT, 4 TLBR, 5 BRTL, 6 TRBL, 7 BLTR (the reference's launch order) */
void oracle_aggregate_path(const uint8_t* cost, const uint8_t* img, int W, int H, int D,
                           int P1, int P2, int costScale, int path, uint16_t* L);
void oracle_wta_left(const uint16_t* S, int W, int H, int D, int m, int uniquenessRatio,
                     int16_t* disp, int16_t* wta_int);
void oracle_wta_right(const uint16_t* S, int W, int H, int D, int m, int16_t* dispR);
void oracle_lr_check(int16_t* dispL, const int16_t* dispR, int W, int H, int maxDiff);
void oracle_speckle(int16_t* disp, int W, int H, int minSimilar, int maxDiff);
void oracle_median(const int16_t* in, int16_t* out, int W, int H);

/* Whole matcher as StereoSGBM::compute runs it (src/sgbm_gpu.cu:147-195) on grayscale input.
 * Returns 0, or 1 for an argument the reference rejects (empty, D<1, D>512). */
int oracle_sgbm(const oracle_params* p, const uint8_t* left, const uint8_t* right, int W, int H,
                int16_t* out, const oracle_dumps* dumps);

/* Same computation split over nthreads host threads: the cost, path sum and WTA in bands of
 * rows, the paths one per thread. Every stage computes each element exactly as oracle_sgbm
 * does, so outputs and dumps are identical to it. Used as the checker at full image sizes and
 * as the reported CPU baseline. */
int oracle_sgbm_mt(const oracle_params* p, const uint8_t* left, const uint8_t* right, int W,
                   int H, int16_t* out, int nthreads, const oracle_dumps* dumps);

#ifdef __cplusplus
}
#endif
