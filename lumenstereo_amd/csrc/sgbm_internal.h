// sgbm_internal.h - shared declarations between the HIP kernels and the C-ABI host code.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace lumen {

// Lanes that carry one scanline's disparity vector: one DPP row of the 64-wide wavefront.
constexpr int kScanLanes = 16;
constexpr int kMaxPaths = 8;
constexpr int kMaxDisparityRange = 512;
// The fused sum/WTA kernel keeps one row of right-WTA state in LDS: up to 14 bytes per pixel
// (8-byte keys with one padding entry per 4, plus two int16 maps) within 64 KiB.
constexpr int kMaxWidth = 4600;

// Elements of the aggregation kernel's trash buffer (64 lanes x 32 entries).
constexpr int kTrashElems = 64 * 32;

// Disparity-innermost layouts are padded to a multiple of 64 entries (16 lanes x 4).
inline int padded_range(int D) { return (D + 63) / 64 * 64; }

struct PathsArgs {
    const uint8_t* cost;   // [H][W][Dp] uint8
    const uint8_t* img;    // left raw image [H][W]
    uint16_t* L;           // [npaths][H][W][Dp] uint16
    uint16_t* trash;       // kTrashElems uint16: sink for stores past the end of a scanline
    size_t pathStride;     // H*W*Dp
    int W, H, D, Dp;
    int P1, costScale;
    int npaths;
    int waveStart[kMaxPaths + 1];  // first wave of each path in the grid
    uint16_t p2tab[256];           // t16(P2adj(|dI|)) for |dI| = 0..255
};

struct WtaArgs {
    const uint16_t* L;
    size_t pathStride;
    int npaths;
    int W, H, D, Dp, m;
    int uniquenessRatio;
    int lrMaxDiff;  // 16 * disp12MaxDiff
    int doLR;
    int16_t* dispL;   // left WTA before the LR check
    int16_t* dispR;   // right WTA
    int16_t* dispLR;  // after the LR check (equal to dispL when the check is off)
    int16_t* wtaInt;  // integer best disparity of the left WTA
};

// Number of paths and waves the aggregation grid needs.
int scanline_count(int path, int W, int H);

hipError_t launch_cost_census(const uint8_t* left, const uint8_t* right, uint8_t* cost, int W,
                              int H, int D, int Dp, int m, hipStream_t s);
hipError_t launch_cost_sad(const uint8_t* left, const uint8_t* right, uint8_t* cost, int W, int H,
                           int D, int Dp, int m, int blockSize, int preFilterCap, hipStream_t s);
hipError_t launch_paths(PathsArgs& a, hipStream_t s);
hipError_t launch_wta(const WtaArgs& a, hipStream_t s);
hipError_t launch_speckle(const int16_t* in, int16_t* out, int W, int H, int minSimilar,
                          int maxDiff, hipStream_t s);
hipError_t launch_median(const int16_t* in, int16_t* out, int W, int H, hipStream_t s);
hipError_t launch_path_sum(const uint16_t* L, size_t pathStride, int npaths, uint16_t* S,
                           size_t n, hipStream_t s);

}  // namespace lumen
