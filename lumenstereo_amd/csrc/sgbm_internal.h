// sgbm_internal.h - shared declarations between the HIP kernels and the host code.
#pragma once
#include <hip/hip_ext.h>
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
// Elements of the aggregation kernels' trash buffer (64 lanes x 32 entries).
constexpr int kTrashElems = 64 * 32;

// Disparity-innermost layouts are padded to a multiple of 64 entries (16 lanes x 4), except
// that ranges of 65 to 96 take 96 (16 lanes x 6): padding them to 128 would move a third more
// bytes than the range needs. Every padded range Dp thus leaves less padding (Dp - D) than one
// chunk of its slice layout spans (at least 64 entries), so the padding always lies in the
// last chunk.
inline int padded_range(int D) { return D > 64 && D <= 96 ? 96 : (D + 63) / 64 * 64; }

// Vector stores per aggregation step and lane (K: entries per lane). 16-lane scanlines: K/8
// 16-byte stores, K/4 8-byte ones, or (K = 6) one 12-byte one. 32- and 64-lane scanlines: one
// 4-, 8- or 16-byte store (K = 2, 4, 8), two 16-byte ones (K = 16), K/4 8-byte ones (K = 12) or
// K/2 4-byte ones (K = 6, 10, 14).
__host__ __device__ constexpr int path_stores(int K, int LPS = 16) {
    return LPS >= 32 ? (K == 2 || K == 4 || K == 8 ? 1 : K == 16 ? 2 : K % 4 == 0 ? K / 4 : K / 2)
                     : (K % 8 == 0 ? K / 8 : K % 4 == 0 ? K / 4 : 1);
}

// Prefetch depths the aggregation supports for slice K: each lane of a DPP row fetches the
// image bytes of at most one step of a block (PF <= 16), and the stores of one block must fit
// the 6-bit vmcnt the loop waits on (PF * stores < 64).
__host__ __device__ constexpr bool valid_prefetch(int K, int PF, int LPS = 16) {
    return (PF == 4 || PF == 8 || PF == 16) && PF * path_stores(K, LPS) < 64;
}

// Whether 32-lane scanlines can run slice K (disparities per 16 lanes) with prefetch depth PF:
// each lane then holds K/2 disparities, a whole number of packed pairs (not for K = 6), and a
// block's stores must fit the wait count.
__host__ __device__ constexpr bool wide_scanlines_ok(int K, int PF) {
    return K % 4 == 0 && valid_prefetch(K / 2, PF, 32);
}

// The same for 64-lane scanlines (one per wave, K/4 disparities per lane): Dp a multiple of 128.
__host__ __device__ constexpr bool xwide_scanlines_ok(int K, int PF) {
    return K % 8 == 0 && valid_prefetch(K / 4, PF, 64);
}

__host__ __device__ constexpr int default_prefetch(int K) { return K <= 8 ? 8 : 4; }

// Schedule of one pair: how its work is grouped and ordered, never what is computed. Every
// valid schedule gives bit-identical results.
struct Schedule {
    int aggPrefetch;  // aggregation cost prefetch depth: 4, 8 or 16 (valid_prefetch)
    int aggPriority;  // aggregation wave priority: 0 none, 1 by path kind, 2 by scanline length
    int censusTile;   // census cost segment width: 128, 256 or 512 pixels
    int aggWide;      // paths on 32-lane scanlines (two per wave): 0 none, 1 the straight
                      // pair along the longer image dimension, 2 all four straight paths,
                      // 3 every path; 4, 5, 6 the same sets on 64-lane scanlines (one per
                      // wave); the others run 16-lane ones
    int aggOccupancy; // aggregation blocks per CU at most: 0 no cap, 2, 3 or 4
};

inline Schedule default_schedule(int K) { return Schedule{default_prefetch(K), 1, 256, 0, 0}; }

// Everything the device stages need to know about one stereo pair.
struct PairJob {
    const uint8_t* left;   // [H][W] grayscale
    const uint8_t* right;  // [H][W] grayscale
    uint8_t* cost;         // [H][W][Dp]
    uint16_t* L;           // [npaths][H][W][Dp]
    int16_t* dispL;        // left WTA before the LR check
    int16_t* dispR;        // right WTA
    int16_t* dispLR;       // after the LR check (equal to dispL when the check is off)
    int16_t* wtaInt;       // integer best disparity of the left WTA
    int16_t* spk;          // after speckle (when enabled)
    int16_t* out;          // final output
    size_t pathStride;     // H*W*Dp
    int W, H, D, Dp, m, K;
    int mode;              // 1 = census, anything else = SAD (src/sgbm_gpu.cu:206)
    int blockSize, preFilterCap;
    int P1, P2, costScale, npaths;
    int uniquenessRatio;
    int doLR, lrMaxDiff;   // 16 * disp12MaxDiff
    int doSpeckle, speckleMin, speckleDiff;
    Schedule sched;
};

// Events a stage launch attaches to its own kernel dispatches (hipExtLaunchKernel): `start`
// takes the start time of the stage's first kernel, `stop` the end time of its last one, so a
// stage time is the kernels' own duration. Null members are not recorded.
struct StageEvents {
    hipEvent_t start = nullptr;
    hipEvent_t stop = nullptr;
};

// Wave ranges of the per-pair aggregation grid: waves waveStart[r] .. waveStart[r+1]-1 run
// path r's scanlines.
struct PathsGrid {
    int waveStart[kMaxPaths + 1];
    int wideMask;   // bit r: path r runs 32-lane scanlines
    int xwideMask;  // bit r: path r runs 64-lane scanlines
};

// The aggregation launch of slice K (sgbm_paths_impl.h; instantiated in sgbm_paths_k*.hip).
template <int K>
hipError_t launch_paths_k(const PairJob& j, const PathsGrid& g, int blocks, size_t lds,
                          uint16_t* trash, hipStream_t s, StageEvents ev);

// Per-pair launches (sgbm_kernels.hip, sgbm_paths.hip).
hipError_t launch_cost(const PairJob& j, hipStream_t s, StageEvents ev = {});
hipError_t launch_paths(const PairJob& j, uint16_t* trash, hipStream_t s, StageEvents ev = {});
// sumOut (may be null): the parity hook's buffer for the fused path sum ([H][W][Dp]).
hipError_t launch_wta(const PairJob& j, hipStream_t s, uint16_t* sumOut = nullptr,
                      StageEvents ev = {});
// mirror (may be null): a second output buffer the median also writes.
hipError_t launch_post(const PairJob& j, hipStream_t s, int16_t* mirror = nullptr,
                       StageEvents ev = {});
// [n][Dp] -> [n][D] copy of a padded volume of 1- or 2-byte entries (parity dumps).
hipError_t launch_strip(const void* src, void* dst, size_t n, int D, int Dp, int elem,
                        hipStream_t s);
// Number of scanlines of SGM path `path` (0..7) over a W x H image.
int scanline_count(int path, int W, int H);

}  // namespace lumen
