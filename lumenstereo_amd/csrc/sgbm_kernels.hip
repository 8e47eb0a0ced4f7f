// sgbm_kernels.hip - per-pair kernels of the matcher (one launch per stage for one pair).
//
// The stage bodies live in sgbm_device.h and are shared with the persistent batched kernels
// (sgbm_batch.hip); these kernels map one grid onto one pair's units of work.
#include "sgbm_device.h"

namespace lumen {

int scanline_count(int path, int W, int H) { return scan_count(path, W, H); }

// ------------------------------------------------------------------------------- cost
__global__ __launch_bounds__(256) void cost_census_kernel(PairJob j) {
    extern __shared__ uint32_t shc[];
    census_cost_segment(j, blockIdx.y, blockIdx.x * kCensusTile, shc);
}

__global__ __launch_bounds__(256) void cost_sad_kernel(PairJob j) {
    extern __shared__ uint8_t shs[];
    sad_cost_segment(j, blockIdx.y, blockIdx.x * kSadTile, shs);
}

// Windows too large for the LDS segment read the images from global memory, one (pixel, d)
// per thread.
__global__ __launch_bounds__(256) void cost_sad_direct_kernel(PairJob j) {
    const size_t item = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)j.W * j.H * j.Dp;
    if (item >= total) return;
    const int d = static_cast<int>(item % j.Dp);
    const size_t pix = item / j.Dp;
    j.cost[item] = sad_cost_direct(j, static_cast<int>(pix % j.W), static_cast<int>(pix / j.W), d);
}

hipError_t launch_cost(const PairJob& j, hipStream_t s) {
    if (j.mode == 1) {
        dim3 grid((j.W + kCensusTile - 1) / kCensusTile, j.H);
        hipLaunchKernelGGL(cost_census_kernel, grid, dim3(256), census_lds_bytes(j.Dp), s, j);
        return hipGetLastError();
    }
    const int r = j.blockSize / 2;
    if (r < 0) {
        // A negative radius gives an empty window: every cost is 255.
        return hipMemsetAsync(j.cost, 0xFF, (size_t)j.W * j.H * j.Dp, s);
    }
    if (sad_lds_bytes(r, j.Dp) <= kSadMaxLds) {
        dim3 grid((j.W + kSadTile - 1) / kSadTile, j.H);
        hipLaunchKernelGGL(cost_sad_kernel, grid, dim3(256), sad_lds_bytes(r, j.Dp), s, j);
    } else {
        const size_t total = (size_t)j.W * j.H * j.Dp;
        hipLaunchKernelGGL(cost_sad_direct_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256),
                           0, s, j);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- paths
// All paths of the pair in one launch: wave w takes scanlines 4w'..4w'+3 of the path whose
// range of waves contains w (paths in order, so the long horizontal scanlines start first).
struct PathsGrid {
    int waveStart[kMaxPaths + 1];
};

template <int K, bool PAD>
__global__ __launch_bounds__(256) void sgm_paths_kernel(PairJob j, PathsGrid g, uint16_t* trash) {
    const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (wave >= g.waveStart[j.npaths]) return;
    int path = 0;
#pragma unroll
    for (int r = 1; r < kMaxPaths; ++r)
        if (r < j.npaths && wave >= g.waveStart[r]) path = r;
    const int s0 = (wave - g.waveStart[path]) * 4;
    // The launch ends when its longest scanlines end: the longer straight paths get the higher
    // issue priority within their SIMD, diagonals the lowest.
    const bool longest = j.W >= j.H ? path < 2 : (path >= 2 && path < 4);
    if (longest)
        __builtin_amdgcn_s_setprio(3);
    else if (path < 4)
        __builtin_amdgcn_s_setprio(2);
    else
        __builtin_amdgcn_s_setprio(1);
    paths_wave<K, PAD>(j, path, s0, trash);
}

template <int K>
static hipError_t launch_paths_k(const PairJob& j, const PathsGrid& g, int blocks,
                                 uint16_t* trash, hipStream_t s) {
    if (j.D == j.Dp)
        hipLaunchKernelGGL((sgm_paths_kernel<K, false>), dim3(blocks), dim3(256), 0, s, j, g, trash);
    else
        hipLaunchKernelGGL((sgm_paths_kernel<K, true>), dim3(blocks), dim3(256), 0, s, j, g, trash);
    return hipGetLastError();
}

hipError_t launch_paths(const PairJob& j, uint16_t* trash, hipStream_t s) {
    PathsGrid g;
    int waves = 0;
    for (int r = 0; r < j.npaths; ++r) {
        g.waveStart[r] = waves;
        waves += (scan_count(r, j.W, j.H) + 3) / 4;
    }
    for (int r = j.npaths; r <= kMaxPaths; ++r) g.waveStart[r] = waves;
    const int blocks = (waves + 3) / 4;
    switch (j.K) {
        case 4: return launch_paths_k<4>(j, g, blocks, trash, s);
        case 8: return launch_paths_k<8>(j, g, blocks, trash, s);
        case 12: return launch_paths_k<12>(j, g, blocks, trash, s);
        case 16: return launch_paths_k<16>(j, g, blocks, trash, s);
        case 20: return launch_paths_k<20>(j, g, blocks, trash, s);
        case 24: return launch_paths_k<24>(j, g, blocks, trash, s);
        case 28: return launch_paths_k<28>(j, g, blocks, trash, s);
        case 32: return launch_paths_k<32>(j, g, blocks, trash, s);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------------------- WTA
template <int K, bool PAD>
__global__ __launch_bounds__(256) void wta_row_kernel(PairJob j) {
    extern __shared__ unsigned long long rbest[];
    wta_row<K, PAD>(j, blockIdx.x, rbest);
}

template <int K>
static hipError_t launch_wta_k(const PairJob& j, hipStream_t s) {
    const size_t lds = wta_lds_bytes<K>(j.W);
    if (j.D == j.Dp)
        hipLaunchKernelGGL((wta_row_kernel<K, false>), dim3(j.H), dim3(256), lds, s, j);
    else
        hipLaunchKernelGGL((wta_row_kernel<K, true>), dim3(j.H), dim3(256), lds, s, j);
    return hipGetLastError();
}

hipError_t launch_wta(const PairJob& j, hipStream_t s) {
    if (j.W > kMaxWidth) return hipErrorInvalidValue;
    switch (j.K) {
        case 4: return launch_wta_k<4>(j, s);
        case 8: return launch_wta_k<8>(j, s);
        case 12: return launch_wta_k<12>(j, s);
        case 16: return launch_wta_k<16>(j, s);
        case 20: return launch_wta_k<20>(j, s);
        case 24: return launch_wta_k<24>(j, s);
        case 28: return launch_wta_k<28>(j, s);
        case 32: return launch_wta_k<32>(j, s);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------------------- post
__global__ __launch_bounds__(256) void speckle_kernel(PairJob j) {
    const int x = blockIdx.x * 16 + (threadIdx.x & 15), y = blockIdx.y * 16 + (threadIdx.x >> 4);
    if (x >= j.W || y >= j.H) return;
    speckle_px(j.dispLR, j.spk, j.W, j.H, x, y, j.speckleMin, j.speckleDiff);
}

__global__ __launch_bounds__(256) void median3x3_kernel(PairJob j) {
    __shared__ int16_t tile[kMedTH + 2][kMedTW + 2];
    median_tile(j.doSpeckle ? j.spk : j.dispLR, j.out, j.W, j.H, blockIdx.x * kMedTW,
                blockIdx.y * kMedTH, tile);
}

hipError_t launch_post(const PairJob& j, hipStream_t s) {
    if (j.doSpeckle) {
        dim3 grid((j.W + 15) / 16, (j.H + 15) / 16);
        hipLaunchKernelGGL(speckle_kernel, grid, dim3(256), 0, s, j);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    dim3 grid((j.W + kMedTW - 1) / kMedTW, (j.H + kMedTH - 1) / kMedTH);
    hipLaunchKernelGGL(median3x3_kernel, grid, dim3(256), 0, s, j);
    return hipGetLastError();
}

// Sum of all path costs mod 2^16 into a padded volume (parity dumps only).
__global__ __launch_bounds__(256) void path_sum_kernel(const uint16_t* __restrict__ L,
                                                       size_t pathStride, int npaths,
                                                       uint16_t* __restrict__ S, size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    uint32_t acc = 0;
    for (int r = 0; r < npaths; ++r) acc += L[(size_t)r * pathStride + i];
    S[i] = static_cast<uint16_t>(acc & 0xFFFFu);
}

hipError_t launch_path_sum(const uint16_t* L, size_t pathStride, int npaths, uint16_t* S,
                           size_t n, hipStream_t s) {
    hipLaunchKernelGGL(path_sum_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, L,
                       pathStride, npaths, S, n);
    return hipGetLastError();
}

}  // namespace lumen
