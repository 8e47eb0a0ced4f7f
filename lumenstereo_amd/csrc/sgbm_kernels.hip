// sgbm_kernels.hip - per-pair kernels of the matcher (one launch per stage for one pair):
// cost, sum/WTA/LR, post-processing. The aggregation kernels are in sgbm_paths.hip.
//
// The stage bodies live in sgbm_device.h and are shared with the persistent batched kernels
// (sgbm_batch.hip); these kernels map one grid onto one pair's units of work.
#include "sgbm_device.h"

namespace lumen {

// ------------------------------------------------------------------------------- cost
template <int T>
__global__ __launch_bounds__(256) void cost_census_kernel(PairJob j) {
    extern __shared__ uint32_t shc[];
    census_cost_segment<T>(j, blockIdx.y, blockIdx.x * T, shc);
}

template <int T>
static hipError_t launch_census(const PairJob& j, hipStream_t s, StageEvents ev) {
    dim3 grid((j.W + T - 1) / T, j.H);
    hipExtLaunchKernelGGL(cost_census_kernel<T>, grid, dim3(256),
                          static_cast<uint32_t>(census_lds_bytes(j.Dp, T)), s, ev.start, ev.stop,
                          0, j);
    return hipGetLastError();
}

// R: the window radius for block sizes 3, 5 and 7, 0 for any other. At least six waves per
// SIMD (at most 80 VGPRs; four for R = 3, which needs more than 96 to run without spills)
// keep a whole 640x480 frame's segments resident at once.
// Block (row, segment): consecutive blocks are consecutive rows, so the segments at the image
// borders, which take longer, are dealt round-robin over the XCDs like the others.
template <int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(R == 3 ? 4 : 6))) void cost_sad_kernel(PairJob j) {
    extern __shared__ uint8_t shs[];
    const int x0 = blockIdx.y * sad_segment_px(R > 0 ? R : j.blockSize / 2, j.Dp);
    if constexpr (R > 0)
        sad_cost_runs<R>(j, blockIdx.x, x0, reinterpret_cast<uint32_t*>(shs));
    else
        sad_cost_segment_generic(j, blockIdx.x, x0, shs);
}

// Windows too large for the LDS segment read the images from global memory, one (pixel, d)
// per thread.
__global__ __launch_bounds__(256) void cost_sad_direct_kernel(PairJob j) {
    const size_t item = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)j.W * j.H * j.Dp;
    if (item >= total) return;
    const int d = static_cast<int>(item % j.Dp);
    const size_t pix = item / j.Dp;
    gmem(j.cost)[item] = sad_cost_direct(j, static_cast<int>(pix % j.W), static_cast<int>(pix / j.W), d);
}

// Every cost 255 (an empty SAD window). A kernel rather than a memset node, so the stage can
// be captured into a graph.
__global__ __launch_bounds__(256) void cost_fill_kernel(PairJob j) {
    const size_t n16 = (size_t)j.W * j.H * j.Dp / 16;  // Dp is a multiple of 64
    LUMEN_G u32x4* dst = reinterpret_cast<LUMEN_G u32x4*>(gmem(j.cost));
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n16; i += (size_t)gridDim.x * 256)
        dst[i] = u32x4{~0u, ~0u, ~0u, ~0u};
}

hipError_t launch_cost(const PairJob& j, hipStream_t s, StageEvents ev) {
    if (j.mode == 1) {
        switch (j.sched.censusTile) {
            case 128: return launch_census<128>(j, s, ev);
            case 256: return launch_census<256>(j, s, ev);
            case 512: return launch_census<512>(j, s, ev);
            default: return hipErrorInvalidValue;
        }
    }
    const int r = j.blockSize / 2;
    if (r < 0) {
        // A negative radius gives an empty window: every cost is 255.
        hipExtLaunchKernelGGL(cost_fill_kernel, dim3(1024), dim3(256), 0, s, ev.start, ev.stop, 0,
                              j);
        return hipGetLastError();
    }
    if (sad_lds_bytes(r, j.Dp) <= kSadMaxLds) {
        const int px = sad_segment_px(r, j.Dp);
        dim3 grid(j.H, (j.W + px - 1) / px);
        const uint32_t lds = static_cast<uint32_t>(sad_lds_bytes(r, j.Dp));
        switch (sad_runs_radius(r) ? r : 0) {
            case 1: hipExtLaunchKernelGGL(cost_sad_kernel<1>, grid, dim3(256), lds, s, ev.start, ev.stop, 0, j); break;
            case 2: hipExtLaunchKernelGGL(cost_sad_kernel<2>, grid, dim3(256), lds, s, ev.start, ev.stop, 0, j); break;
            case 3: hipExtLaunchKernelGGL(cost_sad_kernel<3>, grid, dim3(256), lds, s, ev.start, ev.stop, 0, j); break;
            default: hipExtLaunchKernelGGL(cost_sad_kernel<0>, grid, dim3(256), lds, s, ev.start, ev.stop, 0, j); break;
        }
    } else {
        const size_t total = (size_t)j.W * j.H * j.Dp;
        hipExtLaunchKernelGGL(cost_sad_direct_kernel, dim3((unsigned)((total + 255) / 256)),
                              dim3(256), 0, s, ev.start, ev.stop, 0, j);
    }
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- WTA
// One block per row; NT threads (NT/16 pixels in flight per step).
template <int K, bool PAD, int NP, int NT>
__global__ __launch_bounds__(NT) void wta_row_kernel(PairJob j) {
    extern __shared__ unsigned long long rbest[];
    wta_row<K, PAD, NP>(j, blockIdx.x, rbest);
}

// The parity hook's form: the same rows, also storing each pixel's path sum to sumOut.
template <int K, bool PAD, int NP>
__global__ __launch_bounds__(256) void wta_row_dump_kernel(PairJob j, uint16_t* sumOut) {
    extern __shared__ unsigned long long rbest[];
    wta_row<K, PAD, NP, true>(j, blockIdx.x, rbest, sumOut);
}

// Frames with fewer rows than kWtaWideRows (about two per CU) give each row 512 threads, so a
// row takes half the sequential steps; the arithmetic and the atomics' outcome do not depend
// on the block size.
constexpr int kWtaWideRows = 512;

template <int K, int NP>
static hipError_t launch_wta_kn(const PairJob& j, hipStream_t s, uint16_t* sumOut,
                                StageEvents ev) {
    const uint32_t lds = static_cast<uint32_t>(wta_lds_bytes<K>(j.W));
    const bool pad = j.D != j.Dp;
    if (sumOut) {
        if (pad)
            hipExtLaunchKernelGGL((wta_row_dump_kernel<K, true, NP>), dim3(j.H), dim3(256), lds, s,
                                  ev.start, ev.stop, 0, j, sumOut);
        else
            hipExtLaunchKernelGGL((wta_row_dump_kernel<K, false, NP>), dim3(j.H), dim3(256), lds, s,
                                  ev.start, ev.stop, 0, j, sumOut);
        return hipGetLastError();
    }
    if constexpr (K <= 8) {
        if (j.H < kWtaWideRows) {
            if (pad)
                hipExtLaunchKernelGGL((wta_row_kernel<K, true, NP, 512>), dim3(j.H), dim3(512), lds,
                                      s, ev.start, ev.stop, 0, j);
            else
                hipExtLaunchKernelGGL((wta_row_kernel<K, false, NP, 512>), dim3(j.H), dim3(512), lds,
                                      s, ev.start, ev.stop, 0, j);
            return hipGetLastError();
        }
    }
    if (pad) {
        hipExtLaunchKernelGGL((wta_row_kernel<K, true, NP, 256>), dim3(j.H), dim3(256), lds, s,
                              ev.start, ev.stop, 0, j);
    } else {
        hipExtLaunchKernelGGL((wta_row_kernel<K, false, NP, 256>), dim3(j.H), dim3(256), lds, s,
                              ev.start, ev.stop, 0, j);
    }
    return hipGetLastError();
}

template <int K>
static hipError_t launch_wta_k(const PairJob& j, hipStream_t s, uint16_t* sumOut,
                               StageEvents ev) {
    return j.npaths == 8 ? launch_wta_kn<K, 8>(j, s, sumOut, ev)
                         : launch_wta_kn<K, 4>(j, s, sumOut, ev);
}

hipError_t launch_wta(const PairJob& j, hipStream_t s, uint16_t* sumOut, StageEvents ev) {
    if (j.W > kMaxWidth) return hipErrorInvalidValue;
    switch (j.K) {
        case 4: return launch_wta_k<4>(j, s, sumOut, ev);
        case 6: return launch_wta_k<6>(j, s, sumOut, ev);
        case 8: return launch_wta_k<8>(j, s, sumOut, ev);
        case 12: return launch_wta_k<12>(j, s, sumOut, ev);
        case 16: return launch_wta_k<16>(j, s, sumOut, ev);
        case 20: return launch_wta_k<20>(j, s, sumOut, ev);
        case 24: return launch_wta_k<24>(j, s, sumOut, ev);
        case 28: return launch_wta_k<28>(j, s, sumOut, ev);
        case 32: return launch_wta_k<32>(j, s, sumOut, ev);
        default: return hipErrorInvalidValue;
    }
}

// ------------------------------------------------------------------------------- post
constexpr int kSpkTW = 64, kSpkTH = 16;

__global__ __launch_bounds__(256) void speckle_kernel(PairJob j) {
    __shared__ int16_t tile[(kSpkTH + 4) * (kSpkTW + 4)];
    speckle_tile<kSpkTW, kSpkTH>(gmem(j.dispLR), gmem(j.spk), j.W, j.H, blockIdx.x * kSpkTW,
                                 blockIdx.y * kSpkTH, j.speckleMin, j.speckleDiff, tile);
}

__global__ __launch_bounds__(256) void median3x3_kernel(PairJob j, int16_t* mirror) {
    __shared__ int16_t tile[kMedTH + 2][kMedTW + 2];
    median_tile(gmem(j.doSpeckle ? j.spk : j.dispLR), gmem(j.out), j.W, j.H, blockIdx.x * kMedTW,
                blockIdx.y * kMedTH, tile, mirror ? gmem(mirror) : nullptr);
}

hipError_t launch_post(const PairJob& j, hipStream_t s, int16_t* mirror, StageEvents ev) {
    hipEvent_t first = ev.start;
    if (j.doSpeckle) {
        dim3 grid((j.W + kSpkTW - 1) / kSpkTW, (j.H + kSpkTH - 1) / kSpkTH);
        hipExtLaunchKernelGGL(speckle_kernel, grid, dim3(256), 0, s, first, nullptr, 0, j);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        first = nullptr;
    }
    dim3 grid((j.W + kMedTW - 1) / kMedTW, (j.H + kMedTH - 1) / kMedTH);
    hipExtLaunchKernelGGL(median3x3_kernel, grid, dim3(256), 0, s, first, ev.stop, 0, j, mirror);
    return hipGetLastError();
}

// [n][Dp] -> [n][D] copy of a padded volume of ELEM-byte entries (parity dumps only): the
// dump then leaves the device as one contiguous copy.
template <typename T>
__global__ __launch_bounds__(256) void strip_pad_kernel(const T* __restrict__ src, T* __restrict__ dst,
                                                        size_t total, int D, int Dp) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= total) return;
    const size_t p = i / D;
    dst[i] = src[p * Dp + (i - p * D)];
}

hipError_t launch_strip(const void* src, void* dst, size_t n, int D, int Dp, int elem,
                        hipStream_t s) {
    const size_t total = n * D;
    const dim3 grid((unsigned)((total + 255) / 256));
    if (elem == 1)
        hipLaunchKernelGGL(strip_pad_kernel<uint8_t>, grid, dim3(256), 0, s,
                           static_cast<const uint8_t*>(src), static_cast<uint8_t*>(dst), total, D, Dp);
    else
        hipLaunchKernelGGL(strip_pad_kernel<uint16_t>, grid, dim3(256), 0, s,
                           static_cast<const uint16_t*>(src), static_cast<uint16_t*>(dst), total, D, Dp);
    return hipGetLastError();
}

}  // namespace lumen
