// depth_map.hip - disparity -> depth and the depth colour map (the stages after the matcher).
//
// Both are elementwise over the W x H map, so both are bound by HBM: depth reads 2 bytes and
// writes 4 per pixel, the colour map reads 4 and writes 3, and the fused form reads 2 and
// writes 3 (plus 4 when the depth map is kept). Each thread takes four consecutive pixels: one
// 8-byte disparity load, one 16-byte depth load or store and one 12-byte colour store, so
// every wave moves contiguous runs. The 256-entry colour table is staged in LDS once per block
// as packed R | G << 8 | B << 16 words, since the lanes' lookups are data-dependent.
#include "lumen_depth.h"
#include "scene_common.h"
#include "turbo_table.h"

#include <hip/hip_runtime.h>

#include <cstring>
#include <string>

namespace lumen {
namespace {

constexpr int kBlock = 256;
constexpr int kPixPerThread = 4;

struct PackedTable {
    uint32_t v[256];
};

constexpr PackedTable make_packed_table() {
    PackedTable t{};
    for (int i = 0; i < 256; ++i)
        t.v[i] = (uint32_t)kTurboR[i] | ((uint32_t)kTurboG[i] << 8) | ((uint32_t)kTurboB[i] << 16);
    return t;
}

__constant__ PackedTable kTurboPacked = make_packed_table();

struct DepthArgs {
    float f, B, doffs;      // float form (src/depth_map.cu:74-101)
    double fd, Bd, doffsd;  // host form (src/depth_map.cu:142-160)
};

// src/depth_map.cu:84-100: 0 for d <= 0; d/16 + doffs; (f*B)/denom when denom > 0. HOST selects
// the host function's arithmetic: the quotient in double, rounded to float once.
template <bool HOST>
__device__ __forceinline__ float depth_of(int d, const DepthArgs& a) {
    if (d <= 0) return 0.0f;
    if constexpr (HOST) {
        const float denom = static_cast<float>(d) / 16.0f + static_cast<float>(a.doffsd);
        return denom > 0.0f ? static_cast<float>(a.fd * a.Bd / static_cast<double>(denom)) : 0.0f;
    } else {
        const float denom = static_cast<float>(d) / 16.0f + a.doffs;
        return denom > 0.0f ? (a.f * a.B) / denom : 0.0f;
    }
}

// src/depth_map.cu:120-131. fminf/fmaxf return the non-NaN operand, so NaN lands on entry 0.
__device__ __forceinline__ int colour_index(float d, float minD, float maxD) {
    if (d <= 0.0f || d > maxD * 2.0f) return 0;
    float n = (d - minD) / (maxD - minD);
    n = fmaxf(0.0f, fminf(1.0f, n));
    n = 1.0f - n;
    return static_cast<int>(n * 255.0f);
}

__device__ __forceinline__ void stage_table(uint32_t* tab) {
    tab[threadIdx.x] = kTurboPacked.v[threadIdx.x];
    __syncthreads();
}

// Four pixels' colours as three little-endian words: R0 G0 B0 R1 | G1 B1 R2 G2 | B2 R3 G3 B3.
__device__ __forceinline__ void store_rgb4(uint8_t* dst, const uint32_t (&c)[4]) {
    const uint32_t w0 = c[0] | (c[1] << 24);
    const uint32_t w1 = (c[1] >> 8) | (c[2] << 16);
    const uint32_t w2 = (c[2] >> 16) | (c[3] << 8);
    uint32_t* d = reinterpret_cast<uint32_t*>(dst);
    d[0] = w0;
    d[1] = w1;
    d[2] = w2;
}

__device__ __forceinline__ void store_rgb1(uint8_t* dst, uint32_t c) {
    dst[0] = static_cast<uint8_t>(c);
    dst[1] = static_cast<uint8_t>(c >> 8);
    dst[2] = static_cast<uint8_t>(c >> 16);
}

// VEC: the pointers are aligned for the 4-pixel accesses (checked on the host); otherwise, and
// for the last partial group, pixels are handled one by one.
template <bool HOST, bool VEC>
__global__ void __launch_bounds__(kBlock) depth_kernel(const int16_t* __restrict__ disp,
                                                       float* __restrict__ depth, long long n,
                                                       DepthArgs a) {
    const long long i = ((long long)blockIdx.x * kBlock + threadIdx.x) * kPixPerThread;
    if (i >= n) return;
    if (VEC && i + kPixPerThread <= n) {
        const uint2 raw = *reinterpret_cast<const uint2*>(disp + i);
        const int d0 = static_cast<int16_t>(raw.x & 0xFFFFu), d1 = static_cast<int16_t>(raw.x >> 16);
        const int d2 = static_cast<int16_t>(raw.y & 0xFFFFu), d3 = static_cast<int16_t>(raw.y >> 16);
        *reinterpret_cast<float4*>(depth + i) =
            float4{depth_of<HOST>(d0, a), depth_of<HOST>(d1, a), depth_of<HOST>(d2, a),
                   depth_of<HOST>(d3, a)};
        return;
    }
    for (int k = 0; k < kPixPerThread && i + k < n; ++k) depth[i + k] = depth_of<HOST>(disp[i + k], a);
}

template <bool VEC>
__global__ void __launch_bounds__(kBlock) colorize_kernel(const float* __restrict__ depth,
                                                          uint8_t* __restrict__ rgb, long long n,
                                                          float minD, float maxD) {
    __shared__ uint32_t tab[256];
    stage_table(tab);
    const long long i = ((long long)blockIdx.x * kBlock + threadIdx.x) * kPixPerThread;
    if (i >= n) return;
    if (VEC && i + kPixPerThread <= n) {
        const float4 d = *reinterpret_cast<const float4*>(depth + i);
        const uint32_t c[4] = {tab[colour_index(d.x, minD, maxD)], tab[colour_index(d.y, minD, maxD)],
                               tab[colour_index(d.z, minD, maxD)], tab[colour_index(d.w, minD, maxD)]};
        store_rgb4(rgb + 3 * i, c);
        return;
    }
    for (int k = 0; k < kPixPerThread && i + k < n; ++k)
        store_rgb1(rgb + 3 * (i + k), tab[colour_index(depth[i + k], minD, maxD)]);
}

template <bool VEC, bool KEEP_DEPTH>
__global__ void __launch_bounds__(kBlock) depth_colorize_kernel(const int16_t* __restrict__ disp,
                                                                float* __restrict__ depth,
                                                                uint8_t* __restrict__ rgb,
                                                                long long n, DepthArgs a,
                                                                float minD, float maxD) {
    __shared__ uint32_t tab[256];
    stage_table(tab);
    const long long i = ((long long)blockIdx.x * kBlock + threadIdx.x) * kPixPerThread;
    if (i >= n) return;
    if (VEC && i + kPixPerThread <= n) {
        const uint2 raw = *reinterpret_cast<const uint2*>(disp + i);
        const float z0 = depth_of<false>(static_cast<int16_t>(raw.x & 0xFFFFu), a);
        const float z1 = depth_of<false>(static_cast<int16_t>(raw.x >> 16), a);
        const float z2 = depth_of<false>(static_cast<int16_t>(raw.y & 0xFFFFu), a);
        const float z3 = depth_of<false>(static_cast<int16_t>(raw.y >> 16), a);
        if (KEEP_DEPTH) *reinterpret_cast<float4*>(depth + i) = float4{z0, z1, z2, z3};
        const uint32_t c[4] = {tab[colour_index(z0, minD, maxD)], tab[colour_index(z1, minD, maxD)],
                               tab[colour_index(z2, minD, maxD)], tab[colour_index(z3, minD, maxD)]};
        store_rgb4(rgb + 3 * i, c);
        return;
    }
    for (int k = 0; k < kPixPerThread && i + k < n; ++k) {
        const float z = depth_of<false>(disp[i + k], a);
        if (KEEP_DEPTH) depth[i + k] = z;
        store_rgb1(rgb + 3 * (i + k), tab[colour_index(z, minD, maxD)]);
    }
}

bool aligned(const void* p, size_t a) { return (reinterpret_cast<uintptr_t>(p) & (a - 1)) == 0; }

unsigned grid_for(long long n) {
    return static_cast<unsigned>((n + (long long)kBlock * kPixPerThread - 1) / ((long long)kBlock * kPixPerThread));
}

// Maps of more than 2^31 pixels would overflow the grid; nothing this large is a camera image.
int check_map(const void* a, const void* b, int W, int H) {
    if (!a || !b) {
        set_last_error("Null buffer");
        return LUMEN_EINVAL;
    }
    if (W <= 0 || H <= 0) {
        set_last_error("Disparity map is empty");
        return LUMEN_EINVAL;
    }
    if ((long long)W * H > (1LL << 31)) {
        set_last_error("Map larger than 2^31 pixels");
        return LUMEN_EINVAL;
    }
    return LUMEN_OK;
}

hipError_t launch_depth(const int16_t* disp, float* depth, long long n, const DepthArgs& a,
                        bool host, hipStream_t s) {
    const bool vec = aligned(disp, 8) && aligned(depth, 16);
    const dim3 grid(grid_for(n)), block(kBlock);
    if (host) {
        if (vec)
            depth_kernel<true, true><<<grid, block, 0, s>>>(disp, depth, n, a);
        else
            depth_kernel<true, false><<<grid, block, 0, s>>>(disp, depth, n, a);
    } else {
        if (vec)
            depth_kernel<false, true><<<grid, block, 0, s>>>(disp, depth, n, a);
        else
            depth_kernel<false, false><<<grid, block, 0, s>>>(disp, depth, n, a);
    }
    return hipGetLastError();
}

hipError_t launch_colorize(const float* depth, uint8_t* rgb, long long n, float minD, float maxD,
                           hipStream_t s) {
    const dim3 grid(grid_for(n)), block(kBlock);
    if (aligned(depth, 16) && aligned(rgb, 4))
        colorize_kernel<true><<<grid, block, 0, s>>>(depth, rgb, n, minD, maxD);
    else
        colorize_kernel<false><<<grid, block, 0, s>>>(depth, rgb, n, minD, maxD);
    return hipGetLastError();
}

}  // namespace
}  // namespace lumen

using namespace lumen;

extern "C" int lumen_disparity_to_depth_device(const int16_t* d_disp, float* d_depth, int width,
                                               int height, float focalLength, float baseline,
                                               float disparityOffsetPx, void* stream) {
    if (int rc = check_map(d_disp, d_depth, width, height)) return rc;
    DepthArgs a{focalLength, baseline, disparityOffsetPx, 0.0, 0.0, 0.0};
    hipError_t e = launch_depth(d_disp, d_depth, (long long)width * height, a, false,
                                static_cast<hipStream_t>(stream));
    return e == hipSuccess ? LUMEN_OK : scene_hip_fail(e, "depth kernel launch");
}

extern "C" int lumen_disparity_to_depth(const int16_t* disp, float* depth, int width, int height,
                                        double focalLength, double baseline,
                                        double disparityOffsetPx) {
    if (int rc = check_map(disp, depth, width, height)) return rc;
    const long long n = (long long)width * height;
    SceneStaging& st = scene_staging();
    std::lock_guard<std::mutex> lock(st.mu);
    hipError_t e = st.ensure(0, n * 2);
    if (e == hipSuccess) e = st.ensure(1, n * 4);
    if (e != hipSuccess) return scene_hip_fail(e, "device buffers");
    DepthArgs a{0.0f, 0.0f, 0.0f, focalLength, baseline, disparityOffsetPx};
    auto* dd = static_cast<int16_t*>(st.buf[0]);
    auto* dz = static_cast<float*>(st.buf[1]);
    e = hipMemcpyAsync(dd, disp, n * 2, hipMemcpyHostToDevice, st.stream);
    if (e == hipSuccess) e = launch_depth(dd, dz, n, a, true, st.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(depth, dz, n * 4, hipMemcpyDeviceToHost, st.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st.stream);
    return e == hipSuccess ? LUMEN_OK : scene_hip_fail(e, "disparity to depth");
}

extern "C" int lumen_colorize_depth_device(const float* d_depth, uint8_t* d_rgb, int width,
                                           int height, float minDepth, float maxDepth,
                                           void* stream) {
    if (int rc = check_map(d_depth, d_rgb, width, height)) return rc;
    hipError_t e = launch_colorize(d_depth, d_rgb, (long long)width * height, minDepth, maxDepth,
                                   static_cast<hipStream_t>(stream));
    return e == hipSuccess ? LUMEN_OK : scene_hip_fail(e, "colour map kernel launch");
}

extern "C" int lumen_colorize_depth(const float* depth, uint8_t* rgb, int width, int height,
                                    float minDepth, float maxDepth) {
    if (int rc = check_map(depth, rgb, width, height)) return rc;
    const long long n = (long long)width * height;
    SceneStaging& st = scene_staging();
    std::lock_guard<std::mutex> lock(st.mu);
    hipError_t e = st.ensure(1, n * 4);
    if (e == hipSuccess) e = st.ensure(2, n * 3);
    if (e != hipSuccess) return scene_hip_fail(e, "device buffers");
    auto* dz = static_cast<float*>(st.buf[1]);
    auto* dc = static_cast<uint8_t*>(st.buf[2]);
    e = hipMemcpyAsync(dz, depth, n * 4, hipMemcpyHostToDevice, st.stream);
    if (e == hipSuccess) e = launch_colorize(dz, dc, n, minDepth, maxDepth, st.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(rgb, dc, n * 3, hipMemcpyDeviceToHost, st.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(st.stream);
    return e == hipSuccess ? LUMEN_OK : scene_hip_fail(e, "colour map");
}

extern "C" int lumen_depth_colorize_device(const int16_t* d_disp, float* d_depth, uint8_t* d_rgb,
                                           int width, int height, float focalLength,
                                           float baseline, float disparityOffsetPx,
                                           float minDepth, float maxDepth, void* stream) {
    if (int rc = check_map(d_disp, d_rgb, width, height)) return rc;
    const long long n = (long long)width * height;
    DepthArgs a{focalLength, baseline, disparityOffsetPx, 0.0, 0.0, 0.0};
    const bool vec = aligned(d_disp, 8) && aligned(d_rgb, 4) && (!d_depth || aligned(d_depth, 16));
    const dim3 grid(grid_for(n)), block(kBlock);
    auto s = static_cast<hipStream_t>(stream);
    if (d_depth) {
        if (vec)
            depth_colorize_kernel<true, true><<<grid, block, 0, s>>>(d_disp, d_depth, d_rgb, n, a,
                                                                     minDepth, maxDepth);
        else
            depth_colorize_kernel<false, true><<<grid, block, 0, s>>>(d_disp, d_depth, d_rgb, n, a,
                                                                      minDepth, maxDepth);
    } else {
        if (vec)
            depth_colorize_kernel<true, false><<<grid, block, 0, s>>>(d_disp, nullptr, d_rgb, n, a,
                                                                      minDepth, maxDepth);
        else
            depth_colorize_kernel<false, false><<<grid, block, 0, s>>>(d_disp, nullptr, d_rgb, n,
                                                                       a, minDepth, maxDepth);
    }
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? LUMEN_OK : scene_hip_fail(e, "depth colour kernel launch");
}
