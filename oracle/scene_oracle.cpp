// scene_oracle.cpp - CPU restatement of the rectification remap and the depth stages
// (test infrastructure; see scene_oracle.h for what pins each function).
#include "scene_oracle.h"

#include <cmath>
#include <cstdint>

#include "../lumenstereo_amd/csrc/turbo_table.h"

namespace {

// A map coordinate whose floor is outside +-2^30 (or NaN) has every tap outside any image this
// build accepts; the reference's int conversion is undefined there, and 0 is what any defined
// conversion gives.
bool coord_ok(float v) { return v >= -1073741824.0f && v <= 1073741824.0f; }

}  // namespace

// src/rectification.cu:201-256 (remapKernelSimple): floor, bilinear weights, taps outside the
// image count as 0 (BORDER_CONSTANT), sum in the order written, +0.5 and truncate to uint8.
extern "C" void oracle_remap(const uint8_t* src, const float* mapx, const float* mapy,
                             uint8_t* dst, int W, int H, int C) {
    const long rowStride = (long)W * C;
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            const long idx = (long)y * W + x;
            const float sx = mapx[idx], sy = mapy[idx];
            if (!coord_ok(sx) || !coord_ok(sy)) {
                for (int c = 0; c < C; ++c) dst[y * rowStride + (long)x * C + c] = 0;
                continue;
            }
            const int x0 = (int)std::floor(sx), y0 = (int)std::floor(sy);
            const int x1 = x0 + 1, y1 = y0 + 1;
            const float dx = sx - (float)x0, dy = sy - (float)y0;
            const float w00 = (1.0f - dx) * (1.0f - dy);
            const float w01 = (1.0f - dx) * dy;
            const float w10 = dx * (1.0f - dy);
            const float w11 = dx * dy;
            const bool v00 = x0 >= 0 && x0 < W && y0 >= 0 && y0 < H;
            const bool v01 = x0 >= 0 && x0 < W && y1 >= 0 && y1 < H;
            const bool v10 = x1 >= 0 && x1 < W && y0 >= 0 && y0 < H;
            const bool v11 = x1 >= 0 && x1 < W && y1 >= 0 && y1 < H;
            for (int c = 0; c < C; ++c) {
                const float a = v00 ? (float)src[y0 * rowStride + (long)x0 * C + c] : 0.0f;
                const float b = v01 ? (float)src[y1 * rowStride + (long)x0 * C + c] : 0.0f;
                const float e = v10 ? (float)src[y0 * rowStride + (long)x1 * C + c] : 0.0f;
                const float f = v11 ? (float)src[y1 * rowStride + (long)x1 * C + c] : 0.0f;
                const float val = w00 * a + w01 * b + w10 * e + w11 * f;
                dst[y * rowStride + (long)x * C + c] = (uint8_t)(val + 0.5f);
            }
        }
    }
}

// src/depth_map.cu:74-101 (disparityToDepthKernel): float arithmetic, f*B formed in float.
extern "C" void oracle_disparity_to_depth(const int16_t* disp, float* depth, int W, int H,
                                          float f, float B, float doffs) {
    const float fb = f * B;
    for (long i = 0; i < (long)W * H; ++i) {
        const int16_t d = disp[i];
        if (d <= 0) {
            depth[i] = 0.0f;
            continue;
        }
        const float denom = (float)d / 16.0f + doffs;
        depth[i] = denom > 0.0f ? fb / denom : 0.0f;
    }
}

// src/depth_map.cu:142-160 (disparityToDepth, host): f and B are double, so the quotient is
// formed in double and rounded to float once.
extern "C" void oracle_disparity_to_depth_host(const int16_t* disp, float* depth, int W, int H,
                                               double f, double B, double doffs) {
    for (long i = 0; i < (long)W * H; ++i) {
        const int16_t d = disp[i];
        if (d <= 0) {
            depth[i] = 0.0f;
            continue;
        }
        const float denom = (float)d / 16.0f + (float)doffs;
        depth[i] = denom > 0.0f ? (float)(f * B / (double)denom) : 0.0f;
    }
}

extern "C" uint32_t oracle_turbo_entry(int i) {
    return (uint32_t)lumen::kTurboR[i] | ((uint32_t)lumen::kTurboG[i] << 8) |
           ((uint32_t)lumen::kTurboB[i] << 16);
}

// src/depth_map.cu:106-138 (colorizeDepthKernel): invalid (<= 0 or beyond 2*maxDepth) is
// entry 0; otherwise the inverted, clamped normalised depth times 255, truncated. fminf/fmaxf
// return the non-NaN operand, so a NaN depth lands on entry 0 as well.
extern "C" void oracle_colorize_depth(const float* depth, uint8_t* rgb, int W, int H,
                                      float minDepth, float maxDepth) {
    for (long i = 0; i < (long)W * H; ++i) {
        const float d = depth[i];
        int ci;
        if (d <= 0.0f || d > maxDepth * 2.0f) {
            ci = 0;
        } else {
            float n = (d - minDepth) / (maxDepth - minDepth);
            n = std::fmax(0.0f, std::fmin(1.0f, n));
            n = 1.0f - n;
            ci = (int)(n * 255.0f);
        }
        rgb[3 * i + 0] = lumen::kTurboR[ci];
        rgb[3 * i + 1] = lumen::kTurboG[ci];
        rgb[3 * i + 2] = lumen::kTurboB[ci];
    }
}
