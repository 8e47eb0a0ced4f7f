/*
 * lumen_depth.h - C ABI of the stages after the matcher on MI355X (liblumen_sgbm.so):
 * disparity -> depth and the depth colour map.
 *
 * Same conventions as lumen_sgbm.h: extern "C", plain pointers and sizes, int status,
 * no exception crosses the boundary; errors are reported by lumen_last_error().
 *
 * Disparity maps are int16 in 1/16 px (the matcher's output); depth maps are float metres,
 * 0 where invalid; colour maps are 3 bytes per pixel in R, G, B order (the reference's GPU
 * kernel writes its R table first, src/depth_map.cu:134-137).
 *
 * Threading: every function may be called from several threads. The host-buffer forms share
 * one locked device staging area with lumen_remap; the device forms only enqueue on the
 * caller's stream.
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "lumen_sgbm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces disparityToDepthGPU (src/depth_map.cu:177-197; kernel :74-101): device buffers,
 * depth = (f*B) / (d/16 + doffs) in IEEE single precision with f*B formed in float, 0 where
 * d <= 0 or the denominator is <= 0. Enqueued on `stream` (hipStream_t as void*, NULL = the
 * null stream); does not wait. */
int lumen_disparity_to_depth_device(const int16_t* d_disp, float* d_depth, int width,
                                    int height, float focalLength, float baseline,
                                    float disparityOffsetPx, void* stream);

/* Replaces disparityToDepth (src/depth_map.cu:142-160), the host function callers use
 * (viewer/main.cpp:105, tools/benchmark.cpp:129): host buffers, computed on the device with the
 * host function's arithmetic (f*B / denom in double, rounded to float once). Blocking. */
int lumen_disparity_to_depth(const int16_t* disp, float* depth, int width, int height,
                             double focalLength, double baseline, double disparityOffsetPx);

/* Replaces colorizeDepthGPU (src/depth_map.cu:211-221; kernel :106-138): device buffers.
 * Invalid depth (<= 0 or > 2*maxDepth) takes colour entry 0; otherwise the entry is
 * (int)(255 * (1 - clamp((depth - minDepth) / (maxDepth - minDepth), 0, 1))). The table is the
 * reference's. Enqueued on `stream`; does not wait. */
int lumen_colorize_depth_device(const float* d_depth, uint8_t* d_rgb, int width, int height,
                                float minDepth, float maxDepth, void* stream);

/* Host-buffer form of lumen_colorize_depth_device (the same colours). Blocking. */
int lumen_colorize_depth(const float* depth, uint8_t* rgb, int width, int height,
                         float minDepth, float maxDepth);

/* Both stages in one pass (no reference counterpart): disparity in, colour out, depth written
 * too when d_depth is not NULL. Every value equals lumen_disparity_to_depth_device followed by
 * lumen_colorize_depth_device; the intermediate depth map is not read back from memory.
 * Enqueued on `stream`; does not wait. */
int lumen_depth_colorize_device(const int16_t* d_disp, float* d_depth, uint8_t* d_rgb,
                                int width, int height, float focalLength, float baseline,
                                float disparityOffsetPx, float minDepth, float maxDepth,
                                void* stream);

#ifdef __cplusplus
}
#endif
