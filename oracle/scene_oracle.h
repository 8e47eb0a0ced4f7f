/*
 * scene_oracle.h - CPU restatement of the stages either side of the matcher (test
 * infrastructure): the rectification remap before it, disparity -> depth and the depth colour
 * map after it.
 *
 * THIS IS THE ORACLE, NOT THE PRODUCT. Only tests/ and bench.py's cpu_baseline leg may load it,
 * and only as the checker (or as the CPU baseline reported beside the GPU figure).
 *
 * Each function restates one reference kernel or host function of Pfactorial01/LumenStereo;
 * the file:line it follows is cited on each definition in scene_oracle.cpp. Floating point is
 * IEEE single precision evaluated exactly as written (built with -ffp-contract=off).
 *
 * Pinning (what checks each function independently of this file):
 *   remap            identity map gives the source exactly; half-pixel shift gives the rounded
 *                    mean of four taps (closed form); taps outside the image count as 0; a
 *                    float64 bilinear restatement in the tests agrees within one grey level
 *   depth            closed form f*B/(d/16 + doffs) (float64 within 1 ulp of float32 result),
 *                    invalid (d <= 0, denominator <= 0) gives 0
 *   colorize         colour table equal to the reference's, read from its source into a
 *                    fixture (tests/golden/turbo_table.npz); index closed form on exact depths
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Bilinear remap of a W x H image with `channels` interleaved channels through per-pixel
 * source coordinates mapx/mapy (W x H floats). */
void oracle_remap(const uint8_t* src, const float* mapx, const float* mapy, uint8_t* dst,
                  int W, int H, int channels);

/* disparity (int16, 1/16 px) -> depth (float, metres) as the GPU kernel computes it. */
void oracle_disparity_to_depth(const int16_t* disp, float* depth, int W, int H,
                               float focalLength, float baseline, float doffs);

/* The same conversion as the reference's host function computes it (double f*B / denom). */
void oracle_disparity_to_depth_host(const int16_t* disp, float* depth, int W, int H,
                                    double focalLength, double baseline, double doffs);

/* depth -> RGB (3 bytes per pixel) through the reference's 256-entry colour table. */
void oracle_colorize_depth(const float* depth, uint8_t* rgb, int W, int H, float minDepth,
                           float maxDepth);

/* Colour table entry i (0..255) as R | G << 8 | B << 16. */
uint32_t oracle_turbo_entry(int i);

#ifdef __cplusplus
}
#endif
