/*
 * lumen_evaluate.h - C ABI of the disparity evaluation against Middlebury ground truth
 * (liblumen_sgbm.so, host code only).
 *
 * Replaces the PFM reader and the metric loop of the reference's evaluation tool
 * (tools/evaluate.cpp:45-92 and :171-221). Same conventions as lumen_sgbm.h: int status,
 * message in lumen_last_error().
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "lumen_sgbm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces loadPFM (tools/evaluate.cpp:45-92): a single-channel "Pf" file. *width and *height
 * receive the size; when `data` is not NULL and holds at least width*height floats, it receives
 * the values in native byte order with the rows top to bottom (the file stores them bottom to
 * top; a negative scale marks little-endian data). LUMEN_EINVAL for an unreadable file, another
 * magic (including colour "PF"), a non-positive size, a short payload or a too small `data`. */
int lumen_pfm_load(const char* path, int* width, int* height, float* data, size_t capacity);

/* Writes width*height floats (rows top to bottom) as a little-endian "Pf" file (scale -1).
 * No reference counterpart; the inverse of lumen_pfm_load. */
int lumen_pfm_save(const char* path, const float* data, int width, int height);

/* The metrics the reference's evaluation prints (tools/evaluate.cpp:220-245). */
typedef struct lumen_eval_result {
    int64_t valid;      /* pixels evaluated */
    double valid_pct;   /* 100 * valid / (width*height) */
    double mae, rms;    /* pixels */
    double bad1, bad2, bad4; /* % of valid pixels with error > 1, 2, 4 px */
    float p50, p90, p95, p99; /* errors[n*q/100] of the sorted errors */
} lumen_eval_result;

/* Replaces the metric loop of tools/evaluate.cpp:171-221: disparity in 1/16 px (the matcher's
 * output) against ground truth in px, both width x height. A pixel counts when the ground truth
 * is in (0, 1000] and not infinite and our disparity/16 is > 0; its error |ours - gt| is a
 * float, summed in double (squares formed in float, as written there). NaN errors sort last
 * (the reference's std::sort has no defined order for them). LUMEN_EINVAL when no pixel counts
 * ("No valid pixels to evaluate!"). */
int lumen_evaluate_disparity(const int16_t* disparity, const float* gt, int width, int height,
                             lumen_eval_result* out);

#ifdef __cplusplus
}
#endif
