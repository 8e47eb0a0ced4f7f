/*
 * lumen_rectify.h - C ABI of the stereo rectification remap on MI355X (liblumen_sgbm.so).
 *
 * The stage before the matcher: every output pixel samples the raw image at the source
 * coordinates of a precomputed map with bilinear interpolation. Same conventions as
 * lumen_sgbm.h: extern "C", plain pointers and sizes, int status (LUMEN_OK, LUMEN_EINVAL,
 * LUMEN_EDEVICE, LUMEN_ENOMEM), no exception crosses the boundary. Errors raised without a
 * handle are reported by lumen_last_error().
 *
 * Threading: a rectifier handle must not be used by two threads at once. lumen_remap (the
 * host-buffer form) may be called from several threads: the device staging it shares with the
 * depth and colour host-buffer calls is locked.
 *
 * Arithmetic follows remapKernelSimple (src/rectification.cu:201-256 of the reference): the
 * floor of the map coordinate selects four taps, taps outside the image count as 0, the
 * weights (1-dx)(1-dy), (1-dx)dy, dx(1-dy), dx*dy are applied in that order in IEEE single
 * precision (no fused multiply-add), and the sum + 0.5 is truncated to uint8. A map entry that
 * is NaN or beyond +-2^30 gives 0 (every tap is outside the image; the reference's int
 * conversion is undefined there).
 */
#pragma once
#include <stddef.h>
#include <stdint.h>

#include "lumen_sgbm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lumen_rectifier lumen_rectifier;

/* Replaces StereoRectifier::StereoRectifier(const CameraParams&) (src/rectification.cu:262-284).
 * The four maps are width x height floats in host memory (CameraParams::map0x, map0y, map1x,
 * map1y: left x, left y, right x, right y) and are uploaded once. If map0x or map0y is NULL the
 * pair is taken as pre-rectified, as when the reference's map0x/map0y are empty: rectify then
 * copies its inputs, of any size, and width/height are not checked. map1x/map1y may only be
 * NULL in that case. */
int lumen_rectifier_create(const float* map0x, const float* map0y, const float* map1x,
                           const float* map1y, int width, int height, lumen_rectifier** out);

/* Frees the maps, buffers and stream. NULL is accepted. */
int lumen_rectifier_destroy(lumen_rectifier* r);

/* Replaces StereoRectifier::isPreRectified (include/stereo/rectification.h:68): 1 or 0, or
 * -1 for a NULL handle. */
int lumen_rectifier_is_pre_rectified(const lumen_rectifier* r);

/* Replaces StereoRectifier::rectify (src/rectification.cu:286-352): host images of width x
 * height with `channels` interleaved 8-bit channels (packed rows), written to left_out and
 * right_out (same shape). Blocking. The image size must equal the maps' size (the reference
 * reads past its maps otherwise). Both images are remapped by one launch. */
int lumen_rectifier_rectify(lumen_rectifier* r, const uint8_t* left, const uint8_t* right,
                            int width, int height, int channels, uint8_t* left_out,
                            uint8_t* right_out);

/* Device-resident form (no reference counterpart): packed device images in and out, work
 * enqueued on the handle's stream; waits for it when `sync` is non-zero. */
int lumen_rectifier_rectify_device(lumen_rectifier* r, const uint8_t* d_left,
                                   const uint8_t* d_right, int width, int height, int channels,
                                   uint8_t* d_left_out, uint8_t* d_right_out, int sync);

/* The hipStream_t the handle launches on (as void*). */
int lumen_rectifier_get_stream(lumen_rectifier* r, void** stream_out);

/* Time of the last rectify's remap launch in milliseconds (hipEvents around it on the handle's
 * stream). Waits for it. */
int lumen_rectifier_get_time_ms(lumen_rectifier* r, float* ms_out);

/* Text of the last error on this handle ("" if none). */
const char* lumen_rectifier_last_error(const lumen_rectifier* r);

/* Replaces rectifyImageGPU (src/rectification.cu:363-420): one image and its maps, all in host
 * memory, remapped on the device. Blocking. */
int lumen_remap(const uint8_t* src, const float* mapx, const float* mapy, uint8_t* dst,
                int width, int height, int channels);

/* Device-pointer form of lumen_remap (no reference counterpart), enqueued on `stream` (a
 * hipStream_t as void*, NULL = the null stream); does not wait. */
int lumen_remap_device(const uint8_t* d_src, const float* d_mapx, const float* d_mapy,
                       uint8_t* d_dst, int width, int height, int channels, void* stream);

#ifdef __cplusplus
}
#endif
