/*
 * lumen_config.h - C ABI of the parameter and calibration files (liblumen_sgbm.so, host code
 * only: nothing here touches the GPU).
 *
 * Replaces the reference's YAML parameter files (src/config.cpp:17-65, which use yaml-cpp) and
 * its Middlebury calib.txt reader (src/config.cpp:149-236). yaml-cpp is not used: the parser
 * reads the subset these files are written in - a top-level block mapping of scalars, with
 * comments, blank lines, quoted scalars and document markers - and converts scalars as
 * yaml-cpp's as<int>/as<bool>/as<std::string> do (decimal, 0x hex and 0o octal integers; y/yes/
 * true/on and n/no/false/off in lower, Capitalised or UPPER case). A known key whose value is
 * not a scalar, or not convertible, fails the load as yaml-cpp's exception does.
 *
 * Same conventions as lumen_sgbm.h: int status, message in lumen_last_error().
 */
#pragma once
#include <stdint.h>

#include "lumen_sgbm.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Replaces loadStereoParams (src/config.cpp:17-45). Reads minDisparity, maxDisparity,
 * blockSize, P1, P2, uniquenessRatio, disp12MaxDiff, preFilterCap, useGPU, numDirections and
 * matchingCostMode ("Census" or "census" select Census, any other string SAD) into *params,
 * leaving every field the file does not name (speckleWindowSize and speckleRange are never
 * read, as in the reference) unchanged. Returns LUMEN_OK when the file was read and the result
 * satisfies StereoParams::isValid; LUMEN_EINVAL otherwise (unreadable or malformed file: *params
 * unchanged; invalid result: *params holds what was read, as the reference leaves it). */
int lumen_params_load_yaml(const char* path, lumen_stereo_params* params);

/* Replaces saveStereoParams (src/config.cpp:47-72): writes the eleven keys above in the
 * reference's order as a YAML block mapping. */
int lumen_params_save_yaml(const char* path, const lumen_stereo_params* params);

/* The calibration record of a rectified pair (the numeric part of stereo::CameraParams,
 * include/stereo/stereo_params.h:135-186). Matrices are row-major doubles. */
typedef struct lumen_camera_params {
    int32_t width;
    int32_t height;
    double K0[9];
    double D0[5];
    double K1[9];
    double D1[5];
    double R[9];
    double T[3];
    double baseline;     /* metres */
    double focalLength;  /* pixels */
    int32_t disparityOffset;
    int32_t numDisparities;
} lumen_camera_params;

/* Replaces loadMiddleburyCalib (src/config.cpp:149-236): reads cam0=[..], cam1=[..], doffs=,
 * baseline= (mm), width=, height=, ndisp= lines. K0/K1 take fx, cx, fy, cy from the camera
 * matrices (identity otherwise), D0/D1 are zero, R is the identity, T = (-baseline/1000, 0, 0),
 * baseline is in metres, focalLength = fx of cam0, disparityOffset = (int)(doffs + 0.5),
 * numDisparities = ndisp, or 280 when absent. Fields the file does not set keep the values
 * *params had for width and height. Returns LUMEN_OK when width and height are positive,
 * LUMEN_EINVAL when the file cannot be opened, a number cannot be parsed (the reference's
 * std::stod/stoi would throw) or the size is missing. */
int lumen_middlebury_calib_load(const char* path, lumen_camera_params* params);

#ifdef __cplusplus
}
#endif
