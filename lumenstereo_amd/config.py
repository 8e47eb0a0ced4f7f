"""Parameter and calibration files: Python mirror of the reference's loadStereoParams,
saveStereoParams and loadMiddleburyCalib (include/stereo/stereo_params.h:187-211,
src/config.cpp:17-72 and :149-236), bound to liblumen_sgbm.so (include/lumen_config.h).
Like the reference they return a bool and update the record they are given."""
from __future__ import annotations

import ctypes

import numpy as np

from . import CCameraParams, MatchingCostMode, StereoParams, library
from .scene import CameraParams

__all__ = ["loadStereoParams", "saveStereoParams", "loadMiddleburyCalib", "CCameraParams",
           "last_error"]


def last_error() -> str:
    return library().lumen_last_error().decode()


def loadStereoParams(path: str, params: StereoParams) -> bool:
    """src/config.cpp:17-45: True when the file was read and the parameters are valid."""
    c = params.to_c()
    rc = library().lumen_params_load_yaml(str(path).encode(), ctypes.byref(c))
    q = StereoParams.from_c(c)
    for f in ("minDisparity", "maxDisparity", "blockSize", "P1", "P2", "preFilterCap",
              "uniquenessRatio", "disp12MaxDiff", "useGPU", "numDirections"):
        setattr(params, f, getattr(q, f))
    params.matchingCostMode = MatchingCostMode(c.matchingCostMode)
    return rc == 0


def saveStereoParams(path: str, params: StereoParams) -> bool:
    """src/config.cpp:47-72."""
    c = params.to_c()
    return library().lumen_params_save_yaml(str(path).encode(), ctypes.byref(c)) == 0


def loadMiddleburyCalib(path: str, params: CameraParams) -> bool:
    """src/config.cpp:149-236: fills K0/K1 (3x3), D0/D1 (5x1), R (3x3), T (3x1), baseline (m),
    focalLength, disparityOffset, numDisparities and, when present, width/height."""
    c = CCameraParams()
    c.width, c.height = int(params.width), int(params.height)
    rc = library().lumen_middlebury_calib_load(str(path).encode(), ctypes.byref(c))
    if rc != 0 and not last_error().startswith("Middlebury calib file has no image size"):
        return False
    params.width, params.height = c.width, c.height
    params.K0 = np.array(c.K0, np.float64).reshape(3, 3)
    params.K1 = np.array(c.K1, np.float64).reshape(3, 3)
    params.D0 = np.array(c.D0, np.float64).reshape(5, 1)
    params.D1 = np.array(c.D1, np.float64).reshape(5, 1)
    params.R = np.array(c.R, np.float64).reshape(3, 3)
    params.T = np.array(c.T, np.float64).reshape(3, 1)
    params.baseline = c.baseline
    params.focalLength = c.focalLength
    params.disparityOffset = c.disparityOffset
    params.numDisparities = c.numDisparities
    return rc == 0
