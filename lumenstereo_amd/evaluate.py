"""Evaluation against Middlebury ground truth: Python mirror of the reference's evaluation tool
(tools/evaluate.cpp:45-92 loadPFM, :171-245 metrics), bound to liblumen_sgbm.so
(include/lumen_evaluate.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import CEvalResult, StereoException, library

__all__ = ["loadPFM", "savePFM", "evaluateDisparity"]


def loadPFM(path) -> np.ndarray:
    """A single-channel PFM as float32 H x W, rows top to bottom; raises StereoException."""
    lib = library()
    w, h = ctypes.c_int(), ctypes.c_int()
    p = str(path).encode()
    if lib.lumen_pfm_load(p, ctypes.byref(w), ctypes.byref(h), None, 0) != 0:
        raise StereoException(lib.lumen_last_error().decode())
    out = np.empty((h.value, w.value), np.float32)
    if lib.lumen_pfm_load(p, ctypes.byref(w), ctypes.byref(h), out.ctypes.data, out.size) != 0:
        raise StereoException(lib.lumen_last_error().decode())
    return out


def savePFM(path, data: np.ndarray) -> None:
    a = np.ascontiguousarray(data, np.float32)
    lib = library()
    if a.ndim != 2 or lib.lumen_pfm_save(str(path).encode(), a.ctypes.data, a.shape[1],
                                         a.shape[0]) != 0:
        raise StereoException(lib.lumen_last_error().decode() if a.ndim == 2 else "2-D only")


def evaluateDisparity(disparity: np.ndarray, gt: np.ndarray) -> dict:
    """Metrics of int16 disparity (1/16 px) against float ground truth (px), as the reference's
    tool prints them: valid, valid_pct, mae, rms, bad1, bad2, bad4 (%), p50..p99 (px)."""
    d = np.ascontiguousarray(disparity, np.int16)
    g = np.ascontiguousarray(gt, np.float32)
    if d.shape != g.shape or d.ndim != 2:
        raise StereoException("disparity and ground truth must have the same 2-D shape")
    r = CEvalResult()
    lib = library()
    if lib.lumen_evaluate_disparity(d.ctypes.data, g.ctypes.data, d.shape[1], d.shape[0],
                                    ctypes.byref(r)) != 0:
        raise StereoException(lib.lumen_last_error().decode())
    return {k: getattr(r, k) for k, _ in CEvalResult._fields_}
