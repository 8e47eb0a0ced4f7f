"""MI355X SGBM matcher: Python host mirror of the reference's stereo::StereoSGBM interface.

The compute path is liblumen_sgbm.so (hand-written gfx950 HIP kernels behind a C ABI, declared
in include/lumen_sgbm.h). This module only binds that ABI with ctypes and mirrors the
reference's types so that Python callers and tests read like the reference's own:

* ``StereoParams`` / ``MatchingCostMode`` mirror include/stereo/stereo_params.h:20-104.
* ``StereoSGBM`` mirrors include/stereo/sgbm_gpu.h:29-91 (compute, setParams, getParams and the
  four timing getters).
* ``StereoException`` is raised for the validation errors of src/sgbm_gpu.cu:148-166, and
  ``GpuException`` for runtime/device errors (the reference's CudaException,
  include/stereo/common.h:31-48).

There is no fallback: if the shared library is missing, importing the binding raises.
"""
from __future__ import annotations

import ctypes
import dataclasses
import enum
import os
import threading
from pathlib import Path

import numpy as np

__all__ = [
    "MatchingCostMode",
    "StereoParams",
    "StereoSGBM",
    "StereoException",
    "GpuException",
    "kSgbmMaxDisparityRange",
    "middleburyMatcherMaxDisparity",
    "synth_pair",
    "library",
    "LIB_PATH",
    "STAGE",
    "BatchSGBM",
    "ServeSGBM",
    "now_ns",
    "tuning_load",
    "tuning_clear",
]

LIB_PATH = Path(__file__).resolve().parent / "liblumen_sgbm.so"

# include/stereo/sgbm_limits.h:9
kSgbmMaxDisparityRange = 512

LUMEN_OK, LUMEN_EINVAL, LUMEN_EDEVICE, LUMEN_ENOMEM, LUMEN_ETIMEDOUT = 0, 1, 2, 3, 4

STAGE = {
    "cost": 0,
    "path": 1,
    "sum": 2,
    "wta_int": 3,
    "disp_left": 4,
    "disp_right": 5,
    "after_lr": 6,
    "after_speckle": 7,
    "output": 8,
}


def middleburyMatcherMaxDisparity(calibNumDisparities: int, fallbackWhenMissing: int = 96) -> int:
    """include/stereo/sgbm_limits.h:18-22."""
    ndisp = calibNumDisparities if calibNumDisparities > 0 else fallbackWhenMissing
    return min(ndisp, kSgbmMaxDisparityRange)


class StereoException(RuntimeError):
    """Invalid input or parameters (stereo::StereoException, include/stereo/common.h:105-110)."""


class GpuException(RuntimeError):
    """Device/runtime failure. Replaces the reference's CudaException
    (include/stereo/common.h:31-48); ``error_code`` is the library status code."""

    def __init__(self, msg: str, error_code: int = LUMEN_EDEVICE):
        super().__init__(msg)
        self.error_code = error_code


class MatchingCostMode(enum.IntEnum):
    """include/stereo/stereo_params.h:20-23."""

    SAD = 0
    Census = 1


@dataclasses.dataclass
class StereoParams:
    """include/stereo/stereo_params.h:57-104 (same fields, same defaults)."""

    minDisparity: int = 0
    maxDisparity: int = 128
    blockSize: int = 5
    P1: int = 200
    P2: int = 800
    preFilterCap: int = 31
    uniquenessRatio: int = 10
    disp12MaxDiff: int = 1
    speckleWindowSize: int = 100
    speckleRange: int = 32
    useGPU: bool = True
    useSharedMemory: bool = True
    numDirections: int = 4
    matchingCostMode: MatchingCostMode = MatchingCostMode.SAD

    def autoComputePenalties(self, numChannels: int = 1) -> None:
        """include/stereo/stereo_params.h:88-91."""
        self.P1 = 8 * numChannels * self.blockSize * self.blockSize
        self.P2 = 32 * numChannels * self.blockSize * self.blockSize

    def isValid(self) -> bool:
        """include/stereo/stereo_params.h:96-103 (not called by compute, as in the reference)."""
        if self.blockSize < 1 or self.blockSize % 2 == 0:
            return False
        if self.minDisparity < 0:
            return False
        if self.maxDisparity <= self.minDisparity:
            return False
        if self.P2 <= self.P1:
            return False
        if self.uniquenessRatio < 0 or self.uniquenessRatio > 100:
            return False
        return True

    def to_c(self) -> "CParams":
        return CParams(
            int(self.minDisparity), int(self.maxDisparity), int(self.blockSize), int(self.P1),
            int(self.P2), int(self.preFilterCap), int(self.uniquenessRatio),
            int(self.disp12MaxDiff), int(self.speckleWindowSize), int(self.speckleRange),
            int(bool(self.useGPU)), int(bool(self.useSharedMemory)), int(self.numDirections),
            int(self.matchingCostMode),
        )

    @classmethod
    def from_c(cls, c: "CParams") -> "StereoParams":
        return cls(
            c.minDisparity, c.maxDisparity, c.blockSize, c.P1, c.P2, c.preFilterCap,
            c.uniquenessRatio, c.disp12MaxDiff, c.speckleWindowSize, c.speckleRange,
            bool(c.useGPU), bool(c.useSharedMemory), c.numDirections,
            MatchingCostMode(c.matchingCostMode) if c.matchingCostMode in (0, 1) else c.matchingCostMode,
        )


class CParams(ctypes.Structure):
    """lumen_stereo_params (include/lumen_sgbm.h)."""

    _fields_ = [(name, ctypes.c_int32) for name in (
        "minDisparity", "maxDisparity", "blockSize", "P1", "P2", "preFilterCap",
        "uniquenessRatio", "disp12MaxDiff", "speckleWindowSize", "speckleRange", "useGPU",
        "useSharedMemory", "numDirections", "matchingCostMode")]


class CPairDesc(ctypes.Structure):
    """lumen_pair_desc (include/lumen_sgbm.h)."""

    _fields_ = [("params", CParams), ("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("left", ctypes.c_void_p), ("right", ctypes.c_void_p), ("out", ctypes.c_void_p)]


class CSchedule(ctypes.Structure):
    """lumen_schedule (include/lumen_sgbm.h)."""

    _fields_ = [("agg_prefetch", ctypes.c_int32), ("agg_priority", ctypes.c_int32),
                ("census_tile", ctypes.c_int32), ("agg_wide", ctypes.c_int32),
                ("agg_occupancy", ctypes.c_int32)]


class CServerConfig(ctypes.Structure):
    """lumen_server_config (include/lumen_sgbm.h)."""

    _fields_ = [("pool_bytes", ctypes.c_size_t), ("lanes", ctypes.c_int32),
                ("max_batch", ctypes.c_int32), ("device_io", ctypes.c_int32),
                ("max_bypass", ctypes.c_int32), ("batch_budget_us", ctypes.c_int32),
                ("pinned_mb", ctypes.c_int32), ("large_us", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


class CRequestInfo(ctypes.Structure):
    """lumen_request_info (include/lumen_sgbm.h)."""

    _fields_ = [("status", ctypes.c_int32), ("lane", ctypes.c_int32),
                ("batch_size", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("t_submit_ns", ctypes.c_int64), ("t_admit_ns", ctypes.c_int64),
                ("t_start_ns", ctypes.c_int64), ("t_done_ns", ctypes.c_int64)]


class CServerStats(ctypes.Structure):
    """lumen_server_stats (include/lumen_sgbm.h)."""

    _fields_ = [(n, ctypes.c_uint64) for n in (
        "submitted", "completed", "failed", "batches", "max_batch_seen", "max_lanes_busy",
        "pool_bytes", "pool_high_water", "max_waiting", "direct_starts", "unstaged")]


class CCameraParams(ctypes.Structure):
    """lumen_camera_params (include/lumen_config.h)."""

    _fields_ = [("width", ctypes.c_int32), ("height", ctypes.c_int32),
                ("K0", ctypes.c_double * 9), ("D0", ctypes.c_double * 5),
                ("K1", ctypes.c_double * 9), ("D1", ctypes.c_double * 5),
                ("R", ctypes.c_double * 9), ("T", ctypes.c_double * 3),
                ("baseline", ctypes.c_double), ("focalLength", ctypes.c_double),
                ("disparityOffset", ctypes.c_int32), ("numDisparities", ctypes.c_int32)]


class CEvalResult(ctypes.Structure):
    """lumen_eval_result (include/lumen_evaluate.h)."""

    _fields_ = [("valid", ctypes.c_int64), ("valid_pct", ctypes.c_double),
                ("mae", ctypes.c_double), ("rms", ctypes.c_double), ("bad1", ctypes.c_double),
                ("bad2", ctypes.c_double), ("bad4", ctypes.c_double), ("p50", ctypes.c_float),
                ("p90", ctypes.c_float), ("p95", ctypes.c_float), ("p99", ctypes.c_float)]


_lib = None

# Every symbol include/lumen_*.h declares, with its ctypes signature.
_c_int, _c_void_p, _c_size_t = ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t
_SIGNATURES = {
    "lumen_sgbm_default_params": (_c_int, [ctypes.POINTER(CParams)]),
    "lumen_sgbm_create": (_c_int, [ctypes.POINTER(CParams), ctypes.POINTER(_c_void_p)]),
    "lumen_sgbm_destroy": (_c_int, [_c_void_p]),
    "lumen_sgbm_set_params": (_c_int, [_c_void_p, ctypes.POINTER(CParams)]),
    "lumen_sgbm_get_params": (_c_int, [_c_void_p, ctypes.POINTER(CParams)]),
    "lumen_sgbm_compute": (_c_int, [_c_void_p, _c_void_p, _c_size_t, _c_void_p, _c_size_t, _c_int,
                                    _c_int, _c_int, _c_void_p, _c_size_t]),
    "lumen_sgbm_compute_device": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_int,
                                           _c_void_p, _c_int]),
    "lumen_sgbm_compute_device_noalloc": (_c_int, [ctypes.POINTER(CParams), _c_void_p, _c_void_p,
                                                   _c_int, _c_int, _c_void_p, _c_void_p]),
    "lumen_sgbm_get_stream": (_c_int, [_c_void_p, ctypes.POINTER(_c_void_p)]),
    "lumen_device_alloc": (_c_int, [_c_size_t, ctypes.POINTER(_c_void_p)]),
    "lumen_device_free": (_c_int, [_c_void_p]),
    "lumen_device_copy": (_c_int, [_c_void_p, _c_void_p, _c_size_t, _c_int]),
    "lumen_device_zero": (_c_int, [_c_void_p, _c_size_t]),
    "lumen_sgbm_get_timings": (_c_int, [_c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "lumen_sgbm_get_timings_avg": (_c_int, [_c_void_p, _c_int, ctypes.POINTER(ctypes.c_float)]),
    "lumen_sgbm_set_timing": (_c_int, [_c_void_p, _c_int]),
    "lumen_sgbm_last_error": (ctypes.c_char_p, [_c_void_p]),
    "lumen_last_error": (ctypes.c_char_p, []),
    "lumen_sgbm_dump_stage": (_c_int, [_c_void_p, _c_int, _c_int, _c_void_p, _c_size_t]),
    "lumen_sgbm_num_paths": (_c_int, [_c_void_p]),
    "lumen_sgbm_footprint": (_c_int, [ctypes.POINTER(CParams), _c_int, _c_int,
                                      ctypes.POINTER(_c_size_t)]),
    "lumen_synth_pair": (_c_int, [_c_int, _c_int, ctypes.c_float, ctypes.c_float,
                                  ctypes.c_uint64, ctypes.c_float, _c_void_p, _c_void_p,
                                  _c_void_p]),
    "lumen_batch_create": (_c_int, [_c_size_t, ctypes.POINTER(_c_void_p)]),
    "lumen_batch_destroy": (_c_int, [_c_void_p]),
    "lumen_batch_arena_bytes": (_c_int, [_c_void_p, ctypes.POINTER(_c_size_t)]),
    "lumen_batch_run": (_c_int, [_c_void_p, ctypes.POINTER(CPairDesc), _c_int, _c_int]),
    "lumen_batch_get_timings": (_c_int, [_c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "lumen_batch_last_error": (ctypes.c_char_p, [_c_void_p]),
    "lumen_now_ns": (ctypes.c_int64, []),
    "lumen_sgbm_set_schedule": (_c_int, [_c_void_p, ctypes.POINTER(CSchedule)]),
    "lumen_sgbm_get_schedule": (_c_int, [_c_void_p, _c_int, _c_int, ctypes.POINTER(CSchedule)]),
    "lumen_tuning_load": (_c_int, [ctypes.c_char_p, ctypes.POINTER(_c_int)]),
    "lumen_tuning_clear": (_c_int, []),
    "lumen_server_create": (_c_int, [ctypes.POINTER(CServerConfig), ctypes.POINTER(_c_void_p)]),
    "lumen_server_destroy": (_c_int, [_c_void_p]),
    "lumen_server_submit": (_c_int, [_c_void_p, ctypes.POINTER(CPairDesc),
                                     ctypes.POINTER(ctypes.c_uint64)]),
    "lumen_server_wait": (_c_int, [_c_void_p, ctypes.c_uint64, ctypes.c_int64,
                                   ctypes.POINTER(CRequestInfo)]),
    "lumen_server_drain": (_c_int, [_c_void_p]),
    "lumen_server_get_stats": (_c_int, [_c_void_p, ctypes.POINTER(CServerStats)]),
    "lumen_server_last_error": (ctypes.c_char_p, [_c_void_p]),
    # include/lumen_config.h
    "lumen_params_load_yaml": (_c_int, [ctypes.c_char_p, ctypes.POINTER(CParams)]),
    "lumen_params_save_yaml": (_c_int, [ctypes.c_char_p, ctypes.POINTER(CParams)]),
    "lumen_middlebury_calib_load": (_c_int, [ctypes.c_char_p, ctypes.POINTER(CCameraParams)]),
    # include/lumen_evaluate.h
    "lumen_pfm_load": (_c_int, [ctypes.c_char_p, ctypes.POINTER(_c_int), ctypes.POINTER(_c_int),
                                _c_void_p, _c_size_t]),
    "lumen_pfm_save": (_c_int, [ctypes.c_char_p, _c_void_p, _c_int, _c_int]),
    "lumen_evaluate_disparity": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int,
                                          ctypes.POINTER(CEvalResult)]),
    # include/lumen_rectify.h
    "lumen_rectifier_create": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_int,
                                        _c_int, ctypes.POINTER(_c_void_p)]),
    "lumen_rectifier_destroy": (_c_int, [_c_void_p]),
    "lumen_rectifier_is_pre_rectified": (_c_int, [_c_void_p]),
    "lumen_rectifier_rectify": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_int, _c_int,
                                         _c_void_p, _c_void_p]),
    "lumen_rectifier_rectify_device": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_int,
                                                _c_int, _c_void_p, _c_void_p, _c_int]),
    "lumen_rectifier_get_stream": (_c_int, [_c_void_p, ctypes.POINTER(_c_void_p)]),
    "lumen_rectifier_get_time_ms": (_c_int, [_c_void_p, ctypes.POINTER(ctypes.c_float)]),
    "lumen_rectifier_last_error": (ctypes.c_char_p, [_c_void_p]),
    "lumen_remap": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_int, _c_int, _c_int]),
    "lumen_remap_device": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_void_p, _c_int, _c_int,
                                    _c_int, _c_void_p]),
    # include/lumen_depth.h
    "lumen_disparity_to_depth_device": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int,
                                                 ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                                 _c_void_p]),
    "lumen_disparity_to_depth": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, ctypes.c_double,
                                          ctypes.c_double, ctypes.c_double]),
    "lumen_colorize_depth_device": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, ctypes.c_float,
                                             ctypes.c_float, _c_void_p]),
    "lumen_colorize_depth": (_c_int, [_c_void_p, _c_void_p, _c_int, _c_int, ctypes.c_float,
                                      ctypes.c_float]),
    "lumen_depth_colorize_device": (_c_int, [_c_void_p, _c_void_p, _c_void_p, _c_int, _c_int,
                                             ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                             ctypes.c_float, ctypes.c_float, _c_void_p]),
}


def library() -> ctypes.CDLL:
    """Load liblumen_sgbm.so (built in-tree by __graft_entry__.build()). Raises if absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (there is no CPU fallback)")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def synth_pair(width: int, height: int, dmin: float, dmax: float, seed: int,
               lowtex_frac: float = 0.0, with_gt: bool = False):
    """Deterministic synthetic pair (see include/lumen_sgbm.h, lumen_synth_pair)."""
    lib = library()
    left = np.empty((height, width), np.uint8)
    right = np.empty((height, width), np.uint8)
    gt = np.empty((height, width), np.float32) if with_gt else None
    rc = lib.lumen_synth_pair(width, height, float(dmin), float(dmax), seed & (2**64 - 1),
                              float(lowtex_frac), _ptr(left), _ptr(right),
                              _ptr(gt) if gt is not None else None)
    if rc != LUMEN_OK:
        raise StereoException("lumen_synth_pair failed")
    return (left, right, gt) if with_gt else (left, right)


class StereoSGBM:
    """GPU SGBM matcher (include/stereo/sgbm_gpu.h:29-91)."""

    def __init__(self, params: StereoParams | None = None):
        self._lib = library()
        self._h = ctypes.c_void_p()
        p = (params or StereoParams()).to_c()
        rc = self._lib.lumen_sgbm_create(ctypes.byref(p), ctypes.byref(self._h))
        if rc != LUMEN_OK:
            raise GpuException(self._lib.lumen_last_error().decode(), rc)
        self._times = np.zeros(6, np.float32)
        self._shape = (0, 0)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.lumen_sgbm_destroy(h)
        self._h = None

    @property
    def handle(self) -> ctypes.c_void_p:
        return self._h

    def _raise(self, rc: int):
        msg = self._lib.lumen_sgbm_last_error(self._h).decode()
        if rc == LUMEN_EINVAL:
            raise StereoException(msg)
        raise GpuException(msg, rc)

    def setParams(self, params: StereoParams) -> None:
        p = params.to_c()
        rc = self._lib.lumen_sgbm_set_params(self._h, ctypes.byref(p))
        if rc != LUMEN_OK:
            self._raise(rc)

    def getParams(self) -> StereoParams:
        p = CParams()
        self._lib.lumen_sgbm_get_params(self._h, ctypes.byref(p))
        return StereoParams.from_c(p)

    def compute(self, left: np.ndarray, right: np.ndarray) -> np.ndarray:
        """Disparity (int16, 1/16 px, -1 invalid) of a rectified pair: uint8 HxW grayscale or
        HxWx3 BGR (src/sgbm_gpu.cu:147-195)."""
        left = np.asarray(left)
        right = np.asarray(right)
        if left.size == 0 or right.size == 0:
            raise StereoException("Input images are empty")
        if left.shape != right.shape:
            raise StereoException("Left and right images must have same size")
        if left.dtype != np.uint8 or right.dtype != np.uint8:
            raise StereoException("Input images must be uint8")
        left = np.ascontiguousarray(left)
        right = np.ascontiguousarray(right)
        h, w = left.shape[:2]
        ch = 1 if left.ndim == 2 else left.shape[2]
        out = np.empty((h, w), np.int16)
        rc = self._lib.lumen_sgbm_compute(self._h, _ptr(left), w * ch, _ptr(right), w * ch, w, h,
                                          ch, _ptr(out), w * 2)
        if rc != LUMEN_OK:
            self._raise(rc)
        self._shape = (h, w)
        self._lib.lumen_sgbm_get_timings(self._h, self._times.ctypes.data_as(
            ctypes.POINTER(ctypes.c_float)))
        return out

    def compute_device(self, d_left: int, d_right: int, width: int, height: int, d_out: int,
                       sync: bool = True) -> None:
        """Device-pointer form: inputs and output already in HBM (see lumen_sgbm_compute_device)."""
        rc = self._lib.lumen_sgbm_compute_device(self._h, d_left, d_right, width, height, d_out,
                                                 int(sync))
        if rc != LUMEN_OK:
            self._raise(rc)
        self._shape = (height, width)

    def stream(self) -> int:
        s = ctypes.c_void_p()
        rc = self._lib.lumen_sgbm_get_stream(self._h, ctypes.byref(s))
        if rc != LUMEN_OK:
            self._raise(rc)
        return s.value or 0

    def timings(self, ncalls: int = 1) -> np.ndarray:
        """[total, cost, aggregation, post, sum+WTA+LR, upload] in ms (averaged over ncalls)."""
        t = np.zeros(6, np.float32)
        fp = t.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
        rc = (self._lib.lumen_sgbm_get_timings(self._h, fp) if ncalls == 1
              else self._lib.lumen_sgbm_get_timings_avg(self._h, ncalls, fp))
        if rc != LUMEN_OK:
            self._raise(rc)
        return t

    def setTiming(self, level: int) -> None:
        """What later computes time: 2 every stage (the default), 1 only the total
        (getLastComputeTimeMs), 0 nothing. Stage events cost host and stream time."""
        rc = self._lib.lumen_sgbm_set_timing(self._h, int(level))
        if rc != LUMEN_OK:
            self._raise(rc)

    def dump(self, stage: str, path: int = 0) -> np.ndarray:
        """One stage of the last compute in the reference's layout (parity hook)."""
        p = self.getParams()
        st = STAGE[stage]
        shape_hw = self._shape
        D = p.maxDisparity - p.minDisparity
        if st == 0:
            a = np.empty(shape_hw + (D,), np.uint8)
        elif st in (1, 2):
            a = np.empty(shape_hw + (D,), np.uint16)
        else:
            a = np.empty(shape_hw, np.int16)
        rc = self._lib.lumen_sgbm_dump_stage(self._h, st, path, _ptr(a), a.nbytes)
        if rc != LUMEN_OK:
            self._raise(rc)
        return a

    def numPaths(self) -> int:
        return self._lib.lumen_sgbm_num_paths(self._h)

    def setSchedule(self, schedule: dict | None) -> None:
        """Overrides the tuned schedule: keys agg_prefetch, agg_priority, census_tile and
        optionally agg_wide and agg_occupancy (None restores the tuned one). Results never
        depend on it."""
        if schedule is None:
            rc = self._lib.lumen_sgbm_set_schedule(self._h, None)
        else:
            c = CSchedule(int(schedule["agg_prefetch"]), int(schedule["agg_priority"]),
                          int(schedule["census_tile"]), int(schedule.get("agg_wide", 0)),
                          int(schedule.get("agg_occupancy", 0)))
            rc = self._lib.lumen_sgbm_set_schedule(self._h, ctypes.byref(c))
        if rc != LUMEN_OK:
            self._raise(rc)

    def getSchedule(self, width: int, height: int) -> dict:
        """The schedule a width x height pair would run with."""
        c = CSchedule()
        rc = self._lib.lumen_sgbm_get_schedule(self._h, width, height, ctypes.byref(c))
        if rc != LUMEN_OK:
            raise StereoException("invalid size or parameters")
        return {"agg_prefetch": c.agg_prefetch, "agg_priority": c.agg_priority,
                "census_tile": c.census_tile, "agg_wide": c.agg_wide,
                "agg_occupancy": c.agg_occupancy}

    # Timing getters of include/stereo/sgbm_gpu.h:52-55.
    def getLastComputeTimeMs(self) -> float:
        return float(self._times[0])

    def getCostComputeTimeMs(self) -> float:
        return float(self._times[1])

    def getAggregationTimeMs(self) -> float:
        return float(self._times[2])

    def getPostProcessTimeMs(self) -> float:
        return float(self._times[3])


class BatchSGBM:
    """Batched form: many pairs per stage in persistent kernels replayed as one hipGraph
    (lumen_batch_* in include/lumen_sgbm.h). Outputs are bit-identical to StereoSGBM."""

    def __init__(self, arena_bytes: int = 0):
        self._lib = library()
        self._h = ctypes.c_void_p()
        rc = self._lib.lumen_batch_create(int(arena_bytes), ctypes.byref(self._h))
        if rc != LUMEN_OK:
            raise GpuException(self._lib.lumen_batch_last_error(None).decode(), rc)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.lumen_batch_destroy(h)
        self._h = None

    def arena_bytes(self) -> int:
        v = ctypes.c_size_t()
        self._lib.lumen_batch_arena_bytes(self._h, ctypes.byref(v))
        return v.value

    def _raise(self, rc):
        msg = self._lib.lumen_batch_last_error(self._h).decode()
        if rc == LUMEN_EINVAL:
            raise StereoException(msg)
        raise GpuException(msg, rc)

    def run(self, pairs):
        """pairs: iterable of (StereoParams, left uint8 HxW, right uint8 HxW). Returns the int16
        disparity maps, in order."""
        pairs = list(pairs)
        descs = (CPairDesc * max(len(pairs), 1))()
        keep, outs = [], []
        for i, (p, l, r) in enumerate(pairs):
            l = np.ascontiguousarray(l, np.uint8)
            r = np.ascontiguousarray(r, np.uint8)
            if l.shape != r.shape or l.ndim != 2:
                raise StereoException("Left and right images must have same size")
            o = np.empty(l.shape, np.int16)
            keep += [l, r]
            outs.append(o)
            descs[i] = CPairDesc(p.to_c(), l.shape[1], l.shape[0], _ptr(l), _ptr(r), _ptr(o))
        rc = self._lib.lumen_batch_run(self._h, descs, len(pairs), 0)
        if rc != LUMEN_OK:
            self._raise(rc)
        return outs

    def run_device(self, descs, count: int):
        """descs: a ctypes array of CPairDesc whose pointers are device memory."""
        rc = self._lib.lumen_batch_run(self._h, descs, count, 1)
        if rc != LUMEN_OK:
            self._raise(rc)

    def timings(self) -> np.ndarray:
        """[total, upload, compute, download] of the last run in ms."""
        t = np.zeros(4, np.float32)
        self._lib.lumen_batch_get_timings(self._h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return t


def now_ns() -> int:
    """CLOCK_MONOTONIC in ns, the clock of the serving form's request times."""
    return int(library().lumen_now_ns())


class ServeSGBM:
    """Serving form (lumen_server_* in include/lumen_sgbm.h): requests are admitted into a
    device pool by their working-set size and run in continuous batches over several lanes.
    Results are bit-identical to StereoSGBM."""

    def __init__(self, pool_bytes: int = 0, lanes: int = 2, max_batch: int = 64,
                 device_io: bool = False, max_bypass: int = 8, batch_budget_us: int = 2000,
                 pinned_mb: int = 0, large_us: int = 0):
        self._lib = library()
        self._h = ctypes.c_void_p()
        self._device_io = bool(device_io)
        cfg = CServerConfig(int(pool_bytes), int(lanes), int(max_batch), int(device_io),
                            int(max_bypass), int(batch_budget_us), int(pinned_mb), int(large_us),
                            0)
        rc = self._lib.lumen_server_create(ctypes.byref(cfg), ctypes.byref(self._h))
        if rc != LUMEN_OK:
            raise GpuException(self._lib.lumen_server_last_error(None).decode(), rc)
        # ticket -> (left, right, out) of host requests: the server reads and writes these
        # buffers until a wait returns the ticket's status, so they live here until then.
        self._live = {}
        self._waiting = set()
        self._lock = threading.Lock()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.lumen_server_destroy(h)
        self._h = None

    def close(self) -> None:
        """Finishes every submitted request and frees the server."""
        self.__del__()

    def _raise(self, rc):
        msg = self._lib.lumen_server_last_error(self._h).decode()
        if rc == LUMEN_EINVAL:
            raise StereoException(msg)
        raise GpuException(msg, rc)

    def submit(self, params: StereoParams, left: np.ndarray, right: np.ndarray) -> int:
        """Queues one host-memory pair; returns its ticket. Never waits for the device."""
        left = np.ascontiguousarray(left, np.uint8)
        right = np.ascontiguousarray(right, np.uint8)
        if left.shape != right.shape or left.ndim != 2:
            raise StereoException("Left and right images must have same size")
        out = np.empty(left.shape, np.int16)
        d = CPairDesc(params.to_c(), left.shape[1], left.shape[0], _ptr(left), _ptr(right),
                      _ptr(out))
        t = ctypes.c_uint64()
        rc = self._lib.lumen_server_submit(self._h, ctypes.byref(d), ctypes.byref(t))
        if rc != LUMEN_OK:
            self._raise(rc)
        self._live[t.value] = (left, right, out)
        return t.value

    def submit_desc(self, desc: CPairDesc) -> int:
        """Queues a pair given as a CPairDesc (device pointers when device_io is set)."""
        t = ctypes.c_uint64()
        rc = self._lib.lumen_server_submit(self._h, ctypes.byref(desc), ctypes.byref(t))
        if rc != LUMEN_OK:
            self._raise(rc)
        return t.value

    def wait(self, ticket: int, timeout_s: float | None = None):
        """Waits for a ticket; returns (disparity or None for device requests, info dict).

        Raises TimeoutError when the request is not done within timeout_s: the ticket stays
        valid (the request may still be queued or running, so its buffers are kept) and may be
        waited on again. A ticket takes one waiter at a time."""
        with self._lock:
            if ticket in self._waiting:
                raise StereoException("ticket already has a waiter")
            self._waiting.add(ticket)
        info = CRequestInfo()
        tmo = -1 if timeout_s is None else int(timeout_s * 1e6)
        try:
            rc = self._lib.lumen_server_wait(self._h, ticket, tmo, ctypes.byref(info))
        finally:
            with self._lock:
                self._waiting.discard(ticket)
        if rc == LUMEN_ETIMEDOUT:
            raise TimeoutError(f"ticket {ticket} not done after {timeout_s} s")
        # Any other status means the server released the ticket: its buffers are free again.
        held = self._live.pop(ticket, None)
        if rc != LUMEN_OK:
            self._raise(rc)
        d = {k: getattr(info, k) for k, _ in CRequestInfo._fields_ if k != "reserved"}
        return (held[2] if held is not None else None), d

    def drain(self) -> None:
        self._lib.lumen_server_drain(self._h)

    def stats(self) -> dict:
        st = CServerStats()
        self._lib.lumen_server_get_stats(self._h, ctypes.byref(st))
        return {k: int(getattr(st, k)) for k, _ in CServerStats._fields_}


def tuning_load(path: str | None = None) -> int:
    """Replaces the process-wide tuning table (None: tuning.tsv next to the library). Returns
    the number of entries; raises StereoException for a malformed file."""
    n = ctypes.c_int()
    lib = library()
    rc = lib.lumen_tuning_load(str(path).encode() if path else None, ctypes.byref(n))
    if rc != LUMEN_OK:
        raise StereoException(lib.lumen_last_error().decode())
    return n.value


def tuning_clear() -> None:
    """Empties the tuning table: every pair runs the default schedule."""
    library().lumen_tuning_clear()
