"""The stages either side of the matcher: rectification before it, depth and the depth colour
map after it. Python mirror of the reference's include/stereo/rectification.h (:25-111) and
include/stereo/depth_map.h (:14-58), bound to liblumen_sgbm.so (include/lumen_rectify.h,
include/lumen_depth.h). There is no CPU fallback.

* ``CameraParams`` mirrors the calibration record of include/stereo/stereo_params.h:135-186
  with numpy arrays in place of cv::Mat.
* ``StereoRectifier`` mirrors the reference class (src/rectification.cu:262-352);
  ``rectifyImageGPU`` the one-image function (:363-420).
* ``disparityToDepth`` mirrors the host function (src/depth_map.cu:142-160) and
  ``disparityToDepthGPU`` / ``colorizeDepthGPU`` the device functions (:177-221), whose buffers
  are device pointers (ints, e.g. ``tensor.data_ptr()``).
* ``colorizeDepth`` gives the colours of the reference's GPU colour map. The reference's host
  colorizeDepth calls OpenCV's COLORMAP_TURBO, which is not available here (an API delta).
"""
from __future__ import annotations

import ctypes
import dataclasses

import numpy as np

from . import GpuException, LUMEN_EINVAL, LUMEN_OK, StereoException, library

__all__ = [
    "CameraParams",
    "StereoRectifier",
    "rectifyImageGPU",
    "remapDevice",
    "disparityToDepth",
    "disparityToDepthGPU",
    "colorizeDepth",
    "colorizeDepthGPU",
    "depthColorizeGPU",
]


def _raise_global(rc: int):
    msg = library().lumen_last_error().decode()
    if rc == LUMEN_EINVAL:
        raise StereoException(msg)
    raise GpuException(msg, rc)


def _f32(a, shape) -> np.ndarray:
    a = np.ascontiguousarray(a, np.float32)
    if a.shape != shape:
        raise StereoException(f"map shape {a.shape} does not match the image {shape}")
    return a


@dataclasses.dataclass
class CameraParams:
    """include/stereo/stereo_params.h:135-186: image size, intrinsics K0/K1, distortion D0/D1,
    extrinsics R/T, baseline and focal length, Middlebury disparity offset and count, and the
    precomputed rectification maps (float32 H x W each; None when pre-rectified)."""

    width: int = 0
    height: int = 0
    K0: np.ndarray | None = None
    D0: np.ndarray | None = None
    K1: np.ndarray | None = None
    D1: np.ndarray | None = None
    R: np.ndarray | None = None
    T: np.ndarray | None = None
    baseline: float = 0.0
    focalLength: float = 0.0
    disparityOffset: int = 0
    numDisparities: int = 0
    map0x: np.ndarray | None = None
    map0y: np.ndarray | None = None
    map1x: np.ndarray | None = None
    map1y: np.ndarray | None = None

    def computeDerivedValues(self) -> None:
        """include/stereo/stereo_params.h:166-173: baseline = |T|, focal = (fx + fy) / 2."""
        if self.T is not None and np.size(self.T):
            self.baseline = float(np.linalg.norm(np.asarray(self.T, np.float64)))
        if self.K0 is not None and np.size(self.K0):
            k = np.asarray(self.K0, np.float64)
            self.focalLength = float((k[0, 0] + k[1, 1]) / 2.0)

    def isValid(self) -> bool:
        """include/stereo/stereo_params.h:178-181."""
        return all(m is not None and np.size(m) for m in (self.K0, self.K1, self.R, self.T)) \
            and self.width > 0 and self.height > 0


class StereoRectifier:
    """GPU stereo rectifier (include/stereo/rectification.h:40-88). The maps are uploaded once
    at construction; rectify() remaps both images in one launch."""

    def __init__(self, params: CameraParams):
        self._lib = library()
        self._params = params
        self._h = ctypes.c_void_p()
        pre = params.map0x is None or params.map0y is None or not np.size(params.map0x) \
            or not np.size(params.map0y)
        shape = (params.height, params.width)
        keep = [None] * 4 if pre else [_f32(m, shape) for m in (params.map0x, params.map0y,
                                                                 params.map1x, params.map1y)]
        ptrs = [m.ctypes.data if m is not None else None for m in keep]
        rc = self._lib.lumen_rectifier_create(*ptrs, int(params.width), int(params.height),
                                              ctypes.byref(self._h))
        if rc != LUMEN_OK:
            _raise_global(rc)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            self._lib.lumen_rectifier_destroy(h)
        self._h = None

    def _raise(self, rc):
        msg = self._lib.lumen_rectifier_last_error(self._h).decode()
        if rc == LUMEN_EINVAL:
            raise StereoException(msg)
        raise GpuException(msg, rc)

    def isPreRectified(self) -> bool:
        return self._lib.lumen_rectifier_is_pre_rectified(self._h) == 1

    def getParams(self) -> CameraParams:
        return self._params

    def rectify(self, leftRaw: np.ndarray, rightRaw: np.ndarray):
        """src/rectification.cu:286-352: returns (leftRect, rightRect), same shape and type.
        A pre-rectified rectifier returns copies of its inputs."""
        l = np.ascontiguousarray(leftRaw, np.uint8)
        r = np.ascontiguousarray(rightRaw, np.uint8)
        if l.size == 0 or r.size == 0:
            raise StereoException("Input images are empty")
        if l.shape != r.shape:
            raise StereoException("Left and right images must have same size")
        if self.isPreRectified():
            return l.copy(), r.copy()
        h, w = l.shape[:2]
        ch = 1 if l.ndim == 2 else l.shape[2]
        lo, ro = np.empty_like(l), np.empty_like(r)
        rc = self._lib.lumen_rectifier_rectify(self._h, l.ctypes.data, r.ctypes.data, w, h, ch,
                                               lo.ctypes.data, ro.ctypes.data)
        if rc != LUMEN_OK:
            self._raise(rc)
        return lo, ro

    def rectify_device(self, d_left: int, d_right: int, width: int, height: int, channels: int,
                       d_left_out: int, d_right_out: int, sync: bool = True) -> None:
        """Device-pointer form (lumen_rectifier_rectify_device)."""
        rc = self._lib.lumen_rectifier_rectify_device(self._h, d_left, d_right, width, height,
                                                      channels, d_left_out, d_right_out, int(sync))
        if rc != LUMEN_OK:
            self._raise(rc)

    def stream(self) -> int:
        s = ctypes.c_void_p()
        rc = self._lib.lumen_rectifier_get_stream(self._h, ctypes.byref(s))
        if rc != LUMEN_OK:
            self._raise(rc)
        return s.value or 0

    def last_time_ms(self) -> float:
        t = ctypes.c_float()
        rc = self._lib.lumen_rectifier_get_time_ms(self._h, ctypes.byref(t))
        if rc != LUMEN_OK:
            self._raise(rc)
        return float(t.value)


def rectifyImageGPU(src: np.ndarray, mapX: np.ndarray, mapY: np.ndarray) -> np.ndarray:
    """src/rectification.cu:363-420: one image (H x W or H x W x C uint8) remapped on the
    device through float32 maps of H x W."""
    s = np.ascontiguousarray(src, np.uint8)
    if s.size == 0:
        raise StereoException("Input images are empty")
    h, w = s.shape[:2]
    ch = 1 if s.ndim == 2 else s.shape[2]
    mx, my = _f32(mapX, (h, w)), _f32(mapY, (h, w))
    dst = np.empty_like(s)
    rc = library().lumen_remap(s.ctypes.data, mx.ctypes.data, my.ctypes.data, dst.ctypes.data,
                               w, h, ch)
    if rc != LUMEN_OK:
        _raise_global(rc)
    return dst


def remapDevice(d_src: int, d_mapx: int, d_mapy: int, d_dst: int, width: int, height: int,
                channels: int = 1, stream: int = 0) -> None:
    """Device-pointer remap enqueued on `stream` (lumen_remap_device)."""
    rc = library().lumen_remap_device(d_src, d_mapx, d_mapy, d_dst, width, height, channels,
                                      stream or None)
    if rc != LUMEN_OK:
        _raise_global(rc)


def disparityToDepth(disparity: np.ndarray, focalLength: float, baseline: float,
                     disparityOffsetPx: float = 0.0) -> np.ndarray:
    """src/depth_map.cu:142-160: int16 disparity (1/16 px) -> float32 depth in metres, with the
    host function's arithmetic (double quotient), computed on the device."""
    d = np.ascontiguousarray(disparity, np.int16)
    if d.ndim != 2 or d.size == 0:
        raise StereoException("Disparity map is empty")
    out = np.empty(d.shape, np.float32)
    rc = library().lumen_disparity_to_depth(d.ctypes.data, out.ctypes.data, d.shape[1], d.shape[0],
                                            float(focalLength), float(baseline),
                                            float(disparityOffsetPx))
    if rc != LUMEN_OK:
        _raise_global(rc)
    return out


def disparityToDepthGPU(d_disparity: int, d_depth: int, width: int, height: int,
                        focalLength: float, baseline: float, disparityOffsetPx: float = 0.0,
                        stream: int = 0) -> None:
    """src/depth_map.cu:177-197 on device pointers, enqueued on `stream`."""
    rc = library().lumen_disparity_to_depth_device(d_disparity, d_depth, width, height,
                                                   focalLength, baseline, disparityOffsetPx,
                                                   stream or None)
    if rc != LUMEN_OK:
        _raise_global(rc)


def colorizeDepth(depth: np.ndarray, minDepth: float = 0.5, maxDepth: float = 10.0) -> np.ndarray:
    """Depth (float32 H x W) -> H x W x 3 uint8 (R, G, B) through the reference's GPU colour
    table (src/depth_map.cu:106-138)."""
    z = np.ascontiguousarray(depth, np.float32)
    if z.ndim != 2 or z.size == 0:
        raise StereoException("Depth map is empty")
    out = np.empty(z.shape + (3,), np.uint8)
    rc = library().lumen_colorize_depth(z.ctypes.data, out.ctypes.data, z.shape[1], z.shape[0],
                                        float(minDepth), float(maxDepth))
    if rc != LUMEN_OK:
        _raise_global(rc)
    return out


def colorizeDepthGPU(d_depth: int, d_rgb: int, width: int, height: int, minDepth: float,
                     maxDepth: float, stream: int = 0) -> None:
    """src/depth_map.cu:211-221 on device pointers, enqueued on `stream`."""
    rc = library().lumen_colorize_depth_device(d_depth, d_rgb, width, height, minDepth, maxDepth,
                                               stream or None)
    if rc != LUMEN_OK:
        _raise_global(rc)


def depthColorizeGPU(d_disparity: int, d_depth: int | None, d_rgb: int, width: int, height: int,
                     focalLength: float, baseline: float, disparityOffsetPx: float,
                     minDepth: float, maxDepth: float, stream: int = 0) -> None:
    """Both stages in one pass (lumen_depth_colorize_device); d_depth may be None."""
    rc = library().lumen_depth_colorize_device(d_disparity, d_depth or None, d_rgb, width, height,
                                               focalLength, baseline, disparityOffsetPx,
                                               minDepth, maxDepth, stream or None)
    if rc != LUMEN_OK:
        _raise_global(rc)
