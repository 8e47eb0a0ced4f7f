"""ctypes binding of the CPU oracle (oracle/sgbm_oracle.cpp).

TEST INFRASTRUCTURE, NOT THE PRODUCT: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module, and only as the checker (or as the timed CPU baseline).
The oracle restates the reference kernels of Pfactorial01/LumenStereo (src/sgbm_kernels.cu,
src/sgbm_gpu.cu; src/rectification.cu and src/depth_map.cu for the stages either side of the
matcher) single-threaded; see the headers of sgbm_oracle.h and scene_oracle.h for what pins
each function.
Parity for the speckle stage is unpinned (the reference kernel races).
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboracle.so"
PATH_NAMES = ["LR", "RL", "TB", "BT", "TLBR", "BRTL", "TRBL", "BLTR"]

_lib = None


class OracleParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "minDisparity", "maxDisparity", "blockSize", "P1", "P2", "preFilterCap",
        "uniquenessRatio", "disp12MaxDiff", "speckleWindowSize", "speckleRange", "useGPU",
        "useSharedMemory", "numDirections", "matchingCostMode")]


class OracleDumps(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in (
        "cost", "paths", "sum", "wta_int", "disp_left", "disp_right", "after_lr",
        "after_speckle")]


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing; run __graft_entry__.build()")
        l = ctypes.CDLL(str(LIB_PATH))
        vp, ci = ctypes.c_void_p, ctypes.c_int
        sigs = {
            "oracle_prefilter": [vp, ci, ci, ci, vp],
            "oracle_census": [vp, ci, ci, vp],
            "oracle_cost_sad": [vp, vp, ci, ci, ci, ci, ci, vp],
            "oracle_cost_census": [vp, vp, ci, ci, ci, ci, vp],
            "oracle_cost_scale": [ci, ci, ci],
            "oracle_num_paths": [ci],
            "oracle_aggregate_path": [vp, vp, ci, ci, ci, ci, ci, ci, ci, vp],
            "oracle_wta_left": [vp, ci, ci, ci, ci, ci, vp, vp],
            "oracle_wta_right": [vp, ci, ci, ci, ci, vp],
            "oracle_lr_check": [vp, vp, ci, ci, ci],
            "oracle_speckle": [vp, ci, ci, ci, ci],
            "oracle_median": [vp, vp, ci, ci],
            "oracle_sgbm": [ctypes.POINTER(OracleParams), vp, vp, ci, ci, vp,
                            ctypes.POINTER(OracleDumps)],
            "oracle_sgbm_mt": [ctypes.POINTER(OracleParams), vp, vp, ci, ci, vp, ci,
                               ctypes.POINTER(OracleDumps)],
            # scene_oracle.h
            "oracle_remap": [vp, vp, vp, vp, ci, ci, ci],
            "oracle_disparity_to_depth": [vp, vp, ci, ci, ctypes.c_float, ctypes.c_float,
                                          ctypes.c_float],
            "oracle_disparity_to_depth_host": [vp, vp, ci, ci, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_double],
            "oracle_colorize_depth": [vp, vp, ci, ci, ctypes.c_float, ctypes.c_float],
            "oracle_turbo_entry": [ci],
        }
        void_fns = ("oracle_prefilter", "oracle_census", "oracle_cost_sad", "oracle_cost_census",
                    "oracle_aggregate_path", "oracle_wta_left", "oracle_wta_right",
                    "oracle_lr_check", "oracle_speckle", "oracle_median", "oracle_remap",
                    "oracle_disparity_to_depth", "oracle_disparity_to_depth_host",
                    "oracle_colorize_depth")
        for name, args in sigs.items():
            f = getattr(l, name)
            f.argtypes = args
            f.restype = (None if name in void_fns else
                         ctypes.c_uint32 if name == "oracle_turbo_entry" else ci)
        _lib = l
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def params_struct(p) -> OracleParams:
    """Accepts a lumenstereo_amd.StereoParams-like object or a dict."""
    get = (lambda k: p[k]) if isinstance(p, dict) else (lambda k: getattr(p, k))
    return OracleParams(*[int(get(n)) for n, _ in OracleParams._fields_])


def prefilter(img, cap):
    img = _c(img, np.uint8)
    out = np.empty_like(img)
    lib().oracle_prefilter(_p(img), img.shape[1], img.shape[0], cap, _p(out))
    return out


def census(img):
    img = _c(img, np.uint8)
    out = np.empty(img.shape, np.uint32)
    lib().oracle_census(_p(img), img.shape[1], img.shape[0], _p(out))
    return out


def cost_sad(fl, fr, D, m, block_size):
    fl, fr = _c(fl, np.uint8), _c(fr, np.uint8)
    H, W = fl.shape
    out = np.empty((H, W, D), np.uint8)
    lib().oracle_cost_sad(_p(fl), _p(fr), W, H, D, m, block_size, _p(out))
    return out


def cost_census(cl, cr, D, m):
    cl, cr = _c(cl, np.uint32), _c(cr, np.uint32)
    H, W = cl.shape
    out = np.empty((H, W, D), np.uint8)
    lib().oracle_cost_census(_p(cl), _p(cr), W, H, D, m, _p(out))
    return out


def cost_scale(mode, block_size, cap):
    return lib().oracle_cost_scale(int(mode), block_size, cap)


def aggregate_path(cost, img, P1, P2, cs, path):
    cost, img = _c(cost, np.uint8), _c(img, np.uint8)
    H, W, D = cost.shape
    out = np.empty((H, W, D), np.uint16)
    lib().oracle_aggregate_path(_p(cost), _p(img), W, H, D, P1, P2, cs, path, _p(out))
    return out


def wta_left(S, m, uniq):
    S = _c(S, np.uint16)
    H, W, D = S.shape
    disp = np.empty((H, W), np.int16)
    wi = np.empty((H, W), np.int16)
    lib().oracle_wta_left(_p(S), W, H, D, m, uniq, _p(disp), _p(wi))
    return disp, wi


def wta_right(S, m):
    S = _c(S, np.uint16)
    H, W, D = S.shape
    disp = np.empty((H, W), np.int16)
    lib().oracle_wta_right(_p(S), W, H, D, m, _p(disp))
    return disp


def lr_check(dl, dr, max_diff16):
    dl = _c(dl, np.int16).copy()
    dr = _c(dr, np.int16)
    lib().oracle_lr_check(_p(dl), _p(dr), dl.shape[1], dl.shape[0], max_diff16)
    return dl


def speckle(d, min_similar, max_diff16):
    d = _c(d, np.int16).copy()
    lib().oracle_speckle(_p(d), d.shape[1], d.shape[0], min_similar, max_diff16)
    return d


def median(d):
    d = _c(d, np.int16)
    out = np.empty_like(d)
    lib().oracle_median(_p(d), _p(out), d.shape[1], d.shape[0])
    return out


def sgbm(params, left, right, dumps: bool = False, threads: int = 1):
    """Full matcher. Returns the output, or (output, dict of stage arrays) when dumps=True.
    threads > 1 runs the threaded form (identical results, for full-size frames)."""
    left, right = _c(left, np.uint8), _c(right, np.uint8)
    H, W = left.shape
    ps = params_struct(params)
    D = ps.maxDisparity - ps.minDisparity
    out = np.empty((H, W), np.int16)
    d = None
    st = {}
    if dumps:
        npaths = lib().oracle_num_paths(ps.numDirections)
        st = {
            "cost": np.empty((H, W, D), np.uint8),
            "paths": np.empty((npaths, H, W, D), np.uint16),
            "sum": np.empty((H, W, D), np.uint16),
            "wta_int": np.empty((H, W), np.int16),
            "disp_left": np.empty((H, W), np.int16),
            "disp_right": np.empty((H, W), np.int16),
            "after_lr": np.empty((H, W), np.int16),
            "after_speckle": np.empty((H, W), np.int16),
        }
        d = OracleDumps(*[_p(st[n]) for n, _ in OracleDumps._fields_])
    dref = ctypes.byref(d) if d is not None else None
    if threads > 1:
        rc = lib().oracle_sgbm_mt(ctypes.byref(ps), _p(left), _p(right), W, H, _p(out),
                                  int(threads), dref)
    else:
        rc = lib().oracle_sgbm(ctypes.byref(ps), _p(left), _p(right), W, H, _p(out), dref)
    if rc != 0:
        raise ValueError("oracle_sgbm rejected the arguments")
    return (out, st) if dumps else out


def sgbm_mt(params, left, right, nthreads):
    """The threaded form without dumps (the reported CPU baseline)."""
    return sgbm(params, left, right, dumps=False, threads=int(nthreads))


def remap(src, mapx, mapy):
    """src/rectification.cu:201-256: src H x W (or H x W x C) uint8, float32 maps H x W."""
    src = _c(src, np.uint8)
    H, W = src.shape[:2]
    C = 1 if src.ndim == 2 else src.shape[2]
    mapx, mapy = _c(mapx, np.float32), _c(mapy, np.float32)
    assert mapx.shape == (H, W) and mapy.shape == (H, W)
    out = np.empty_like(src)
    lib().oracle_remap(_p(src), _p(mapx), _p(mapy), _p(out), W, H, C)
    return out


def disparity_to_depth(disp, f, B, doffs=0.0):
    """src/depth_map.cu:74-101 (the GPU kernel's float arithmetic)."""
    disp = _c(disp, np.int16)
    out = np.empty(disp.shape, np.float32)
    lib().oracle_disparity_to_depth(_p(disp), _p(out), disp.shape[1], disp.shape[0], f, B, doffs)
    return out


def disparity_to_depth_host(disp, f, B, doffs=0.0):
    """src/depth_map.cu:142-160 (the host function's double quotient)."""
    disp = _c(disp, np.int16)
    out = np.empty(disp.shape, np.float32)
    lib().oracle_disparity_to_depth_host(_p(disp), _p(out), disp.shape[1], disp.shape[0], f, B,
                                         doffs)
    return out


def colorize_depth(depth, min_depth, max_depth):
    """src/depth_map.cu:106-138: H x W float32 -> H x W x 3 uint8 (R, G, B)."""
    depth = _c(depth, np.float32)
    out = np.empty(depth.shape + (3,), np.uint8)
    lib().oracle_colorize_depth(_p(depth), _p(out), depth.shape[1], depth.shape[0], min_depth,
                                max_depth)
    return out


def turbo_table():
    """The oracle's colour table as a 256 x 3 uint8 array (R, G, B)."""
    v = np.array([lib().oracle_turbo_entry(i) for i in range(256)], np.uint32)
    return np.stack([v & 0xFF, (v >> 8) & 0xFF, (v >> 16) & 0xFF], axis=1).astype(np.uint8)
