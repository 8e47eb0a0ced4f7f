"""Host-side checks of the C ABI and its Python mirror (no GPU work).

Covers: the library loads and exports every function include/lumen_*.h declares; defaults
and helpers mirror the reference's StereoParams (tests/test_sgbm.cpp:28-75); validation errors
are raised before any device call with the reference's messages (src/sgbm_gpu.cu:148-166);
footprint arithmetic; the synthetic scene generator is deterministic.
"""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import lumenstereo_amd as ls

HEADERS = sorted((Path(__file__).resolve().parent.parent / "include").glob("lumen_*.h"))


def declared_functions():
    text = "".join(h.read_text() for h in HEADERS)
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(lumen_\w+)\s*\(", text, re.M)))


def test_header_symbols_exported():
    names = declared_functions()
    assert len(HEADERS) >= 3 and len(names) >= 45
    lib = ctypes.CDLL(str(ls.LIB_PATH))
    for n in names:
        assert hasattr(lib, n), f"{n} declared in include/ but not exported"
    # and every declared symbol has a Python signature
    assert set(names) <= set(ls._SIGNATURES), set(names) - set(ls._SIGNATURES)


def test_defaults_match_reference():
    p = ls.StereoParams()
    assert (p.minDisparity, p.maxDisparity, p.blockSize) == (0, 128, 5)
    assert (p.P1, p.P2, p.preFilterCap) == (200, 800, 31)
    assert (p.uniquenessRatio, p.disp12MaxDiff) == (10, 1)
    assert (p.speckleWindowSize, p.speckleRange, p.numDirections) == (100, 32, 4)
    assert p.matchingCostMode == ls.MatchingCostMode.SAD
    assert p.isValid()
    c = ls.CParams()
    assert ls.library().lumen_sgbm_default_params(ctypes.byref(c)) == 0
    assert ls.StereoParams.from_c(c) == p


def test_is_valid_and_penalties():
    p = ls.StereoParams(blockSize=4)
    assert not p.isValid()
    p = ls.StereoParams(minDisparity=100, maxDisparity=50)
    assert not p.isValid()
    p = ls.StereoParams(minDisparity=-10)
    assert not p.isValid()
    p = ls.StereoParams(P1=100, P2=50)
    assert not p.isValid()
    p = ls.StereoParams()
    p.autoComputePenalties(1)
    assert (p.P1, p.P2) == (200, 800)
    assert ls.middleburyMatcherMaxDisparity(280) == 280
    assert ls.middleburyMatcherMaxDisparity(1000) == 512
    assert ls.middleburyMatcherMaxDisparity(0) == 96


def test_params_roundtrip_without_gpu():
    m = ls.StereoSGBM(ls.StereoParams(maxDisparity=64, numDirections=8))
    assert m.getParams().maxDisparity == 64
    assert m.numPaths() == 8
    m.setParams(ls.StereoParams(numDirections=2))
    assert m.numPaths() == 4
    assert (m.timings() == 0).all()


@pytest.mark.parametrize("params,shape,msg", [
    (dict(), (0, 0), "Input images are empty"),
    (dict(minDisparity=5, maxDisparity=5), (8, 8),
     "Invalid disparity range: maxDisparity must be greater than minDisparity"),
    (dict(maxDisparity=544), (8, 8),
     "Disparity range (maxDisparity - minDisparity) exceeds GPU limit (512). Reduce the search range."),
    (dict(minDisparity=-4, maxDisparity=60), (8, 8), "minDisparity must be >= 0"),
    (dict(maxDisparity=16), (4, 4601), "Image width exceeds 4600"),
    (dict(maxDisparity=16), (3, 1), "Image width must be at least 2 for the SAD prefilter"),
    # Ranges whose difference does not fit 32 bits are still classified by their true value.
    (dict(minDisparity=-2**31, maxDisparity=2**31 - 1), (8, 8),
     "Disparity range (maxDisparity - minDisparity) exceeds GPU limit (512). Reduce the search range."),
    (dict(minDisparity=10, maxDisparity=-2**31 + 5), (8, 8),
     "Invalid disparity range: maxDisparity must be greater than minDisparity"),
])
def test_validation_messages(params, shape, msg):
    m = ls.StereoSGBM(ls.StereoParams(**params))
    img = np.zeros(shape, np.uint8)
    with pytest.raises(ls.StereoException) as e:
        m.compute(img, img)
    assert str(e.value) == msg


def test_size_mismatch():
    m = ls.StereoSGBM()
    with pytest.raises(ls.StereoException, match="same size"):
        m.compute(np.zeros((100, 100), np.uint8), np.zeros((100, 200), np.uint8))


def test_dump_before_compute_is_an_error():
    m = ls.StereoSGBM()
    m._shape = (4, 4)
    with pytest.raises(ls.StereoException):
        m.dump("output")


def test_footprint():
    lib = ls.library()
    out = ctypes.c_size_t()
    p = ls.StereoParams(maxDisparity=128, numDirections=8,
                        matchingCostMode=ls.MatchingCostMode.Census).to_c()
    assert lib.lumen_sgbm_footprint(ctypes.byref(p), 1920, 1080, ctypes.byref(out)) == 0
    n = 1920 * 1080
    assert out.value == n * 128 * (1 + 2 * 8) + n * 2 + n * 12
    p = ls.StereoParams(maxDisparity=290, numDirections=4).to_c()   # padded to 320
    assert lib.lumen_sgbm_footprint(ctypes.byref(p), 100, 10, ctypes.byref(out)) == 0
    assert out.value == 1000 * 320 * 9 + 1000 * 14


def test_footprint_and_schedule_reject_extreme_ranges():
    lib = ls.library()
    out = ctypes.c_size_t()
    for lo, hi in [(-2**31, 2**31 - 1), (10, -2**31 + 5), (0, 0)]:
        p = ls.StereoParams(minDisparity=lo, maxDisparity=hi).to_c()
        assert lib.lumen_sgbm_footprint(ctypes.byref(p), 64, 64, ctypes.byref(out)) == ls.LUMEN_EINVAL
    m = ls.StereoSGBM(ls.StereoParams(minDisparity=-2**31, maxDisparity=2**31 - 1))
    sched = ls.CSchedule()
    assert lib.lumen_sgbm_get_schedule(m._h, 64, 64, ctypes.byref(sched)) == ls.LUMEN_EINVAL


def test_synth_deterministic_and_shaped():
    a = ls.synth_pair(120, 80, 4, 60, seed=99, with_gt=True)
    b = ls.synth_pair(120, 80, 4, 60, seed=99, with_gt=True)
    c = ls.synth_pair(120, 80, 4, 60, seed=100)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert not np.array_equal(a[0], c[0])
    left, right, gt = a
    assert abs(float(left.mean()) - 128) < 6 and abs(float(left.std()) - 40) < 6
    assert gt.min() >= 4 and gt.max() <= 60
    low = ls.synth_pair(200, 100, 10, 280, seed=5, lowtex_frac=0.2)[0]
    assert low.std() < 40


def test_device_memory_argument_checks():
    """lumen_device_* (the DeviceBuffer ABI) refuse bad arguments before any device call, and
    treat empty requests as no-ops, as the reference's CudaBuffer does for size 0."""
    lib = ls.library()
    p = ctypes.c_void_p(1234)
    assert lib.lumen_device_alloc(0, ctypes.byref(p)) == ls.LUMEN_OK and p.value is None
    assert lib.lumen_device_alloc(16, None) == ls.LUMEN_EINVAL
    assert lib.lumen_device_free(None) == ls.LUMEN_OK
    assert lib.lumen_device_copy(None, None, 0, 0) == ls.LUMEN_OK
    assert lib.lumen_device_copy(None, None, 8, 0) == ls.LUMEN_EINVAL
    buf = (ctypes.c_uint8 * 8)()
    assert lib.lumen_device_copy(buf, buf, 8, 7) == ls.LUMEN_EINVAL
    assert b"direction" in lib.lumen_last_error()
    assert lib.lumen_device_zero(None, 0) == ls.LUMEN_OK
    assert lib.lumen_device_zero(None, 4) == ls.LUMEN_EINVAL
