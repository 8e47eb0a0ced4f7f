"""Stage-by-stage parity of the HIP path (through the C ABI) against the CPU oracle.

Every stage is compared bit-exactly: cost volume, each path's aggregated costs, their sum,
integer WTA, left/right subpixel maps, after the LR check, after speckle, final output.
"""
import numpy as np
import pytest

import lumenstereo_amd as ls
import oracle

pytestmark = pytest.mark.gpu

CASES = [
    # (W, H, minD, maxD, mode, paths, extra params)
    (96, 64, 0, 32, 0, 4, {}),
    (96, 64, 0, 32, 1, 8, {}),
    (130, 77, 0, 40, 0, 8, {}),          # padded D (40 -> 64), odd sizes
    (130, 77, 3, 43, 1, 4, {}),          # minDisparity > 0
    (160, 90, 0, 64, 1, 8, {"uniquenessRatio": 10}),
    (160, 90, 0, 64, 0, 4, {"disp12MaxDiff": -1}),
    (120, 80, 0, 96, 0, 4, {"blockSize": 3}),        # D = 96: one run of six per lane
    (130, 70, 5, 85, 1, 8, {"uniquenessRatio": 10}),  # D = 80 padded to 96, minDisparity > 0
    (4600, 5, 0, 90, 1, 8, {}),           # widest image with the 96 range's right-WTA table
    (120, 80, 0, 48, 0, 4, {"blockSize": 7, "preFilterCap": 15}),
    (100, 70, 0, 32, 1, 4, {"P1": 30000, "P2": 60000}),   # forces uint16 wrap
    (100, 70, 0, 32, 0, 8, {"speckleWindowSize": 20, "speckleRange": 2}),
    (200, 40, 0, 130, 1, 8, {}),          # D > W/2, padded to 192
    (64, 200, 0, 16, 0, 8, {}),           # tall image, short diagonals
    (40, 520, 0, 16, 1, 8, {}),           # 520 rows: the WTA's 256-thread row blocks
    (90, 60, 0, 24, 0, 4, {"blockSize": 9}),     # SAD window with the generic radius
    (90, 60, 4, 36, 0, 8, {"blockSize": 1}),     # 1x1 SAD window (divisor 1)
    (70, 50, 0, 16, 0, 4, {"blockSize": 201}),   # SAD window too large for the LDS segment
    (150, 64, 0, 200, 1, 4, {}),          # 16 disparities per lane
    (120, 40, 5, 295, 0, 4, {}),          # D = 290: 20 per lane, padded to 320
    (100, 30, 0, 420, 1, 4, {}),          # D = 420: 28 per lane (7 chunks of 4), padded to 448
    (96, 24, 0, 512, 0, 8, {}),           # D = 512, the largest range: 32 per lane
    (4600, 5, 0, 320, 1, 4, {}),          # the widest image: right-WTA row state fills LDS
    (4600, 5, 7, 263, 0, 8, {}),          # widest image, 16 per lane, minDisparity > 0
    # Prefilter caps of 128 and more: g + cap exceeds 255 and wraps as the uint8 store does.
    (96, 64, 0, 48, 0, 4, {"preFilterCap": 200, "blockSize": 3}),
    (96, 64, 0, 48, 0, 8, {"preFilterCap": 200}),
    (96, 64, 0, 48, 0, 4, {"preFilterCap": 1000, "blockSize": 7}),
    (96, 64, 0, 48, 0, 4, {"preFilterCap": 1000, "blockSize": 5}),
    (96, 64, 0, 48, 0, 4, {"preFilterCap": 200, "blockSize": 9}),
    (70, 50, 0, 16, 0, 4, {"blockSize": -3}),    # empty window: every cost 255 (fill kernel)
    # Wide enough for SAD segments away from the image borders (staged by column groups).
    (700, 40, 0, 64, 0, 4, {"blockSize": 3}),
    (700, 40, 0, 64, 0, 8, {"blockSize": 7, "preFilterCap": 200}),
]


def _params(minD, maxD, mode, paths, extra):
    p = ls.StereoParams(minDisparity=minD, maxDisparity=maxD, numDirections=paths,
                        matchingCostMode=ls.MatchingCostMode(mode), uniquenessRatio=0,
                        speckleWindowSize=0)
    for k, v in extra.items():
        setattr(p, k, v)
    return p


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}_d{c[2]}-{c[3]}_m{c[4]}_p{c[5]}_{len(c[6])}")
def test_stage_parity(gpu_available, case):
    W, H, minD, maxD, mode, paths, extra = case
    left, right = ls.synth_pair(W, H, max(minD, 2), maxD - 4, seed=W * 1000 + H)
    p = _params(minD, maxD, mode, paths, extra)
    want, st = oracle.sgbm(p, left, right, dumps=True)
    m = ls.StereoSGBM(p)
    got = m.compute(left, right)
    np.testing.assert_array_equal(m.dump("cost"), st["cost"], err_msg="cost volume")
    for r in range(m.numPaths()):
        np.testing.assert_array_equal(m.dump("path", r), st["paths"][r],
                                      err_msg=f"path {oracle.PATH_NAMES[r]}")
    np.testing.assert_array_equal(m.dump("sum"), st["sum"], err_msg="path sum")
    np.testing.assert_array_equal(m.dump("wta_int"), st["wta_int"], err_msg="integer WTA")
    np.testing.assert_array_equal(m.dump("disp_left"), st["disp_left"], err_msg="left WTA")
    np.testing.assert_array_equal(m.dump("disp_right"), st["disp_right"], err_msg="right WTA")
    np.testing.assert_array_equal(m.dump("after_lr"), st["after_lr"], err_msg="LR check")
    np.testing.assert_array_equal(m.dump("after_speckle"), st["after_speckle"], err_msg="speckle")
    np.testing.assert_array_equal(got, want, err_msg="output")


def test_repeat_and_reuse(gpu_available):
    """Buffers are reused across calls of different sizes; results must not depend on history."""
    p = _params(0, 48, 1, 8, {})
    m = ls.StereoSGBM(p)
    for (W, H) in [(128, 96), (64, 48), (128, 96), (150, 50)]:
        left, right = ls.synth_pair(W, H, 2, 40, seed=W + H)
        got = m.compute(left, right)
        np.testing.assert_array_equal(got, oracle.sgbm(p, left, right))


def test_bgr_input(gpu_available):
    rng = np.random.default_rng(3)
    left, right = ls.synth_pair(96, 64, 2, 28, seed=11)
    lc = np.stack([left, np.roll(left, 1, 1), left[::-1, ::-1]], axis=2).copy()
    rc = np.stack([right, np.roll(right, 1, 1), right[::-1, ::-1]], axis=2).copy()
    p = _params(0, 32, 0, 4, {})
    got = ls.StereoSGBM(p).compute(lc, rc)
    assert got.dtype == np.int16 and got.shape == (64, 96)
    # OpenCV's fixed-point BGR->gray, written out here for the check.
    gl = ((lc[..., 0].astype(np.int32) * 1868 + lc[..., 1].astype(np.int32) * 9617 +
           lc[..., 2].astype(np.int32) * 4899 + 8192) >> 14).astype(np.uint8)
    gr = ((rc[..., 0].astype(np.int32) * 1868 + rc[..., 1].astype(np.int32) * 9617 +
           rc[..., 2].astype(np.int32) * 4899 + 8192) >> 14).astype(np.uint8)
    np.testing.assert_array_equal(got, oracle.sgbm(p, gl, gr))
    del rng


def test_compute_device_buffers(gpu_available):
    """lumen_sgbm_compute_device: images and output in the caller's device buffers. The output
    equals the oracle's, and the dump hook sees the same output (the handle keeps its own copy);
    a second, smaller frame through the same handle does the same."""
    import torch
    p = _params(0, 48, 1, 8, {"uniquenessRatio": 10})
    m = ls.StereoSGBM(p)
    for (W, H, seed) in [(160, 90, 1), (96, 40, 2)]:
        left, right = ls.synth_pair(W, H, 2, 40, seed=seed)
        dl = torch.from_numpy(left).cuda()
        dr = torch.from_numpy(right).cuda()
        do = torch.full((H, W), 12345, dtype=torch.int16, device="cuda")
        torch.cuda.synchronize()
        m.compute_device(dl.data_ptr(), dr.data_ptr(), W, H, do.data_ptr(), sync=True)
        got = do.cpu().numpy()
        np.testing.assert_array_equal(got, oracle.sgbm(p, left, right))
        np.testing.assert_array_equal(m.dump("output"), got)


def test_compute_device_unaligned_images(gpu_available):
    """Device images that do not start on a 4-byte boundary (views into larger buffers) give
    the output of aligned copies: the SAD cost then stages every segment column by column."""
    import torch
    W, H = 700, 48
    left, right = ls.synth_pair(W, H, 2, 60, seed=9)
    for bs in (3, 5, 7):
        p = _params(0, 64, 0, 4, {"blockSize": bs})
        m = ls.StereoSGBM(p)
        want = m.compute(left, right)
        for off in (1, 2, 3):
            bl = torch.zeros(W * H + 8, dtype=torch.uint8, device="cuda")
            br = torch.zeros(W * H + 8, dtype=torch.uint8, device="cuda")
            bl[off:off + W * H] = torch.from_numpy(left.reshape(-1)).cuda()
            br[off:off + W * H] = torch.from_numpy(right.reshape(-1)).cuda()
            do = torch.empty((H, W), dtype=torch.int16, device="cuda")
            torch.cuda.synchronize()
            m.compute_device(bl.data_ptr() + off, br.data_ptr() + off, W, H, do.data_ptr(), sync=True)
            np.testing.assert_array_equal(do.cpu().numpy(), want)


def test_bgr_input_large(gpu_available):
    """A 1080p BGR pair (converted in parallel bands on the host) gives exactly the output of
    its gray conversion."""
    left, right = ls.synth_pair(1920, 1080, 4, 60, seed=21)
    rng = np.random.default_rng(4)
    lc = np.stack([left, rng.integers(0, 256, left.shape, dtype=np.uint8), left[:, ::-1]], axis=2).copy()
    rc = np.stack([right, rng.integers(0, 256, right.shape, dtype=np.uint8), right[:, ::-1]], axis=2).copy()
    gl = ((lc[..., 0].astype(np.int32) * 1868 + lc[..., 1].astype(np.int32) * 9617 +
           lc[..., 2].astype(np.int32) * 4899 + 8192) >> 14).astype(np.uint8)
    gr = ((rc[..., 0].astype(np.int32) * 1868 + rc[..., 1].astype(np.int32) * 9617 +
           rc[..., 2].astype(np.int32) * 4899 + 8192) >> 14).astype(np.uint8)
    m = ls.StereoSGBM(_params(0, 64, 1, 4, {}))
    np.testing.assert_array_equal(m.compute(lc, rc), m.compute(gl, gr))


def test_timing_levels(gpu_available):
    """lumen_sgbm_set_timing: every level computes the same output; level 2 records every stage
    time, level 1 only the total (getLastComputeTimeMs), level 0 none; an unknown level is
    refused."""
    p = _params(0, 64, 1, 8, {})
    left, right = ls.synth_pair(160, 90, 2, 60, seed=5)
    m = ls.StereoSGBM(p)
    want = oracle.sgbm(p, left, right)
    for level in (2, 1, 0, 2):
        m.setTiming(level)
        np.testing.assert_array_equal(m.compute(left, right), want, err_msg=f"level {level}")
        t = m.timings()
        if level == 2:
            assert t[0] > 0 and t[1] > 0 and t[2] > 0 and t[4] > 0 and t[3] > 0
            assert t[1] + t[2] + t[3] + t[4] <= t[0] * 1.001
        elif level == 1:
            assert t[0] > 0 and t[5] > 0 and not (t[1] or t[2] or t[3] or t[4])
        else:
            assert not t.any()
    with pytest.raises(ls.StereoException, match="timing level"):
        m.setTiming(3)
