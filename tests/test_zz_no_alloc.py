"""lumen_sgbm_compute_device_noalloc: the allocation-free entry point, called through the C ABI.

Two calls in a row, the second with a smaller pair (other cost mode, 8 paths, speckle on), each
bit-identical to lumen_sgbm_compute_device (the handle form) on the same device images; device
memory free before and after the second call is the same (nothing is allocated); a pair larger
than the scratch is refused with LUMEN_EINVAL and a message giving the sizes.
"""
import ctypes

import numpy as np
import pytest

import lumenstereo_amd as ls

pytestmark = pytest.mark.gpu


def _params(D, census, paths, speckle=0):
    return ls.StereoParams(maxDisparity=D, numDirections=paths,
                           matchingCostMode=ls.MatchingCostMode.Census if census
                           else ls.MatchingCostMode.SAD,
                           uniquenessRatio=0, speckleWindowSize=speckle, speckleRange=2)


def test_noalloc_equals_compute_device(gpu_available):
    import torch

    lib = ls.library()
    cases = [(640, 480, _params(64, False, 4), 20240611),    # the 640x480 D=64 SAD 4-path frame
             (320, 200, _params(48, True, 8, speckle=20), 7)]
    bufs = []
    for W, H, p, seed in cases:
        left, right = ls.synth_pair(W, H, 4, p.maxDisparity - 4, seed=seed)
        bufs.append((torch.from_numpy(left).cuda(), torch.from_numpy(right).cuda(),
                     torch.full((H, W), -7, dtype=torch.int16, device="cuda"),
                     torch.full((H, W), -9, dtype=torch.int16, device="cuda")))
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream().cuda_stream
    free = []
    for (W, H, p, _), (dl, dr, d_ref, d_out) in zip(cases, bufs):
        print(f"noalloc {W}x{H} D={p.maxDisparity} paths={p.numDirections}", flush=True)
        ls.StereoSGBM(p).compute_device(dl.data_ptr(), dr.data_ptr(), W, H, d_ref.data_ptr(),
                                        sync=True)
        torch.cuda.synchronize()
        free.append(torch.cuda.mem_get_info()[0])
        cp = p.to_c()
        rc = lib.lumen_sgbm_compute_device_noalloc(ctypes.byref(cp), dl.data_ptr(), dr.data_ptr(),
                                                   W, H, d_out.data_ptr(), stream)
        assert rc == 0, lib.lumen_last_error().decode()
        torch.cuda.synchronize()
        free[-1] -= torch.cuda.mem_get_info()[0]
        np.testing.assert_array_equal(d_out.cpu().numpy(), d_ref.cpu().numpy(),
                                      err_msg=f"{W}x{H}: noalloc vs compute_device")
    # The first call may load the library's scratch; the second allocates nothing.
    assert free[1] == 0, f"device memory changed by {free[1]} bytes"


def test_noalloc_refuses_pairs_beyond_the_scratch(gpu_available):
    import torch

    lib = ls.library()
    W, H = 1920, 1080
    p = _params(128, True, 8)
    d = torch.zeros((H, W), dtype=torch.uint8, device="cuda")
    o = torch.zeros((H, W), dtype=torch.int16, device="cuda")
    torch.cuda.synchronize()
    cp = p.to_c()
    rc = lib.lumen_sgbm_compute_device_noalloc(ctypes.byref(cp), d.data_ptr(), d.data_ptr(), W,
                                               H, o.data_ptr(), None)
    assert rc == ls.LUMEN_EINVAL
    msg = lib.lumen_last_error().decode()
    assert "bytes of working memory" in msg and str(192 << 20) in msg
    bad = _params(64, False, 4)
    bad.minDisparity, bad.maxDisparity = 10, 5
    cb = bad.to_c()
    assert lib.lumen_sgbm_compute_device_noalloc(ctypes.byref(cb), d.data_ptr(), d.data_ptr(),
                                                 64, 48, o.data_ptr(), None) == ls.LUMEN_EINVAL
    assert "disparity range" in lib.lumen_last_error().decode()
