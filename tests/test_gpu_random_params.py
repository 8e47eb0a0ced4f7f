"""Seeded random parameter sweep: 48 small pairs, each with random size, disparity range and
offset, cost mode, path count, block size (odd and even; the reference does not enforce
isValid), penalties (including P2 <= P1), prefilter cap, uniqueness, LR tolerance and speckle
settings, compared with the oracle stage by stage. Catches corner cases that the hand-picked
parity cases miss."""
import numpy as np
import pytest

import lumenstereo_amd as ls
import oracle

pytestmark = pytest.mark.gpu


def random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    W = int(rng.integers(8, 200))
    H = int(rng.integers(3, 72))
    m = int(rng.choice([0, 0, 0, rng.integers(1, 24)]))
    D = int(rng.choice([int(rng.integers(1, 64)), int(rng.integers(64, 200)), 16, 64, 128]))
    p1 = int(rng.integers(1, 400))
    p = ls.StereoParams(
        minDisparity=m, maxDisparity=m + D,
        blockSize=int(rng.choice([1, 3, 4, 5, 6, 7, 9, 11])),
        P1=p1, P2=int(rng.choice([p1 + int(rng.integers(1, 2000)), max(1, p1 - 5)])),
        preFilterCap=int(rng.integers(1, 64)),
        uniquenessRatio=int(rng.choice([0, 0, 5, 15])),
        disp12MaxDiff=int(rng.choice([-1, 0, 1, 3])),
        speckleWindowSize=int(rng.choice([0, 0, 12, 20])), speckleRange=int(rng.choice([1, 4])),
        numDirections=int(rng.choice([4, 8])),
        matchingCostMode=ls.MatchingCostMode(int(rng.integers(0, 2))))
    dmax = max(3.0, min(D + m - 2.0, W / 2.0))
    left, right = ls.synth_pair(W, H, min(2.0, dmax - 1), dmax, seed=5000 + seed,
                                lowtex_frac=float(rng.choice([0.0, 0.3])))
    return p, left, right


@pytest.mark.parametrize("seed", range(48))
def test_random_parameters_match_oracle(gpu_available, seed):
    p, left, right = random_case(seed)
    want, st = oracle.sgbm(p, left, right, dumps=True)
    m = ls.StereoSGBM(p)
    got = m.compute(left, right)
    ctx = f"seed {seed}: {left.shape[1]}x{left.shape[0]} {p}"
    np.testing.assert_array_equal(m.dump("cost"), st["cost"], err_msg="cost " + ctx)
    np.testing.assert_array_equal(m.dump("sum"), st["sum"], err_msg="sum " + ctx)
    np.testing.assert_array_equal(m.dump("disp_left"), st["disp_left"], err_msg="left " + ctx)
    np.testing.assert_array_equal(m.dump("disp_right"), st["disp_right"], err_msg="right " + ctx)
    np.testing.assert_array_equal(got, want, err_msg="output " + ctx)
