"""Seeded random stereo cases shared by the parameter sweeps (imported by the tests, not collected).

Each case draws image size (a third of them 200-700 px wide), disparity range and offset, cost mode, path count, block size (odd
and even: the reference does not enforce isValid), penalties (including P2 <= P1), prefilter
cap, uniqueness, LR tolerance and speckle settings, and a synthetic pair whose true disparities
fit the range.
"""
import numpy as np

import lumenstereo_amd as ls


def random_case(seed):
    rng = np.random.default_rng(1000 + seed)
    W = int(rng.integers(8, 200))
    H = int(rng.integers(3, 72))
    if seed % 3 == 2:
        # Wider rows (several cost segments per row, census segments of 128-512 px) with fewer
        # rows, so the oracle stays quick.
        W = int(rng.integers(200, 700))
        H = int(rng.integers(3, 24))
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
