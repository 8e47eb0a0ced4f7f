"""Every schedule the autotuner may choose gives bit-identical disparity maps (the schedule
only groups and orders work), and a schedule a slice cannot run is refused before any work."""
import itertools

import numpy as np
import pytest

import lumenstereo_amd as ls
import oracle
from lumenstereo_amd import tune

pytestmark = pytest.mark.gpu

CASES = [
    # (W, H, D, census, paths); 32-lane scanlines hold K/2 disparities per lane, 64-lane ones
    # (K a multiple of 8) K/4
    (96, 64, 64, False, 4),     # K = 4 (32-lane: one chunk of 2)
    (130, 70, 128, True, 8),    # K = 8 (32-lane: one chunk of 4; 64-lane: one chunk of 2)
    (120, 60, 170, True, 4),    # K = 12, padded (32-lane: three chunks of 2)
    (150, 50, 290, False, 4),   # K = 20, padded (32-lane: five chunks of 2)
    (100, 40, 384, True, 8),    # K = 24 (32-lane: three chunks of 4; 64-lane: three of 2)
    (60, 90, 40, True, 8),      # K = 4 padded (40 -> 64), taller than wide
    (100, 33, 250, False, 8),   # K = 16 padded (32-lane: one chunk of 8; 64-lane: one of 4)
    (70, 41, 512, True, 4),     # K = 32 (32-lane: one chunk of 16; 64-lane: one of 8)
    (80, 45, 448, False, 8),    # K = 28 (32-lane: seven chunks of 2)
    (110, 60, 96, True, 8),     # K = 6 (Dp = 96: three chunks of 2 on 16 lanes; no 32-lane)
    (90, 70, 75, False, 4),     # K = 6 padded (75 -> 96)
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}_D{c[2]}_{'c' if c[3] else 's'}{c[4]}")
def test_every_candidate_schedule_is_bit_identical(gpu_available, case):
    W, H, D, census, paths = case
    p = tune.params_for(D, census, paths)
    left, right = ls.synth_pair(W, H, 4, max(5.0, 0.9 * D), seed=W + D)
    m = ls.StereoSGBM(p)
    K = tune.padded_range(D) // 16
    m.setSchedule(tune.default_schedule(K))
    ref = m.compute(left, right)
    np.testing.assert_array_equal(ref, oracle.sgbm(p, left, right))
    tiles = (128, 256, 512) if census else (256,)
    n = 0
    # The occupancy cap cycles through its settings over the combinations.
    for (pf, prio), tile, wide in itertools.product(tune.agg_candidates(K), tiles,
                                                    tune.wide_settings(K)):
        occ = (0, 2, 3, 4)[n % 4]
        m.setSchedule({"agg_prefetch": pf, "agg_priority": prio, "census_tile": tile,
                       "agg_wide": wide, "agg_occupancy": occ})
        np.testing.assert_array_equal(
            m.compute(left, right), ref,
            err_msg=f"pf {pf} prio {prio} tile {tile} wide {wide} occupancy {occ}")
        n += 1
    assert n >= 6


@pytest.mark.parametrize("case", [(96, 64, 0, 64, 0, 4, {}), (130, 77, 0, 40, 1, 8, {}),
                                  (64, 150, 3, 131, 0, 8, {}), (100, 70, 0, 32, 1, 8,
                                                                 {"P1": 30000, "P2": 60000}),
                                  (90, 50, 0, 256, 1, 4, {}), (80, 30, 0, 512, 0, 8, {}),
                                  (90, 50, 0, 192, 1, 8, {}), (70, 40, 5, 295, 0, 4, {}),
                                  (60, 30, 0, 448, 1, 8, {}), (66, 35, 0, 384, 0, 4,
                                                               {"P1": 30000, "P2": 60000}),
                                  (100, 60, 0, 128, 1, 8, {}), (75, 44, 2, 102, 0, 8, {}),
                                  (50, 80, 0, 230, 1, 4, {}), (64, 36, 1, 385, 0, 8, {})],
                         ids=lambda c: f"{c[0]}x{c[1]}_d{c[2]}-{c[3]}_m{c[4]}_p{c[5]}")
def test_wide_scanlines_every_path_equals_oracle(gpu_available, case):
    """Every path on 32-lane scanlines, and where the padded range is a multiple of 128 on
    64-lane ones: each path's aggregated costs equal the oracle's (wide scanlines, padded
    ranges, taller-than-wide frames, minDisparity > 0, forced uint16 wrap)."""
    W, H, minD, maxD, mode, paths, extra = case
    p = ls.StereoParams(minDisparity=minD, maxDisparity=maxD, numDirections=paths,
                        matchingCostMode=ls.MatchingCostMode(mode), uniquenessRatio=0,
                        speckleWindowSize=0)
    for k, v in extra.items():
        setattr(p, k, v)
    left, right = ls.synth_pair(W, H, max(minD, 2), maxD - 4, seed=W * 31 + H)
    want, st = oracle.sgbm(p, left, right, dumps=True)
    m = ls.StereoSGBM(p)
    K = tune.padded_range(maxD - minD) // 16
    for wide in tune.wide_settings(K)[1:]:
        m.setSchedule({"agg_prefetch": 8, "agg_priority": 1, "census_tile": 256,
                       "agg_wide": wide})
        got = m.compute(left, right)
        for r in range(m.numPaths()):
            np.testing.assert_array_equal(m.dump("path", r), st["paths"][r],
                                          err_msg=f"wide {wide} path {oracle.PATH_NAMES[r]}")
        np.testing.assert_array_equal(got, want, err_msg=f"wide {wide} output")


def test_unsupported_schedule_is_refused(gpu_available):
    p = tune.params_for(290, True, 4)  # K = 20: 16 steps of 5 stores do not fit vmcnt
    m = ls.StereoSGBM(p)
    left, right = ls.synth_pair(64, 32, 4, 200, seed=1)
    m.setSchedule({"agg_prefetch": 16, "agg_priority": 1, "census_tile": 256})
    with pytest.raises(ls.StereoException, match="not supported"):
        m.compute(left, right)
    m.setSchedule({"agg_prefetch": 8, "agg_priority": 1, "census_tile": 100})
    with pytest.raises(ls.StereoException, match="census tile"):
        m.compute(left, right)
    m.setSchedule({"agg_prefetch": 4, "agg_priority": 1, "census_tile": 256, "agg_wide": 4})
    with pytest.raises(ls.StereoException, match="64-lane scanlines need a padded range"):
        m.compute(left, right)
    m.setSchedule({"agg_prefetch": 4, "agg_priority": 1, "census_tile": 256, "agg_wide": 7})
    with pytest.raises(ls.StereoException, match="lane setting"):
        m.compute(left, right)
    m.setSchedule({"agg_prefetch": 4, "agg_priority": 1, "census_tile": 256, "agg_wide": 0,
                   "agg_occupancy": 1})
    with pytest.raises(ls.StereoException, match="occupancy"):
        m.compute(left, right)
    m.setSchedule(None)
    np.testing.assert_array_equal(m.compute(left, right), oracle.sgbm(p, left, right))
