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
    # (W, H, D, census, paths)
    (96, 64, 64, False, 4),     # K = 4
    (130, 70, 128, True, 8),    # K = 8
    (120, 60, 170, True, 4),    # K = 12, padded
    (150, 50, 290, False, 4),   # K = 20, padded
    (100, 40, 384, True, 8),    # K = 24
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}_D{c[2]}_{'c' if c[3] else 's'}{c[4]}")
def test_every_candidate_schedule_is_bit_identical(gpu_available, case):
    W, H, D, census, paths = case
    p = tune.params_for(D, census, paths)
    left, right = ls.synth_pair(W, H, 4, max(5.0, 0.9 * D), seed=W + D)
    m = ls.StereoSGBM(p)
    K = (D + 63) // 64 * 4
    m.setSchedule(tune.default_schedule(K))
    ref = m.compute(left, right)
    np.testing.assert_array_equal(ref, oracle.sgbm(p, left, right))
    tiles = (128, 256, 512) if census else (256,)
    n = 0
    for (pf, prio), tile in itertools.product(tune.agg_candidates(K), tiles):
        m.setSchedule({"agg_prefetch": pf, "agg_priority": prio, "census_tile": tile})
        np.testing.assert_array_equal(m.compute(left, right), ref,
                                      err_msg=f"pf {pf} prio {prio} tile {tile}")
        n += 1
    assert n >= 6


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
    m.setSchedule(None)
    np.testing.assert_array_equal(m.compute(left, right), oracle.sgbm(p, left, right))
