"""Seeded random parameter sweep (cases from random_cases.py): each pair is compared with the
oracle stage by stage through the per-pair path, and random mixes of the same cases are run
through the batched and serving forms, which must equal the per-pair path pair by pair. Catches
corner cases that the hand-picked parity cases miss."""
import numpy as np
import pytest

import lumenstereo_amd as ls
import oracle
from random_cases import random_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed", range(96))
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


@pytest.mark.parametrize("seed", range(24))
def test_random_large_ranges_match_oracle(gpu_available, seed):
    """Disparity ranges of 193-512 (slices of 12 to 32 per lane, one to seven chunks on 16 or 32
    lanes), each under the tuned schedule and two random valid schedules."""
    import random
    from lumenstereo_amd import tune
    rng = random.Random(7000 + seed)
    W, H = rng.randint(40, 260), rng.randint(5, 40)
    m = rng.choice([0, 0, rng.randint(1, 20)])
    D = rng.randint(193, 512 - m)
    p1 = rng.randint(1, 400)
    p = ls.StereoParams(minDisparity=m, maxDisparity=m + D, blockSize=rng.choice([1, 3, 5, 7, 9]),
                        P1=p1, P2=rng.choice([p1 + rng.randint(1, 2000), max(1, p1 - 5)]),
                        preFilterCap=rng.randint(1, 63), uniquenessRatio=rng.choice([0, 10]),
                        disp12MaxDiff=rng.choice([-1, 1]), speckleWindowSize=0,
                        numDirections=rng.choice([4, 8]),
                        matchingCostMode=ls.MatchingCostMode(rng.randint(0, 1)))
    left, right = ls.synth_pair(W, H, 2.0, max(3.0, min(D + m - 2.0, W / 2.0)), seed=7000 + seed)
    want = oracle.sgbm(p, left, right)
    K = tune.padded_range(D) // 16
    m_ = ls.StereoSGBM(p)
    np.testing.assert_array_equal(m_.compute(left, right), want, err_msg=f"tuned {W}x{H} {p}")
    for _ in range(2):
        pf, prio = rng.choice(list(tune.agg_candidates(K)))
        sched = {"agg_prefetch": pf, "agg_priority": prio,
                 "census_tile": rng.choice([128, 256, 512]),
                 "agg_wide": rng.choice(list(tune.wide_settings(K))),
                 "agg_occupancy": rng.choice([0, 2, 3, 4])}
        m_.setSchedule(sched)
        np.testing.assert_array_equal(m_.compute(left, right), want, err_msg=f"{sched} {W}x{H} {p}")


@pytest.mark.parametrize("batch", range(4))
def test_random_batches_match_per_pair(gpu_available, batch):
    # 12 random cases per batch; a batch mixes every cost mode, path count, window and slice
    # width the draw gives, so the persistent kernels see them side by side.
    pairs = [random_case(100 + 12 * batch + i) for i in range(12)]
    got = ls.BatchSGBM(arena_bytes=256 << 20).run(pairs)
    for i, ((p, l, r), g) in enumerate(zip(pairs, got)):
        np.testing.assert_array_equal(g, ls.StereoSGBM(p).compute(l, r),
                                      err_msg=f"batch {batch} pair {i}: {l.shape[1]}x{l.shape[0]} {p}")


def test_random_mix_served_matches_per_pair(gpu_available):
    pairs = [random_case(200 + i) for i in range(32)]
    srv = ls.ServeSGBM(pool_bytes=1 << 30, lanes=3, max_batch=8)
    tickets = [srv.submit(p, l, r) for p, l, r in pairs]
    outs = []
    for t in tickets:
        out, info = srv.wait(t)
        assert info["status"] == 0
        outs.append(out)
    srv.close()
    for i, ((p, l, r), g) in enumerate(zip(pairs, outs)):
        np.testing.assert_array_equal(g, ls.StereoSGBM(p).compute(l, r),
                                      err_msg=f"served pair {i}: {l.shape[1]}x{l.shape[0]} {p}")
