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
