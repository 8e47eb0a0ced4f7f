"""Batched form: every pair of a mixed batch must be bit-identical to the per-pair path."""
import numpy as np
import pytest

import lumenstereo_amd as ls
import oracle

pytestmark = pytest.mark.gpu


def mixed_pairs(n, seed, sizes=((96, 64), (130, 77), (64, 48), (200, 90), (160, 120))):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        W, H = sizes[i % len(sizes)]
        D = int(rng.choice([16, 32, 40, 64, 96]))
        p = ls.StereoParams(maxDisparity=D, numDirections=int(rng.choice([4, 8])),
                            matchingCostMode=ls.MatchingCostMode(int(rng.integers(0, 2))),
                            uniquenessRatio=int(rng.choice([0, 10])),
                            speckleWindowSize=int(rng.choice([0, 0, 12])), speckleRange=2,
                            disp12MaxDiff=int(rng.choice([-1, 1])),
                            blockSize=int(rng.choice([3, 5, 7])))
        left, right = ls.synth_pair(W, H, 2, max(4, D - 6), seed=seed * 100 + i)
        out.append((p, left, right))
    return out


def plain(pairs):
    return [ls.StereoSGBM(p).compute(l, r) for p, l, r in pairs]


def test_batch_matches_plain_and_oracle(gpu_available):
    pairs = mixed_pairs(12, 1)
    b = ls.BatchSGBM(arena_bytes=256 << 20)
    got = b.run(pairs)
    want = plain(pairs)
    for i, (g, w) in enumerate(zip(got, want)):
        np.testing.assert_array_equal(g, w, err_msg=f"pair {i}")
    for i in (0, 3, 7):
        p, l, r = pairs[i]
        np.testing.assert_array_equal(got[i], oracle.sgbm(p, l, r), err_msg=f"pair {i} vs oracle")


def test_batch_replay_and_different_batches(gpu_available):
    b = ls.BatchSGBM(arena_bytes=256 << 20)
    a = mixed_pairs(6, 2)
    c = mixed_pairs(9, 3)
    first = b.run(a)
    second = b.run(a)            # graph replay
    third = b.run(c)             # different composition, same graph when it fits
    fourth = b.run(a)
    for x, y in zip(first, second):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(first, fourth):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(third, plain(c)):
        np.testing.assert_array_equal(x, y)


def test_batch_large_slices_and_special_windows(gpu_available):
    # D=200 (K=16) and D=290 (K=20) go to the big-slice kernels; blockSize 65 (r=32) and
    # blockSize -3 (empty window) take the per-pair cost kernels inside a batch.
    pairs = []
    for i, (D, bs, mode) in enumerate([(200, 5, 1), (290, 5, 0), (48, 65, 0), (32, -3, 0),
                                       (64, 5, 1)]):
        p = ls.StereoParams(maxDisparity=D, numDirections=8 if i % 2 else 4, blockSize=bs,
                            matchingCostMode=ls.MatchingCostMode(mode), uniquenessRatio=0,
                            speckleWindowSize=0)
        left, right = ls.synth_pair(320, 40, 2, min(D - 6, 250), seed=40 + i)
        pairs.append((p, left, right))
    got = ls.BatchSGBM(arena_bytes=512 << 20).run(pairs)
    for i, (g, w) in enumerate(zip(got, plain(pairs))):
        np.testing.assert_array_equal(g, w, err_msg=f"pair {i}")


def test_batch_splits_when_arena_is_small(gpu_available):
    import ctypes
    pairs = mixed_pairs(8, 4)
    need = []
    for p, l, _ in pairs:
        v = ctypes.c_size_t()
        ls.library().lumen_sgbm_footprint(ctypes.byref(p.to_c()), l.shape[1], l.shape[0],
                                          ctypes.byref(v))
        need.append(v.value)
    b = ls.BatchSGBM(arena_bytes=int(max(need) * 1.6) + (1 << 20))   # one or two pairs at a time
    got = b.run(pairs)
    for i, (g, w) in enumerate(zip(got, plain(pairs))):
        np.testing.assert_array_equal(g, w, err_msg=f"pair {i}")


def test_batch_of_one_pair_between_larger_batches(gpu_available):
    """A batch of one pair takes the per-pair kernels; runs of one pair and of several on one
    engine, in any order, all equal the per-pair path."""
    b = ls.BatchSGBM(arena_bytes=512 << 20)
    singles = mixed_pairs(4, 6)
    p = ls.StereoParams(maxDisparity=48, blockSize=65, uniquenessRatio=0, speckleWindowSize=0)
    left, right = ls.synth_pair(120, 40, 2, 40, seed=61)
    singles.append((p, left, right))          # SAD window beyond the LDS segment
    many = mixed_pairs(7, 7)
    want_many = plain(many)
    for i, pr in enumerate(singles):
        np.testing.assert_array_equal(b.run([pr])[0], plain([pr])[0], err_msg=f"single {i}")
        for k, (g, w) in enumerate(zip(b.run(many), want_many)):
            np.testing.assert_array_equal(g, w, err_msg=f"after single {i}: pair {k}")


def test_host_batch_in_pipelined_chunks(gpu_available):
    """Host-memory batches of 16 or more pairs run as four chunks alternating over two engines
    (2 to 15 pairs: two chunks); every output equals the per-pair path, on repeated runs."""
    b = ls.BatchSGBM(arena_bytes=1 << 30)
    for n, seed in ((20, 8), (5, 9), (17, 10)):
        pairs = mixed_pairs(n, seed)
        want = plain(pairs)
        for rep in range(2):
            for i, (g, w) in enumerate(zip(b.run(pairs), want)):
                np.testing.assert_array_equal(g, w, err_msg=f"{n} pairs, run {rep}: pair {i}")
    t = b.timings()
    assert t[0] > 0 and t[2] > 0


def test_batch_rejects_invalid_pair(gpu_available):
    pairs = mixed_pairs(2, 5)
    bad = ls.StereoParams(minDisparity=10, maxDisparity=10)
    pairs.append((bad, pairs[0][1], pairs[0][2]))
    with pytest.raises(ls.StereoException, match="pair 2"):
        ls.BatchSGBM(arena_bytes=64 << 20).run(pairs)


def test_batch_high_prefilter_caps(gpu_available):
    """SAD with prefilter caps of 128 and more (the prefiltered byte wraps as the uint8 store
    does) at block sizes 3, 5, 7 and 9, batched, equals the per-pair path and the oracle."""
    pairs = []
    for i, (cap, bs) in enumerate([(200, 3), (200, 5), (1000, 7), (1000, 5), (200, 9), (128, 5)]):
        p = ls.StereoParams(maxDisparity=48, numDirections=4 + 4 * (i % 2), preFilterCap=cap,
                            blockSize=bs, uniquenessRatio=0, speckleWindowSize=0)
        l, r = ls.synth_pair(96 + 8 * i, 64, 2, 40, seed=500 + i)
        pairs.append((p, l, r))
    got = ls.BatchSGBM(arena_bytes=256 << 20).run(pairs)
    for i, ((p, l, r), g) in enumerate(zip(pairs, got)):
        np.testing.assert_array_equal(g, ls.StereoSGBM(p).compute(l, r), err_msg=f"pair {i}")
        np.testing.assert_array_equal(g, oracle.sgbm(p, l, r), err_msg=f"pair {i} vs oracle")
