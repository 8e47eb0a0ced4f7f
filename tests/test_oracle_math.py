"""Pins the oracle's semantics with what the mathematics fixes: closed forms, known answers,
invariants and brute force on tiny inputs. Each check is chosen so that a plausible
transcription mistake (a dropped truncation, a wrong neighbour, a transposed operand, an
off-by-one border) fails it.
"""
import itertools

import numpy as np
import pytest

import oracle

rng = np.random.default_rng(1234)


# ----------------------------------------------------------------------------- prefilter
def test_prefilter_ramp_closed_form():
    img = np.tile((np.arange(20) * 3).astype(np.uint8), (4, 1))
    f = oracle.prefilter(img, 31)
    assert (f[:, 1:-1] == 31 + 6).all()      # central difference 3(x+1) - 3(x-1)
    assert (f[:, 0] == 31 + 3).all() and (f[:, -1] == 31 + 3).all()  # one-sided edges


def test_prefilter_clipping_and_cap_floor():
    img = np.array([[0, 255, 0, 255, 0]], np.uint8)
    np.testing.assert_array_equal(oracle.prefilter(img, 31), [[62, 31, 31, 31, 0]])
    # preFilterCap <= 0 behaves as cap = 1 (src/sgbm_gpu.cu:227)
    np.testing.assert_array_equal(oracle.prefilter(img, 0), [[2, 1, 1, 1, 0]])


# ----------------------------------------------------------------------------- census
def test_census_constant_and_border():
    c = oracle.census(np.full((9, 11), 77, np.uint8))
    assert (c[2:-2, 2:-2] == 0xFFFFFF).all()      # every neighbour >= centre
    assert (c[:2] == 0).all() and (c[-2:] == 0).all()
    assert (c[:, :2] == 0).all() and (c[:, -2:] == 0).all()


@pytest.mark.parametrize("dy,dx,bit", [(-2, -2, 0), (-2, 2, 4), (-1, -2, 5), (0, -1, 11),
                                       (0, 1, 12), (1, 0, 16), (2, 2, 23)])
def test_census_bit_order(dy, dx, bit):
    img = np.full((5, 5), 10, np.uint8)
    img[2, 2] = 100
    img[2 + dy, 2 + dx] = 200          # the only neighbour >= centre
    assert oracle.census(img)[2, 2] == 1 << bit


def test_census_dark_centre():
    img = np.full((5, 5), 10, np.uint8)
    img[2, 2] = 0
    assert oracle.census(img)[2, 2] == 0xFFFFFF


@pytest.mark.parametrize("h,want", [(0, 0), (1, 11), (12, 128), (23, 244), (24, 255)])
def test_census_hamming_table(h, want):
    W = 8
    cl = np.zeros((5, W), np.uint32)
    cr = np.zeros((5, W), np.uint32)
    cl[2, 5] = (1 << h) - 1                       # h set bits against code 0
    cost = oracle.cost_census(cl, cr, 2, 0)
    assert cost[2, 5, 0] == want
    assert cost[1, 5, 0] == 255                   # border row
    assert cost[2, 5, 1] == (0 if h == 0 else want)  # rx = 4 is inside too


def test_census_cost_out_of_range_is_255():
    cl = np.zeros((6, 10), np.uint32)
    cr = np.zeros((6, 10), np.uint32)
    cost = oracle.cost_census(cl, cr, 6, 1)
    # x - (m + d) < 2 -> 255
    for x in range(2, 8):
        for d in range(6):
            assert cost[2, x, d] == (0 if x - 1 - d >= 2 else 255)


# ----------------------------------------------------------------------------- SAD
def _sad_integral(fl, fr, D, m, r):
    """Box sums via integral images (a different algorithm from the oracle's tap loop)."""
    H, W = fl.shape
    out = np.empty((H, W, D), np.uint8)
    for d in range(D):
        a = m + d
        ad = np.zeros((H, W), np.int64)
        val = np.zeros((H, W), np.int64)
        for x in range(W):
            if 0 <= x - a < W:
                ad[:, x] = np.abs(fl[:, x].astype(np.int64) - fr[:, x - a].astype(np.int64))
                val[:, x] = 1
        ii = np.pad(ad.cumsum(0).cumsum(1), ((1, 0), (1, 0)))
        iv = np.pad(val.cumsum(0).cumsum(1), ((1, 0), (1, 0)))
        for y in range(H):
            y0, y1 = max(0, y - r), min(H, y + r + 1)
            for x in range(W):
                x0, x1 = max(0, x - r), min(W, x + r + 1)
                s = ii[y1, x1] - ii[y0, x1] - ii[y1, x0] + ii[y0, x0]
                n = iv[y1, x1] - iv[y0, x1] - iv[y1, x0] + iv[y0, x0]
                out[y, x, d] = s // n if n else 255
    return out


@pytest.mark.parametrize("bs,m", [(5, 0), (3, 2), (1, 0), (7, 1)])
def test_sad_matches_integral_images(bs, m):
    fl = rng.integers(0, 63, (9, 13)).astype(np.uint8)
    fr = rng.integers(0, 63, (9, 13)).astype(np.uint8)
    np.testing.assert_array_equal(oracle.cost_sad(fl, fr, 6, m, bs), _sad_integral(fl, fr, 6, m, bs // 2))


def test_sad_invariants():
    fl = rng.integers(0, 63, (10, 16)).astype(np.uint8)
    c = oracle.cost_sad(fl, fl, 5, 0, 5)
    assert (c[:, :, 0] == 0).all()                       # identical images at d = 0
    shifted = np.zeros_like(fl)
    shifted[:, :-3] = fl[:, 3:]                          # right = left moved by 3 px
    c = oracle.cost_sad(fl, shifted, 6, 0, 5)
    assert (c[:, 3 + 2 + 3:-2, 3] == 0).all()            # exact match at the true disparity
    # No valid column at all: x + r < m + d  ->  255
    c = oracle.cost_sad(fl, fl, 8, 0, 3)
    for x in range(16):
        for d in range(8):
            if x + 1 < d:
                assert (c[:, x, d] == 255).all()
            else:
                assert (c[:, x, d] <= 62).all()


def test_cost_scale():
    assert oracle.cost_scale(0, 5, 31) == 25
    assert oracle.cost_scale(1, 5, 31) == 6
    assert oracle.cost_scale(1, 5, 0) == 1
    assert oracle.cost_scale(2, 5, 31) == 6        # any non-SAD mode uses the census formula
    assert oracle.cost_scale(0, 4, 31) == 16


# ----------------------------------------------------------------------------- aggregation
def _p2adj(P1, P2, a, b):
    g = abs(int(a) - int(b))
    return max(P1 + 1, P2 * 8 // max(8, g))


def _brute_path_energy(chat, img, P1, P2):
    """Minimum over all labelings of a 1-D chain: E*(i, d) = min energy of a labeling of
    pixels 0..i ending with label d, with pairwise term 0 / P1 / P2adj."""
    n, D = chat.shape
    E = np.full((n, D), np.inf)
    for labels in itertools.product(range(D), repeat=n):
        e = chat[0, labels[0]]
        E[0, labels[0]] = min(E[0, labels[0]], e)
        for i in range(1, n):
            dd = abs(labels[i] - labels[i - 1])
            pen = 0 if dd == 0 else (P1 if dd == 1 else _p2adj(P1, P2, img[i], img[i - 1]))
            e += chat[i, labels[i]] + pen
            E[i, labels[i]] = min(E[i, labels[i]], e)
    return E


@pytest.mark.parametrize("seed", range(4))
def test_path_recurrence_equals_brute_force_min_energy(seed):
    r = np.random.default_rng(seed)
    W, D, cs = 5, 3, 7
    cost = r.integers(0, 256, (1, W, D)).astype(np.uint8)
    img = r.integers(0, 256, (1, W)).astype(np.uint8)
    P1, P2 = 150, 900
    L = oracle.aggregate_path(cost, img, P1, P2, cs, 0).astype(np.int64)   # LR
    chat = cost[0].astype(np.int64) * cs
    E = _brute_path_energy(chat, img[0], P1, P2)
    assert (L[0, 0] == chat[0]).all()
    for i in range(1, W):
        np.testing.assert_array_equal(L[0, i], E[i] - E[i - 1].min())


@pytest.mark.parametrize("path", [4, 5, 6, 7])
def test_diagonal_paths_equal_brute_force_along_each_diagonal(path):
    r = np.random.default_rng(path)
    H, W, D, cs = 3, 4, 3, 5
    cost = r.integers(0, 256, (H, W, D)).astype(np.uint8)
    img = r.integers(0, 256, (H, W)).astype(np.uint8)
    P1, P2 = 100, 700
    L = oracle.aggregate_path(cost, img, P1, P2, cs, path).astype(np.int64)
    dx, dy = [(1, 1), (-1, -1), (-1, 1), (1, -1)][path - 4]
    for y in range(H):
        for x in range(W):
            if 0 <= x - dx < W and 0 <= y - dy < H:
                continue                       # not an origin
            chain = []
            cx, cy = x, y
            while 0 <= cx < W and 0 <= cy < H:
                chain.append((cy, cx))
                cx, cy = cx + dx, cy + dy
            chat = np.array([cost[p] for p in chain], np.int64) * cs
            im = np.array([img[p] for p in chain])
            E = _brute_path_energy(chat, im, P1, P2)
            np.testing.assert_array_equal(L[chain[0]], chat[0])
            for i in range(1, len(chain)):
                np.testing.assert_array_equal(L[chain[i]], E[i] - E[i - 1].min())


@pytest.mark.parametrize("path", range(8))
def test_single_disparity_and_constant_cost(path):
    cost = rng.integers(0, 256, (6, 7, 1)).astype(np.uint8)
    img = rng.integers(0, 256, (6, 7)).astype(np.uint8)
    L = oracle.aggregate_path(cost, img, 200, 800, 6, path)
    np.testing.assert_array_equal(L[..., 0], (cost[..., 0].astype(np.int64) * 6) & 0xFFFF)
    cost = np.full((6, 7, 5), 17, np.uint8)
    L = oracle.aggregate_path(cost, img, 200, 800, 6, path)
    assert (L == 17 * 6).all()


@pytest.mark.parametrize("path", range(8))
def test_no_wrap_bounds(path):
    cost = rng.integers(0, 256, (8, 9, 6)).astype(np.uint8)
    img = rng.integers(0, 256, (8, 9)).astype(np.uint8)
    L = oracle.aggregate_path(cost, img, 200, 800, 25, path).astype(np.int64)
    chat = cost.astype(np.int64) * 25
    assert (L >= chat).all() and (L <= chat + 800).all()


def test_uint16_wrap_known_answers():
    """Hand-computed: P1 = 65530, P2adj = 65534 (equal image), two pixels, D = 2, cs = 1.
    M = 10, c3 = t16(10 + 65534) = 8; d=0: min(10, -, t16(20+65530)=14, 8) = 8 -> 30+8-10 = 28;
    d=1: min(20, t16(10+65530)=4, -, 8) = 4 -> 40+4-10 = 34. Without the truncations the
    answer would be (30, 40)."""
    img = np.zeros((1, 2), np.uint8)
    cost = np.array([[[10, 20], [30, 40]]], np.uint8)
    L = oracle.aggregate_path(cost, img, 65530, 65534, 1, 0)
    np.testing.assert_array_equal(L[0, 1], [28, 34])
    cost = np.array([[[10, 20], [0, 1]]], np.uint8)       # negative results wrap too
    L = oracle.aggregate_path(cost, img, 65530, 65534, 1, 0)
    np.testing.assert_array_equal(L[0, 1], [65534, 65531])


# ----------------------------------------------------------------------------- WTA
def _wta_vec(s, m=0, uniq=0):
    S = np.array(s, np.uint16).reshape(1, 1, -1)
    d, wi = oracle.wta_left(S, m, uniq)
    return int(d[0, 0]), int(wi[0, 0])


def test_wta_left_known_answers():
    assert _wta_vec([10, 0, 10]) == (16, 1)                # symmetric parabola: exact
    assert _wta_vec([8, 0, 0]) == (24, 1)                  # first minimum, delta = 1/2
    assert _wta_vec([10, 2, 9]) == (16, 1)                 # 1 + 1/30 -> 16.53 -> 16
    assert _wta_vec([9, 2, 10]) == (15, 1)                 # 1 - 1/30 -> 15.47 -> 15
    assert _wta_vec([0, 5, 6]) == (0, 0)                   # border minimum: no parabola
    assert _wta_vec([6, 5, 0]) == (32, 2)
    assert _wta_vec([65535] * 4) == (0, 0)                 # nothing beats the initial value
    assert _wta_vec([4, 1, 9], m=3) == (60, 1)             # (1 - 5/22 + 3) * 16 = 60.36


def test_wta_left_parabola_with_min_disparity():
    # c = (4,1,9): denom = 2*(4+9-2) = 22, delta = (4-9)/22 = -0.22727..; (1 - 0.22727 + 3) * 16
    d, _ = _wta_vec([4, 1, 9], m=3)
    assert d == int(np.float32(np.float32(1.0) + np.float32(-5.0) / np.float32(22.0) + np.float32(3.0)) * 16)


def test_wta_left_uniqueness():
    assert _wta_vec([5, 3, 3, 7], uniq=10)[0] == -1        # tie: second == best
    assert _wta_vec([100, 50, 54, 300], uniq=10)[0] == -1  # 54 <= 50 + 5
    assert _wta_vec([100, 50, 56, 300], uniq=10)[0] != -1  # 56 > 55
    assert _wta_vec([100, 50, 54, 300], uniq=0)[0] != -1   # ratio 0 disables the test
    assert _wta_vec([65535, 50, 65535], uniq=10)[0] != -1  # second == 65535 skips the test


def test_wta_left_brute_force():
    S = rng.integers(0, 400, (6, 7, 9)).astype(np.uint16)
    d, wi = oracle.wta_left(S, 0, 0)
    np.testing.assert_array_equal(wi, S.argmin(axis=2))


def test_wta_right_brute_force():
    H, W, D, m = 4, 12, 5, 1
    S = rng.integers(0, 300, (H, W, D)).astype(np.uint16)
    got = oracle.wta_right(S, m)
    for y in range(H):
        for x in range(W):
            cands = [(int(S[y, x + m + d, d]), d) for d in range(D) if x + m + d < W]
            best = min([c for c in cands if c[0] < 65535], default=None)
            bd = best[1] if best else 0
            sub = np.float32(bd)
            if 0 < bd < D - 1:
                c0, c1, c2 = (np.float32(S[y, x + m + bd, bd + k]) for k in (-1, 0, 1))
                den = np.float32(2) * (c0 + c2 - np.float32(2) * c1)
                if abs(den) > np.float32(1e-6):
                    sub = np.float32(sub + np.clip(np.float32((c0 - c2) / den), -1, 1))
            assert got[y, x] == int(np.trunc(np.float32((sub + np.float32(m)) * np.float32(16))))


# ----------------------------------------------------------------------------- post
def test_lr_check_brute_force():
    dl = rng.integers(-1, 200, (5, 20)).astype(np.int16)
    dr = rng.integers(-1, 200, (5, 20)).astype(np.int16)
    got = oracle.lr_check(dl, dr, 16)
    for y in range(5):
        for x in range(20):
            v = int(dl[y, x])
            if v > 0:
                xr = x - v // 16
                if not (0 <= xr < 20) or dr[y, xr] <= 0 or abs(v - int(dr[y, xr])) > 16:
                    v = -1
            assert got[y, x] == v


def test_median_brute_force_and_hole_fill():
    d = rng.integers(-2, 50, (7, 9)).astype(np.int16)
    got = oracle.median(d)
    for y in range(7):
        for x in range(9):
            v = sorted(int(d[j, i]) for j in range(max(0, y - 1), min(7, y + 2))
                       for i in range(max(0, x - 1), min(9, x + 2)) if d[j, i] > 0)
            assert got[y, x] == (v[len(v) // 2] if v else d[y, x])
    hole = np.array([[1, 2, 3], [4, -1, 6], [7, 8, 9]], np.int16)
    assert oracle.median(hole)[1, 1] == 6                  # upper median of 8 values


def test_speckle_snapshot_known_answer():
    d = np.full((5, 5), -1, np.int16)
    d[2, 2] = 100
    d[0, 0] = 100
    out = oracle.speckle(d, 2, 16)
    assert out[2, 2] == 100            # itself + (0,0) within the 5x5 window
    d[0, 0] = 200
    assert oracle.speckle(d, 2, 16)[2, 2] == -1


def _speckle_padded(d, min_similar, max_diff):
    """Speckle (src/sgbm_kernels.cu:979-1013) counted on a zero-padded copy: 25 shifted
    views of the snapshot instead of a bounds-checked window walk. Zero is not > 0, so padding
    adds nothing, which is what skipping out-of-image neighbours does."""
    H, W = d.shape
    pad = np.zeros((H + 4, W + 4), np.int32)
    pad[2:-2, 2:-2] = d
    c = d.astype(np.int32)
    n = np.zeros((H, W), np.int32)
    for dy in range(5):
        for dx in range(5):
            v = pad[dy:dy + H, dx:dx + W]
            n += (v > 0) & (np.abs(c - v) <= max_diff)
    out = d.copy()
    out[(c > 0) & (n < min_similar)] = -1
    return out


def _speckle_in_place(d, min_similar, max_diff):
    """The reference kernel run by one thread in raster order, updating in place (one of the
    orders its racing threads can produce)."""
    d = d.copy()
    H, W = d.shape
    for y in range(H):
        for x in range(W):
            c = int(d[y, x])
            if c <= 0:
                continue
            n = sum(1 for j in range(max(0, y - 2), min(H, y + 3))
                    for i in range(max(0, x - 2), min(W, x + 3))
                    if d[j, i] > 0 and abs(c - int(d[j, i])) <= max_diff)
            if n < min_similar:
                d[y, x] = -1
    return d


@pytest.mark.parametrize("seed", range(6))
def test_speckle_random_maps_equal_padded_count(seed):
    r = np.random.default_rng(seed)
    H, W = int(r.integers(1, 12)), int(r.integers(1, 14))
    # Values around a few levels, holes and zeros, so counts vary across the threshold.
    d = (r.integers(0, 4, (H, W)) * 40 + r.integers(-6, 7, (H, W))).astype(np.int16)
    d[r.random((H, W)) < 0.2] = -1
    d[r.random((H, W)) < 0.1] = 0
    for min_similar in (1, 4, 9, 25):
        for max_diff in (0, 5, 16):
            np.testing.assert_array_equal(oracle.speckle(d, min_similar, max_diff),
                                          _speckle_padded(d, min_similar, max_diff))


def test_speckle_border_windows():
    # A constant valid map: a pixel's count is the number of in-image pixels of its window,
    # 9 at a corner, 12 next to it, 15 on an edge two pixels in, 25 inside.
    d = np.full((7, 8), 48, np.int16)
    counts = np.array([[min(y, 2) + min(6 - y, 2) + 1 for _ in range(8)] for y in range(7)]) * \
        np.array([[min(x, 2) + min(7 - x, 2) + 1 for x in range(8)] for _ in range(7)])
    assert counts[0, 0] == 9 and counts[0, 1] == 12 and counts[3, 3] == 25
    for min_similar in (9, 10, 12, 13, 15, 16, 25):
        want = np.where(counts >= min_similar, 48, -1).astype(np.int16)
        np.testing.assert_array_equal(oracle.speckle(d, min_similar, 0), want)


def test_speckle_snapshot_differs_from_in_place():
    # One row of six equal values with minSimilar 4: the end pixels see 3 similar pixels, the
    # others 4 or 5. Reading the snapshot, only the ends go. Updating in raster order, each
    # removal drops the next pixel's count below 4 and the whole row goes.
    d = np.full((1, 6), 100, np.int16)
    snap = oracle.speckle(d, 4, 0)
    np.testing.assert_array_equal(snap, [[-1, 100, 100, 100, 100, -1]])
    np.testing.assert_array_equal(_speckle_in_place(d, 4, 0), np.full((1, 6), -1))


def test_speckle_window_mapping_in_pipeline():
    """compute maps speckleWindowSize to min(speckleWindowSize, 25) similar pixels and
    speckleRange to 16*speckleRange (src/sgbm_gpu.cu:480-488): the pipeline's after-speckle map
    equals the padded count applied to its after-LR map with those values."""
    import lumenstereo_amd as ls
    left, right = ls.synth_pair(64, 40, 2, 28, seed=77)
    for sws, rng_ in ((100, 2), (25, 2), (24, 2), (7, 1), (3, 0)):
        p = dict(minDisparity=0, maxDisparity=32, blockSize=5, P1=200, P2=800, preFilterCap=31,
                 uniquenessRatio=0, disp12MaxDiff=1, speckleWindowSize=sws, speckleRange=rng_,
                 useGPU=1, useSharedMemory=1, numDirections=4, matchingCostMode=0)
        _, st = oracle.sgbm(p, left, right, dumps=True)
        np.testing.assert_array_equal(st["after_speckle"],
                                      _speckle_padded(st["after_lr"], min(sws, 25), 16 * rng_))
    # 100 and 25 are the same filter; 24 keeps pixels that 25 removes.
    outs = {}
    for sws in (100, 25, 24):
        p["speckleWindowSize"], p["speckleRange"] = sws, 2
        outs[sws] = oracle.sgbm(p, left, right, dumps=True)[1]["after_speckle"]
    np.testing.assert_array_equal(outs[100], outs[25])
    assert np.count_nonzero(outs[24] != outs[25]) > 0


@pytest.mark.parametrize("case", [(130, 77, 0, 40, 0, 8, {}), (96, 64, 3, 67, 1, 4, {}),
                                  (100, 70, 0, 32, 0, 8, {"speckleWindowSize": 20}),
                                  (90, 40, 0, 290, 1, 4, {"uniquenessRatio": 10})])
def test_threaded_oracle_equals_single_threaded(case):
    """The threaded form (bands of rows, one path per thread) that checks full-size frames
    gives exactly the single-threaded oracle's output and stage dumps."""
    import lumenstereo_amd as ls
    W, H, minD, maxD, mode, paths, extra = case
    left, right = ls.synth_pair(W, H, 2, maxD - minD - 4, seed=W + H)
    p = dict(minDisparity=minD, maxDisparity=maxD, blockSize=5, P1=200, P2=800,
             preFilterCap=31, uniquenessRatio=0, disp12MaxDiff=1, speckleWindowSize=0,
             speckleRange=2, useGPU=1, useSharedMemory=1, numDirections=paths,
             matchingCostMode=mode)
    p.update(extra)
    o1, s1 = oracle.sgbm(p, left, right, dumps=True)
    o7, s7 = oracle.sgbm(p, left, right, dumps=True, threads=7)
    np.testing.assert_array_equal(o1, o7)
    for k in s1:
        np.testing.assert_array_equal(s1[k], s7[k], err_msg=k)


def test_full_pipeline_rejects_bad_ranges():
    p = dict(minDisparity=0, maxDisparity=0, blockSize=5, P1=200, P2=800, preFilterCap=31,
             uniquenessRatio=0, disp12MaxDiff=1, speckleWindowSize=0, speckleRange=32,
             useGPU=1, useSharedMemory=1, numDirections=4, matchingCostMode=0)
    img = np.zeros((8, 8), np.uint8)
    with pytest.raises(ValueError):
        oracle.sgbm(p, img, img)
    p["maxDisparity"] = 600
    with pytest.raises(ValueError):
        oracle.sgbm(p, img, img)
