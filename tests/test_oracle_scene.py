"""Pins of the oracle's rectification remap, disparity -> depth and colour map
(oracle/scene_oracle.cpp) by closed forms, known answers and independent restatements, as
the reference's own tests do (tests/test_rectification.cpp:189-265, tests/test_sgbm.cpp:320-349).
CPU only."""
import numpy as np
import pytest

import oracle

rng = np.random.default_rng(1234)


def _grid(H, W):
    ys, xs = np.mgrid[0:H, 0:W]
    return xs.astype(np.float32), ys.astype(np.float32)


@pytest.mark.parametrize("C", [1, 3])
def test_remap_identity_map_is_exact(C):
    """tests/test_rectification.cpp:189-224: an identity map reproduces the source exactly."""
    H, W = 48, 64
    src = rng.integers(0, 256, (H, W) if C == 1 else (H, W, C), dtype=np.uint8)
    mx, my = _grid(H, W)
    assert np.array_equal(oracle.remap(src, mx, my), src)


def _half_shift_expected(src):
    """Closed form of a (+0.5, +0.5) map: every weight is 1/4, so the value is
    floor(s/4 + 1/2) = (s + 2) // 4 for the sum s of the taps inside the image."""
    H, W = src.shape
    p = np.zeros((H + 1, W + 1), np.int64)
    p[:H, :W] = src
    s = p[:H, :W] + p[1:, :W] + p[:H, 1:] + p[1:, 1:]
    return ((s + 2) // 4).astype(np.uint8)


def test_remap_half_pixel_shift_closed_form():
    """tests/test_rectification.cpp:226-265: gradient image, map shifted by half a pixel."""
    H, W = 100, 100
    src = np.tile((np.arange(W) * 255 // W).astype(np.uint8), (H, 1))
    mx, my = _grid(H, W)
    got = oracle.remap(src, mx + 0.5, my + 0.5)
    assert np.array_equal(got, _half_shift_expected(src))
    noise = rng.integers(0, 256, (37, 53), dtype=np.uint8)
    mx, my = _grid(37, 53)
    assert np.array_equal(oracle.remap(noise, mx + 0.5, my + 0.5), _half_shift_expected(noise))


def test_remap_integer_shift_and_outside():
    H, W = 20, 30
    src = rng.integers(1, 256, (H, W), dtype=np.uint8)
    mx, my = _grid(H, W)
    got = oracle.remap(src, mx + 3, my - 2)
    want = np.zeros_like(src)
    want[2:, : W - 3] = src[:-2, 3:]
    assert np.array_equal(got, want)
    # Entirely outside (beyond one pixel of the border), NaN and huge coordinates give 0.
    for v in (-1.5, W + 0.0, np.nan, 3e9, -3e9, np.inf):
        assert not oracle.remap(src, np.full_like(mx, v), my).any(), v
    # Half outside: x = -0.5 has only its right taps inside, weighted 1/2.
    got = oracle.remap(src, np.full_like(mx, -0.5), my)
    assert np.array_equal(got, ((src[:, :1].astype(np.int64) + 1) // 2).repeat(W, 1).astype(np.uint8))


def _bilinear_f64(src, mx, my):
    H, W = src.shape
    x0 = np.floor(mx.astype(np.float64)).astype(np.int64)
    y0 = np.floor(my.astype(np.float64)).astype(np.int64)
    dx = mx.astype(np.float64) - x0
    dy = my.astype(np.float64) - y0
    acc = np.zeros(src.shape, np.float64)
    for ox, oy, w in ((0, 0, (1 - dx) * (1 - dy)), (0, 1, (1 - dx) * dy), (1, 0, dx * (1 - dy)),
                      (1, 1, dx * dy)):
        xx, yy = x0 + ox, y0 + oy
        ok = (xx >= 0) & (xx < W) & (yy >= 0) & (yy < H)
        acc += np.where(ok, w * src[np.clip(yy, 0, H - 1), np.clip(xx, 0, W - 1)], 0.0)
    return np.floor(acc + 0.5)


def test_remap_random_maps_within_one_level_of_float64():
    """Single-precision rounding can move a value across a .5 boundary, nothing more."""
    H, W = 64, 96
    src = rng.integers(0, 256, (H, W), dtype=np.uint8)
    mx = rng.uniform(-2, W + 1, (H, W)).astype(np.float32)
    my = rng.uniform(-2, H + 1, (H, W)).astype(np.float32)
    got = oracle.remap(src, mx, my).astype(np.int64)
    want = _bilinear_f64(src, mx, my)
    diff = np.abs(got - want)
    assert diff.max() <= 1
    assert np.count_nonzero(diff) <= 0.01 * diff.size


def test_remap_channels_share_taps():
    H, W = 16, 24
    src = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    mx = rng.uniform(-1, W, (H, W)).astype(np.float32)
    my = rng.uniform(-1, H, (H, W)).astype(np.float32)
    got = oracle.remap(src, mx, my)
    for c in range(3):
        assert np.array_equal(got[..., c], oracle.remap(np.ascontiguousarray(src[..., c]), mx, my))


def test_depth_known_answer_and_invalid():
    """tests/test_sgbm.cpp:320-332: f = 1000 px, B = 0.1 m, 100 px of disparity -> 1 m."""
    d = np.array([[1600, 0, -1, 16]], np.int16)
    z = oracle.disparity_to_depth(d, 1000.0, 0.1, 0.0)
    assert z[0, 0] == np.float32(1.0)
    assert z[0, 1] == 0 and z[0, 2] == 0
    assert z[0, 3] == np.float32(np.float32(1000.0) * np.float32(0.1))
    # A denominator <= 0 (negative offset) is invalid.
    z = oracle.disparity_to_depth(np.array([[16, 32]], np.int16), 1000.0, 0.1, -1.0)
    assert z[0, 0] == 0 and z[0, 1] == np.float32(100.0)


def test_depth_within_an_ulp_of_the_exact_quotient():
    d = rng.integers(-50, 32767, (50, 70)).astype(np.int16)
    f, B, off = 4161.221, 0.176252, 209.059
    z = oracle.disparity_to_depth(d, f, B, off).astype(np.float64)
    f32, B32, o32 = (np.float64(np.float32(v)) for v in (f, B, off))
    denom = np.float64(np.float32(d.astype(np.float64) / 16.0 + o32))
    exact = np.where(d > 0, f32 * B32 / denom, 0.0)
    rel = np.abs(z - exact) / np.maximum(exact, 1e-30)
    assert rel.max() <= 1.2e-7


def test_depth_host_form_is_a_double_quotient():
    """src/depth_map.cu:142-160: f and B are double; the quotient is rounded to float once."""
    d = rng.integers(-50, 32767, (40, 60)).astype(np.int16)
    f, B, off = 4161.221, 0.176252, 209.059
    z = oracle.disparity_to_depth_host(d, f, B, off)
    denom = (d.astype(np.float32) / np.float32(16.0) + np.float32(off)).astype(np.float64)
    want = np.where(d > 0, (f * B / denom), 0.0).astype(np.float32)
    assert np.array_equal(z, want)


def test_colour_table_is_the_reference_table():
    fx = np.load(__import__("pathlib").Path(__file__).parent / "golden" / "turbo_table.npz")
    t = oracle.turbo_table()
    assert np.array_equal(t[:, 0], fx["r"]) and np.array_equal(t[:, 1], fx["g"])
    assert np.array_equal(t[:, 2], fx["b"])


def _colour_index_np(z, lo, hi):
    z = z.astype(np.float32)
    n = (z - np.float32(lo)) / (np.float32(hi) - np.float32(lo))
    n = np.fmax(np.float32(0), np.fmin(np.float32(1), n))
    idx = ((np.float32(1) - n) * np.float32(255)).astype(np.int32)
    bad = (z <= 0) | (z > np.float32(hi) * np.float32(2))
    return np.where(bad, 0, idx)


def test_colorize_known_indices_and_invalid():
    t = oracle.turbo_table()
    z = np.array([[1.0, 10.0, 5.5, 0.0, -3.0, 20.5, np.nan, 0.5, 15.0]], np.float32)
    got = oracle.colorize_depth(z, 1.0, 10.0)
    # min depth -> 255, max depth -> 0, midpoint -> (int)127.5 = 127; invalid and NaN -> 0;
    # below min clamps to 255, between max and 2*max clamps to 0.
    idx = [255, 0, 127, 0, 0, 0, 0, 255, 0]
    assert np.array_equal(got[0], t[idx])


def test_colorize_matches_numpy_restatement():
    t = oracle.turbo_table()
    z = rng.uniform(-1, 25, (60, 80)).astype(np.float32)
    z[::7, ::5] = np.nan
    got = oracle.colorize_depth(z, 0.5, 10.0)
    assert np.array_equal(got, t[_colour_index_np(z, 0.5, 10.0)])
