"""Exhaustive checks of the integer identities the SAD cost kernel divides with (no GPU needed).

The kernel never divides: a window sum s of n taps (each tap an absolute difference of two
uint8 values, so s <= 255 n) becomes floor(s / n) through a 24-bit multiply, and the quotient
is read from byte 3 of the product. These tests run every (s, n) the kernel can meet, so a
wrong constant, a dropped rounding-up or an overflow of 32 bits fails them.
"""
import numpy as np


def _byte3(v):
    return (v >> np.uint64(24)) & np.uint64(0xFF)


def test_whole_window_magic():
    # Windows whose taps are all valid: n = (2R+1) * rows with 3 <= n <= 49, and
    # magic = ceil(2^24 / n) (sad_cost_runs).
    for n in range(3, 50):
        magic = ((1 << 24) + n - 1) // n
        assert magic < (1 << 24)
        s = np.arange(0, 255 * n + 1, dtype=np.uint64)
        prod = s * np.uint64(magic)
        assert int(prod.max()) < (1 << 32)
        np.testing.assert_array_equal(_byte3(prod), s // np.uint64(n))


def _div_entry(n):
    # (mul, add) of sad_div_entry in lumenstereo_amd/csrc/sgbm_device.h.
    if n == 0:
        return 0, 0xFF000000
    if n == 1:
        return 0xFFFFFF, 0xFFFFFF
    return 2 * (((1 << 23) + n - 1) // n), 0


def test_checked_window_divisor_table():
    # Border windows: c valid columns of `rows` valid rows, n = c * rows in 0..49; an empty
    # window (s = 0) must read 255.
    seen = set()
    for rows in range(1, 8):
        for c in range(0, 8):
            n = c * rows
            if n > 49 or n in seen:
                continue
            seen.add(n)
            mul, add = _div_entry(n)
            assert mul < (1 << 24) and add < (1 << 32)
            s = np.arange(0, 255 * n + 1, dtype=np.uint64)
            prod = s * np.uint64(mul) + np.uint64(add)
            assert int(prod.max()) < (1 << 32)
            want = s // np.uint64(n) if n else np.full_like(s, 255)
            np.testing.assert_array_equal(_byte3(prod), want)
    assert {0, 1, 2, 49} <= seen
