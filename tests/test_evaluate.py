"""PFM ground truth and the evaluation metrics (host code, CPU only), checked against files
written independently here and a numpy restatement of tools/evaluate.cpp:171-221."""
import struct

import numpy as np
import pytest

from lumenstereo_amd import StereoException
from lumenstereo_amd.evaluate import evaluateDisparity, loadPFM, savePFM

rng = np.random.default_rng(5)


def _write_pfm(path, a, little=True, scale=1.0):
    """PFM as the Middlebury tools write it: header text, then the rows bottom to top."""
    h, w = a.shape
    fmt = "<f" if little else ">f"
    body = b"".join(struct.pack(fmt, float(v)) for v in a[::-1].ravel())
    path.write_bytes(f"Pf\n{w} {h}\n{-scale if little else scale}\n".encode() + body)


@pytest.mark.parametrize("little", [True, False])
def test_load_pfm_both_byte_orders(tmp_path, little):
    a = rng.uniform(0, 300, (7, 11)).astype(np.float32)
    a[2, 3] = np.inf
    f = tmp_path / "d.pfm"
    _write_pfm(f, a, little)
    got = loadPFM(f)
    assert got.shape == (7, 11) and np.array_equal(got, a)


def test_save_load_roundtrip_and_errors(tmp_path):
    a = rng.uniform(-5, 5, (4, 6)).astype(np.float32)
    f = tmp_path / "r.pfm"
    savePFM(f, a)
    assert np.array_equal(loadPFM(f), a)
    assert f.read_bytes().startswith(b"Pf\n6 4\n-1.0\n")
    with pytest.raises(StereoException):
        loadPFM(tmp_path / "missing.pfm")
    (tmp_path / "c.pfm").write_bytes(b"PF\n2 2\n-1.0\n" + b"\0" * 48)
    with pytest.raises(StereoException, match="magic"):
        loadPFM(tmp_path / "c.pfm")
    (tmp_path / "s.pfm").write_bytes(b"Pf\n4 4\n-1.0\n" + b"\0" * 20)
    with pytest.raises(StereoException, match="payload"):
        loadPFM(tmp_path / "s.pfm")


def _metrics_np(disp, gt):
    ours = disp.astype(np.float32) * np.float32(1 / 16)
    keep = ~((gt <= 0) | (gt > 1000) | np.isinf(gt)) & ~(ours <= 0)
    e = np.abs(ours[keep] - gt[keep]).astype(np.float32)
    n = e.size
    srt = np.sort(e)
    return {
        "valid": n, "valid_pct": 100.0 * n / gt.size,
        "mae": float(np.sum(e.astype(np.float64))) / n,
        "rms": float(np.sqrt(np.sum((e * e).astype(np.float64)) / n)),
        "bad1": 100.0 * np.count_nonzero(e > 1) / n, "bad2": 100.0 * np.count_nonzero(e > 2) / n,
        "bad4": 100.0 * np.count_nonzero(e > 4) / n,
        "p50": srt[n * 50 // 100], "p90": srt[n * 90 // 100], "p95": srt[n * 95 // 100],
        "p99": srt[n * 99 // 100],
    }


def test_metrics_match_restatement():
    H, W = 60, 90
    gt = rng.uniform(0.5, 120, (H, W)).astype(np.float32)
    gt[::7, ::3] = np.inf
    gt[1::9, ::4] = 0
    gt[2::11, ::5] = 1500
    disp = (gt * 16 + rng.normal(0, 40, (H, W))).clip(-1, 32767).astype(np.int16)
    disp[::5, 1::6] = -1
    disp[3::8, ::7] = 0
    got = evaluateDisparity(disp, gt)
    want = _metrics_np(disp, gt)
    assert got["valid"] == want["valid"]
    for k in ("valid_pct", "bad1", "bad2", "bad4"):
        assert got[k] == pytest.approx(want[k], rel=1e-12)
    # Sums in a different order differ only in the last bits.
    assert got["mae"] == pytest.approx(want["mae"], rel=1e-12)
    assert got["rms"] == pytest.approx(want["rms"], rel=1e-12)
    for k in ("p50", "p90", "p95", "p99"):
        assert got[k] == want[k]


def test_known_answer_and_no_valid_pixels():
    gt = np.array([[10.0, 20.0, np.inf, 5.0]], np.float32)
    disp = np.array([[16 * 11, 16 * 20 + 8, 100, 0]], np.int16)  # errors 1.0, 0.5; skip; skip
    r = evaluateDisparity(disp, gt)
    assert r["valid"] == 2 and r["mae"] == 0.75 and r["bad1"] == 0.0
    assert r["rms"] == pytest.approx(np.sqrt((1 + 0.25) / 2))
    assert r["p50"] == 1.0 and r["valid_pct"] == 50.0
    with pytest.raises(StereoException, match="No valid pixels"):
        evaluateDisparity(np.zeros((2, 2), np.int16), np.ones((2, 2), np.float32))
