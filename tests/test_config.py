"""Parameter files and Middlebury calibration (host code, CPU only): mirrors the reference's
ConfigTest and CalibrationLoadTest cases (tests/test_rectification.cpp:58-145) and loads the
reference's own data files, kept as fixtures: tests/golden/reference_stereo_params.yaml is
config/stereo_params.yaml and tests/golden/adirondack_calib.txt is
dataset/Adirondack-perfect/calib.txt of Pfactorial01/LumenStereo."""
from pathlib import Path

import numpy as np
import pytest

import lumenstereo_amd as ls
from lumenstereo_amd import config
from lumenstereo_amd.scene import CameraParams

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_stereo_params_save_load(tmp_path):
    """tests/test_rectification.cpp:90-111."""
    original = ls.StereoParams(minDisparity=10, maxDisparity=200, blockSize=7, P1=15, P2=60)
    path = tmp_path / "test_stereo_params.yaml"
    assert config.saveStereoParams(path, original)
    loaded = ls.StereoParams()
    assert config.loadStereoParams(path, loaded)
    for f in ("minDisparity", "maxDisparity", "blockSize", "P1", "P2"):
        assert getattr(loaded, f) == getattr(original, f)


def test_matching_cost_roundtrip(tmp_path):
    """tests/test_rectification.cpp:113-122."""
    original = ls.StereoParams(matchingCostMode=ls.MatchingCostMode.Census)
    path = tmp_path / "census.yaml"
    assert config.saveStereoParams(path, original)
    loaded = ls.StereoParams()
    assert config.loadStereoParams(path, loaded)
    assert loaded.matchingCostMode == ls.MatchingCostMode.Census


def test_saved_text_is_the_reference_key_order(tmp_path):
    path = tmp_path / "p.yaml"
    assert config.saveStereoParams(path, ls.StereoParams(useGPU=False))
    keys = [ln.split(":")[0] for ln in path.read_text().splitlines()]
    assert keys == ["minDisparity", "maxDisparity", "blockSize", "P1", "P2", "preFilterCap",
                    "uniquenessRatio", "disp12MaxDiff", "useGPU", "numDirections",
                    "matchingCostMode"]
    assert "useGPU: false" in path.read_text() and "matchingCostMode: SAD" in path.read_text()


def test_reference_config_file():
    """config/stereo_params.yaml: every key it sets is read; speckle keys are not (the
    reference's loader never reads them, src/config.cpp:21-38)."""
    p = ls.StereoParams(speckleWindowSize=77, speckleRange=5)
    assert config.loadStereoParams(GOLDEN / "reference_stereo_params.yaml", p)
    assert (p.minDisparity, p.maxDisparity, p.blockSize, p.P1, p.P2) == (0, 128, 5, 200, 800)
    assert (p.preFilterCap, p.uniquenessRatio, p.disp12MaxDiff) == (31, 5, 2)
    assert p.useGPU is True and p.numDirections == 8
    assert p.matchingCostMode == ls.MatchingCostMode.SAD
    assert (p.speckleWindowSize, p.speckleRange) == (77, 5)


@pytest.mark.parametrize("text,ok,check", [
    ("maxDisparity: 0x40\n", True, lambda p: p.maxDisparity == 64),
    ("maxDisparity: 0o100\n", True, lambda p: p.maxDisparity == 64),
    ("maxDisparity: 0100\n", True, lambda p: p.maxDisparity == 64),   # stream base detection
    ("maxDisparity: '96'  # quoted\n", True, lambda p: p.maxDisparity == 96),
    ("---\nuseGPU: off\nmatchingCostMode: census\n", True,
     lambda p: not p.useGPU and p.matchingCostMode == ls.MatchingCostMode.Census),
    ("matchingCostMode: \"ncc\"\n", True, lambda p: p.matchingCostMode == ls.MatchingCostMode.SAD),
    ("maxDisparity: 12abc\n", False, lambda p: p.maxDisparity == 128),  # not an int: unchanged
    ("maxDisparity:\n  a: 1\n", False, lambda p: True),                 # nested node
    ("maxDisparity: [1, 2]\n", False, lambda p: True),                  # flow sequence
    ("useGPU: maybe\n", False, lambda p: True),
    ("this is not yaml\n", False, lambda p: True),
    ("blockSize: 4\n", False, lambda p: p.blockSize == 4),  # read, but isValid() fails
])
def test_yaml_scalar_conversions(tmp_path, text, ok, check):
    path = tmp_path / "p.yaml"
    path.write_text(text)
    p = ls.StereoParams()
    assert config.loadStereoParams(path, p) is ok
    assert check(p)


def test_missing_params_file(tmp_path):
    assert not config.loadStereoParams(tmp_path / "nope.yaml", ls.StereoParams())


def test_middlebury_calib_adirondack():
    """tests/test_rectification.cpp:58-79 on the reference's calib.txt."""
    cp = CameraParams()
    assert config.loadMiddleburyCalib(GOLDEN / "adirondack_calib.txt", cp)
    assert (cp.width, cp.height) == (2880, 1988)
    assert cp.focalLength == 4161.221
    assert cp.baseline == pytest.approx(0.176252, abs=1e-12)
    assert cp.disparityOffset == 209 and cp.numDisparities == 280
    assert np.array_equal(cp.K0, [[4161.221, 0, 1445.577], [0, 4161.221, 984.686], [0, 0, 1]])
    assert np.array_equal(cp.K1, [[4161.221, 0, 1654.636], [0, 4161.221, 984.686], [0, 0, 1]])
    assert np.array_equal(cp.R, np.eye(3)) and not cp.D0.any() and not cp.D1.any()
    assert cp.T[0, 0] == -176.252 / 1000.0 and cp.T[1, 0] == 0 and cp.T[2, 0] == 0
    assert cp.isValid()
    cp.computeDerivedValues()  # |T| and (fx + fy) / 2 agree with what the loader set
    assert cp.baseline == pytest.approx(0.176252) and cp.focalLength == 4161.221
    assert ls.middleburyMatcherMaxDisparity(cp.numDisparities) == 280


def test_middlebury_calib_missing_and_defaults(tmp_path):
    """tests/test_rectification.cpp:81-85, plus ndisp absent -> 280 and a missing size."""
    assert not config.loadMiddleburyCalib("/nonexistent/path/calib.txt", CameraParams())
    f = tmp_path / "calib.txt"
    f.write_text("cam0=[100 0 50; 0 100 40; 0 0 1]\ndoffs=3.5\nbaseline=50\nwidth=64\nheight=48\n")
    cp = CameraParams()
    assert config.loadMiddleburyCalib(f, cp)
    assert cp.numDisparities == 280 and cp.disparityOffset == 4 and cp.focalLength == 100
    assert np.array_equal(cp.K1, np.eye(3))
    f.write_text("cam0=[100 0 50; 0 100 40; 0 0 1]\n")
    assert not config.loadMiddleburyCalib(f, CameraParams())
    f.write_text("width=abc\nheight=2\n")
    assert not config.loadMiddleburyCalib(f, CameraParams())
