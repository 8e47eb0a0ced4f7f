"""The oracle against the committed golden fixtures (tests/golden/golden_*.npz).

The fixtures come from tests/golden/make_golden.py, an independent numpy restatement of the
reference's semantics; every stage must agree bit for bit.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle

FIXTURES = sorted((Path(__file__).parent / "golden").glob("golden_*.npz"))
STAGES = ["cost", "paths", "sum", "wta_int", "disp_left", "disp_right", "after_lr",
          "after_speckle"]


def test_fixtures_present():
    assert len(FIXTURES) >= 5


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.stem)
def test_oracle_matches_fixture(path):
    z = np.load(path)
    p = json.loads(str(z["params"]))
    out, st = oracle.sgbm(p, z["left"], z["right"], dumps=True)
    for k in STAGES:
        np.testing.assert_array_equal(st[k], z[k], err_msg=f"{path.stem}: {k}")
    np.testing.assert_array_equal(out, z["output"], err_msg=f"{path.stem}: output")


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: p.stem)
def test_oracle_stage_functions_match_fixture(path):
    """Each stage function of the oracle on its own, fed with the fixture's previous stage."""
    z = np.load(path)
    p = json.loads(str(z["params"]))
    D = p["maxDisparity"] - p["minDisparity"]
    m = p["minDisparity"]
    if p["matchingCostMode"] == 1:
        cost = oracle.cost_census(oracle.census(z["left"]), oracle.census(z["right"]), D, m)
    else:
        cost = oracle.cost_sad(oracle.prefilter(z["left"], p["preFilterCap"]),
                               oracle.prefilter(z["right"], p["preFilterCap"]), D, m,
                               p["blockSize"])
    np.testing.assert_array_equal(cost, z["cost"])
    cs = oracle.cost_scale(p["matchingCostMode"], p["blockSize"], p["preFilterCap"])
    for r in range(z["paths"].shape[0]):
        L = oracle.aggregate_path(z["cost"], z["left"], p["P1"], p["P2"], cs, r)
        np.testing.assert_array_equal(L, z["paths"][r], err_msg=f"path {oracle.PATH_NAMES[r]}")
    dl, wi = oracle.wta_left(z["sum"], m, p["uniquenessRatio"])
    np.testing.assert_array_equal(dl, z["disp_left"])
    np.testing.assert_array_equal(wi, z["wta_int"])
    np.testing.assert_array_equal(oracle.wta_right(z["sum"], m), z["disp_right"])
    if p["disp12MaxDiff"] >= 0:
        np.testing.assert_array_equal(
            oracle.lr_check(z["disp_left"], z["disp_right"], p["disp12MaxDiff"] * 16), z["after_lr"])
    np.testing.assert_array_equal(oracle.median(z["after_speckle"]), z["output"])
