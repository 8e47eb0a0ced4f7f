"""The tuned-schedule table (host side, no GPU work): parsing, validation with the library's
own rules, bucket lookup by width, and the defaults. The table only chooses schedules, which
never change results (tests/test_gpu_tune.py checks that on the GPU)."""
import pytest

import lumenstereo_amd as ls
from lumenstereo_amd import tune

TABLE = ls.LIB_PATH.parent / "tuning.tsv"


@pytest.fixture(autouse=True)
def restore_table():
    yield
    if TABLE.exists():
        ls.tuning_load(str(TABLE))
    else:
        ls.tuning_clear()


def _handle(D, census, paths):
    return ls.StereoSGBM(tune.params_for(D, census, paths))


@pytest.mark.parametrize("D,K", [(64, 4), (128, 8), (192, 12), (290, 20), (512, 32)])
def test_defaults_without_table(D, K):
    ls.tuning_clear()
    s = _handle(D, True, 8).getSchedule(640, 480)
    assert s == tune.default_schedule(K)


def test_bucket_lookup(tmp_path):
    f = tmp_path / "t.tsv"
    f.write_text("# comment\n\n640 128 1 8 16 2 128   # trailing comment\n"
                 "1920 128 1 8 4 0 512\n640 128 0 8 4 2 256\n")
    assert ls.tuning_load(str(f)) == 3
    m = _handle(128, True, 8)
    assert m.getSchedule(320, 240)["agg_prefetch"] == 16      # smallest bucket >= width
    assert m.getSchedule(640, 480)["census_tile"] == 128
    assert m.getSchedule(641, 480)["census_tile"] == 512      # next bucket up
    assert m.getSchedule(4000, 100)["agg_priority"] == 0      # wider than every bucket
    assert _handle(100, True, 8).getSchedule(640, 480)["agg_prefetch"] == 16  # D=100 pads to 128
    assert _handle(128, False, 8).getSchedule(640, 480)["agg_prefetch"] == 4  # SAD entry
    assert _handle(128, True, 4).getSchedule(640, 480) == tune.default_schedule(8)  # no entry


def test_override_beats_table(tmp_path):
    f = tmp_path / "t.tsv"
    f.write_text("640 128 1 8 16 2 128\n")
    ls.tuning_load(str(f))
    m = _handle(128, True, 8)
    m.setSchedule({"agg_prefetch": 4, "agg_priority": 0, "census_tile": 512, "agg_wide": 2})
    assert m.getSchedule(640, 480) == {"agg_prefetch": 4, "agg_priority": 0, "census_tile": 512,
                                       "agg_wide": 2, "agg_occupancy": 0}
    m.setSchedule(None)
    assert m.getSchedule(640, 480)["agg_prefetch"] == 16


@pytest.mark.parametrize("line,msg", [
    ("640 128 1 8 16 2", "expected 7 to 9 integers"),
    ("640 128 1 8 16 2 256 1 2 5", "expected 7 to 9 integers"),
    ("640 128 1 8 16 2 256 x", "expected 7 to 9 integers"),
    ("640 128 1 8 16 2 256 1 x", "expected 7 to 9 integers"),
    ("640 128 1 8 16 2 256 7", "lane setting"),
    ("640 192 1 8 8 2 256 4", "multiple of 128"),
    ("640 128 1 8 16 2 256 1 1", "occupancy"),
    ("640 128 1 8 16 2 256 1 5", "occupancy"),
    ("640 100 1 8 8 1 256", "bad width or padded range"),
    ("640 128 2 8 8 1 256", "bad cost mode"),
    ("640 128 1 6 8 1 256", "bad cost mode or path count"),
    ("640 128 1 8 8 3 256", "priority"),
    ("640 128 1 8 8 1 300", "census tile"),
    ("640 320 1 4 16 1 256", "not supported"),
])
def test_malformed_tables_are_rejected(tmp_path, line, msg):
    f = tmp_path / "t.tsv"
    f.write_text("640 64 0 4 8 1 256\n" + line + "\n")
    ls.tuning_clear()
    with pytest.raises(ls.StereoException, match=msg):
        ls.tuning_load(str(f))
    # the table is unchanged by a failed load
    assert _handle(64, False, 4).getSchedule(640, 480) == tune.default_schedule(4)


def test_prefetch_rule_matches_library(tmp_path):
    """The tuner's candidate filter agrees with the library's validation for every slice."""
    for K in range(4, 33, 4):
        for pf in (4, 8, 16):
            f = tmp_path / f"k{K}_{pf}.tsv"
            f.write_text(f"640 {K * 16} 1 8 {pf} 1 256\n")
            try:
                ls.tuning_load(str(f))
                ok = True
            except ls.StereoException:
                ok = False
            assert ok == tune.valid_prefetch(K, pf), (K, pf)
    for K in range(4, 33, 4):
        assert tune.valid_prefetch(K, tune.default_schedule(K)["agg_prefetch"])


def test_missing_file_is_an_error():
    with pytest.raises(ls.StereoException, match="cannot open"):
        ls.tuning_load("/nonexistent/tuning.tsv")


def test_lane_setting_column(tmp_path):
    """The optional eighth column is the aggregation lane setting (0 when absent); every padded
    range takes every 32-lane setting with every prefetch depth its 16-lane scanlines take, and
    the 64-lane ones where it is a multiple of 128."""
    f = tmp_path / "t.tsv"
    f.write_text("640 128 1 8 16 2 128 3\n640 64 1 8 16 2 128\n640 192 1 8 8 1 256 2\n")
    assert ls.tuning_load(str(f)) == 3
    assert _handle(128, True, 8).getSchedule(640, 480)["agg_wide"] == 3
    assert _handle(64, True, 8).getSchedule(640, 480)["agg_wide"] == 0
    assert _handle(192, True, 8).getSchedule(640, 480)["agg_wide"] == 2
    for Dp in range(64, 513, 64):
        for pf in (4, 8, 16):
            g = tmp_path / f"w{Dp}_{pf}.tsv"
            g.write_text(f"640 {Dp} 1 4 {pf} 1 256 3\n")
            try:
                ls.tuning_load(str(g))
                ok = True
            except ls.StereoException:
                ok = False
            assert ok == tune.valid_prefetch(Dp // 16, pf), (Dp, pf)
            # 64-lane settings only where a lane holds whole packed pairs (Dp = 128 k).
            g.write_text(f"640 {Dp} 1 4 {pf} 1 256 {4 + Dp % 3}\n")
            try:
                ls.tuning_load(str(g))
                ok = True
            except ls.StereoException:
                ok = False
            assert ok == (Dp % 128 == 0 and tune.valid_prefetch(Dp // 16, pf)), (Dp, pf)
            assert (4 + Dp % 3 in tune.wide_settings(Dp // 16)) == (Dp % 128 == 0), Dp


def test_occupancy_column(tmp_path):
    """The optional ninth column caps the aggregation's blocks per CU (0 when absent)."""
    f = tmp_path / "t.tsv"
    f.write_text("640 128 1 8 16 2 128 3 2\n640 64 1 8 16 2 128 1\n640 192 1 8 8 1 256 2 4\n"
                 "640 256 1 8 4 1 256 0 3\n")
    assert ls.tuning_load(str(f)) == 4
    assert _handle(128, True, 8).getSchedule(640, 480)["agg_occupancy"] == 2
    assert _handle(64, True, 8).getSchedule(640, 480)["agg_occupancy"] == 0
    assert _handle(192, True, 8).getSchedule(640, 480)["agg_occupancy"] == 4
    assert _handle(256, True, 8).getSchedule(640, 480) == {
        "agg_prefetch": 4, "agg_priority": 1, "census_tile": 256, "agg_wide": 0,
        "agg_occupancy": 3}


def test_range_of_96(tmp_path):
    """Ranges of 65 to 96 pad to 96 (K = 6): a table key of 96 is accepted with 16-lane
    scanlines and every prefetch depth, and refused with 32-lane ones (three disparities per
    lane of 32 is not a whole number of packed pairs)."""
    assert [tune.padded_range(d) for d in (64, 65, 80, 96, 97, 128, 290)] == [64, 96, 96, 96, 128, 128, 320]
    for pf in (4, 8, 16):
        f = tmp_path / f"r96_{pf}.tsv"
        f.write_text(f"640 96 1 8 {pf} 1 256 0\n")
        assert ls.tuning_load(str(f)) == 1
        assert tune.valid_prefetch(6, pf)
    assert _handle(80, True, 8).getSchedule(640, 480)["agg_prefetch"] == 16
    g = tmp_path / "r96_wide.tsv"
    g.write_text("640 96 1 8 8 1 256 1\n")
    with pytest.raises(ls.StereoException, match="32-lane"):
        ls.tuning_load(str(g))
    assert tune.wide_settings(6) == (0,)
    h = tmp_path / "r160.tsv"
    h.write_text("640 160 1 8 8 1 256\n")
    with pytest.raises(ls.StereoException, match="padded range"):
        ls.tuning_load(str(h))
