"""bench.py's one-line JSON contract: the fields a reader of the benchmark relies on, and the
refusal of more than one GPU."""
import json
import subprocess
import sys

import pytest

from conftest import ROOT


def _run(*args, timeout):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=timeout)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2", timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


@pytest.mark.gpu
def test_json_line(gpu_available):
    r = _run("--workload", "480p-d64-sad-4", "--steps", "3", "--warmup", "1", "--skip-forms",
             "--skip-scene", timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
              "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config",
              "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 3 and d["warmup"] == 1
    assert d["unit"] == "MDE/s" and d["higher_is_better"] is True
    assert d["value"] > 0 and d["ms_per_step"] > 0
    # value is W*H*D per second over the timed steps, so it and ms_per_step agree.
    de = 640 * 480 * 64
    assert d["value"] == pytest.approx(de / (d["ms_per_step"] * 1e-3) / 1e6, rel=0.01)
    assert "workload" in d["config"] and "model" not in d["config"]
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert rf["achieved"] > 0 and rf["frac"] == pytest.approx(rf["achieved"] / rf["peak"], rel=1e-3)
    cb = d["cpu_baseline"]
    assert cb["kind"] == "port" and cb["cores"] >= 1 and cb["value"] > 0
    assert d["parity_vs_oracle_mismatched_pixels"] == 0
