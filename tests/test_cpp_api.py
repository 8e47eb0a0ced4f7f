"""Runs the C++ API test program (tests/cpp/test_stereo_api.cpp), built by
__graft_entry__.build(): host-only cases here, GPU cases under the gpu marker."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "test_stereo_api"


def _run(args):
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built; run __graft_entry__.build()")
    r = subprocess.run([str(BIN), *args], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_api_host_cases():
    _run([])


@pytest.mark.gpu
def test_cpp_api_gpu_cases(gpu_available):
    _run(["--gpu"])
