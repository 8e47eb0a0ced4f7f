"""Runs the C++ API test programs, built by __graft_entry__.build(): tests/cpp/test_stereo_api.cpp
(the reference's C++ tests) and tests/cpp/test_cvmat_overloads.cpp (the cv::Mat overloads against
a cv::Mat test double); host-only cases here, GPU cases under the gpu marker."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "test_stereo_api"
CVMAT_BIN = Path(__file__).resolve().parent / "cpp" / "test_cvmat_overloads"


def _run(args, binary=BIN):
    if not binary.exists():
        pytest.fail(f"{binary} is not built; run __graft_entry__.build()")
    r = subprocess.run([str(binary), *args], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_cpp_api_host_cases():
    _run([])


@pytest.mark.gpu
def test_cpp_api_gpu_cases(gpu_available):
    _run(["--gpu"])


def test_cvmat_overloads_host_cases():
    _run([], CVMAT_BIN)


@pytest.mark.gpu
def test_cvmat_overloads_gpu_cases(gpu_available):
    _run(["--gpu"], CVMAT_BIN)
