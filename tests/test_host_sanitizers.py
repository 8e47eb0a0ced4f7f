"""Host code under the sanitizers (CPU only): the file readers (config.cpp, evaluate.cpp) on
random inputs under AddressSanitizer + UBSan, and the serving loop's pool/admission core
(serve_core.h, host_copy.cpp) under AddressSanitizer + UBSan and under ThreadSanitizer."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "lumenstereo_amd" / "csrc"

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


def _build_and_run(tmp_path, name, sources, flags, includes):
    exe = tmp_path / name
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-pthread", "-fno-omit-frame-pointer", *flags,
           *[f"-I{i}" for i in includes], *map(str, sources), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True, timeout=240)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    assert "WARNING: ThreadSanitizer" not in r.stderr, r.stderr[-4000:]


def test_file_readers_under_asan_ubsan(tmp_path):
    _build_and_run(tmp_path, "fuzz_readers",
                   [ROOT / "tests" / "cpp" / "fuzz_file_readers.cpp", CSRC / "config.cpp",
                    CSRC / "evaluate.cpp"],
                   ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"],
                   [ROOT / "include"])


@pytest.mark.parametrize("san", ["address,undefined", "thread"])
def test_serve_core_under_sanitizer(tmp_path, san):
    _build_and_run(tmp_path, "serve_core_" + san.split(",")[0],
                   [ROOT / "tests" / "cpp" / "test_serve_core.cpp", CSRC / "host_copy.cpp"],
                   [f"-fsanitize={san}"], [CSRC])
