"""Runs the host-only test program of the serving loop's device pool and admission logic
(tests/cpp/test_serve_core.cpp, an injected fake allocator in place of hipMalloc), and of the
staging copy pool (copies and BGR -> gray)."""
import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "cpp" / "test_serve_core"


def test_serve_core_host_logic():
    if not BIN.exists():
        pytest.fail(f"{BIN} is not built; run __graft_entry__.build()")
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok ") == 9, r.stdout
