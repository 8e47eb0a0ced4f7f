"""Build recipe for liblumen_sgbm.so (gfx950) and the oracle's CPU library.

Run ``python -m lumenstereo_amd._build`` (or ``__graft_entry__.build()``). Everything is
compiled in-tree so the shared objects travel with the source tree.
"""
from __future__ import annotations

import concurrent.futures
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "lumenstereo_amd"
CSRC = PKG / "csrc"
LIB = PKG / "liblumen_sgbm.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["sgbm_kernels.hip", "sgbm_paths.hip", "sgbm_paths_k4.hip", "sgbm_paths_k12.hip",
               "sgbm_paths_k20.hip", "sgbm_paths_k28.hip", "sgbm_capi.hip", "sgbm_batch.hip",
               "sgbm_serve.hip", "sgbm_tune.hip", "rectify.hip", "depth_map.hip", "device_mem.hip"]
CXX_SOURCES = ["synth_scene.cpp", "host_copy.cpp", "config.cpp", "evaluate.cpp"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd: list[str]) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True)


def build_library() -> Path:
    inc = ["-I", str(ROOT / "include"), "-I", str(CSRC)]
    objdir = PKG / "_objs"
    objdir.mkdir(exist_ok=True)
    objs = [objdir / (Path(src).stem + ".o") for src in HIP_SOURCES]
    # The translation units are independent: compile them in parallel (bounded by the CPUs).
    jobs = max(1, min(len(HIP_SOURCES), os.cpu_count() or 1, 16))
    with concurrent.futures.ThreadPoolExecutor(jobs) as pool:
        list(pool.map(lambda so: _run([HIPCC, *HIP_FLAGS, *inc, "-c", CSRC / so[0], "-o", so[1]]),
                      zip(HIP_SOURCES, objs)))
    for src in CXX_SOURCES:
        obj = objdir / (Path(src).stem + ".o")
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", *inc, "-c",
              CSRC / src, "-o", obj])
        objs.append(obj)
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-pthread", "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_oracle() -> Path:
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-pthread",
          ORACLE_DIR / "sgbm_oracle.cpp", ORACLE_DIR / "scene_oracle.cpp", "-o", tmp])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


CPP_TEST_SRC = ROOT / "tests" / "cpp" / "test_stereo_api.cpp"
CPP_TEST_BIN = ROOT / "tests" / "cpp" / "test_stereo_api"


def build_cpp_tests() -> Path:
    """The C++ API test program (header-only stereo::StereoSGBM over liblumen_sgbm.so)."""
    tmp = CPP_TEST_BIN.with_suffix(".tmp")
    _run(["g++", "-O1", "-std=c++17", "-I", ROOT / "include", CPP_TEST_SRC, "-L", PKG,
          "-llumen_sgbm", "-Wl,-rpath,$ORIGIN/../../lumenstereo_amd", "-o", tmp])
    os.replace(tmp, CPP_TEST_BIN)
    return CPP_TEST_BIN


CVMAT_TEST_SRC = ROOT / "tests" / "cpp" / "test_cvmat_overloads.cpp"
CVMAT_TEST_BIN = ROOT / "tests" / "cpp" / "test_cvmat_overloads"


def build_cvmat_test() -> Path:
    """The cv::Mat overloads (LUMEN_WITH_OPENCV) against the cv::Mat test double of
    tests/cpp/mock_opencv: OpenCV itself is not installed."""
    tmp = CVMAT_TEST_BIN.with_suffix(".tmp")
    _run(["g++", "-O1", "-std=c++17", "-I", ROOT / "include", "-I",
          ROOT / "tests" / "cpp" / "mock_opencv", CVMAT_TEST_SRC, "-L", PKG, "-llumen_sgbm",
          "-Wl,-rpath,$ORIGIN/../../lumenstereo_amd", "-o", tmp])
    os.replace(tmp, CVMAT_TEST_BIN)
    return CVMAT_TEST_BIN


SERVE_TEST_SRC = ROOT / "tests" / "cpp" / "test_serve_core.cpp"
SERVE_TEST_BIN = ROOT / "tests" / "cpp" / "test_serve_core"


def build_serve_core_test() -> Path:
    """Host-only test program of the serving loop's pool and admission logic (no HIP)."""
    tmp = SERVE_TEST_BIN.with_suffix(".tmp")
    _run(["g++", "-O1", "-std=c++17", "-pthread", "-I", CSRC, SERVE_TEST_SRC, CSRC / "host_copy.cpp",
          "-o", tmp])
    os.replace(tmp, SERVE_TEST_BIN)
    return SERVE_TEST_BIN


def build_all() -> None:
    build_oracle()
    build_library()
    build_cpp_tests()
    build_cvmat_test()
    build_serve_core_test()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
