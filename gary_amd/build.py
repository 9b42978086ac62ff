"""Builds libgary.so (HIP for gfx950 + host C++) in-tree with hipcc, and the CPU oracle.

Usage: python -m gary_amd.build [--jobs N] [--no-oracle]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libgary.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-std=c++17", "-fPIC", f"-I{ROOT / 'include'}", f"-I{CSRC}", "-Wall",
          "-Wno-unused-function", "-Wno-unused-variable"]
# Device code: gfx950 only; packed FP32 instructions disabled for every HIP translation unit.
HIP_FLAGS = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-munsafe-fp-atomics"]


def sources() -> list[Path]:
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [HIPCC, *COMMON]
    if src.suffix == ".hip":
        cmd += HIP_FLAGS
    else:
        cmd += ["-march=x86-64-v3"]
    cmd += ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build_lib(jobs: int = 8) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp),
           "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


CPP_TEST_SRC = ROOT / "tests" / "cpp" / "tri_query_test.cpp"
CPP_TEST_BIN = ROOT / "tests" / "cpp" / "tri_query_test"
CAPTURE_TEST_SRC = ROOT / "tests" / "cpp" / "graph_capture_test.cpp"
CAPTURE_TEST_BIN = ROOT / "tests" / "cpp" / "graph_capture_test"


def build_cpp_test() -> Path:
    """The façade test: plain g++ host code linked against libgary.so and the oracle; and the
    graph-capture test: g++ host code against libgary.so and the HIP runtime API."""
    cmd = ["g++", "-std=c++17", "-O2", f"-I{ROOT / 'include'}", str(CPP_TEST_SRC),
           f"-L{PKG}", "-lgary", f"-L{ROOT / 'oracle'}", "-loracle",
           "-Wl,-rpath,$ORIGIN/../../gary_amd", "-Wl,-rpath,$ORIGIN/../../oracle",
           "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(CPP_TEST_BIN)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"C++ façade test build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    cmd = ["g++", "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", f"-I{ROOT / 'include'}",
           "-I/opt/rocm/include", str(CAPTURE_TEST_SRC), f"-L{PKG}", "-lgary",
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,$ORIGIN/../../gary_amd",
           "-Wl,-rpath,/opt/rocm/lib", "-o", str(CAPTURE_TEST_BIN)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"graph-capture test build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return CPP_TEST_BIN


CONVERT_SRC = PKG / "cli" / "convert.cpp"
CONVERT_BIN = PKG / "gary_convert"


def build_convert() -> Path:
    """The graph conversion tool (host code, g++) linked against libgary.so."""
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(CONVERT_SRC),
           f"-L{PKG}", "-lgary", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath-link,/opt/rocm/lib",
           "-o", str(CONVERT_BIN)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"gary_convert build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return CONVERT_BIN


def build_oracle() -> Path:
    r = subprocess.run(["make", "-B", "-C", str(ROOT / "oracle")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}\n{r.stderr}")
    return ROOT / "oracle" / "liboracle.so"


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args(argv)
    print(build_lib(a.jobs))
    print(build_convert())
    if not a.no_oracle:
        print(build_oracle())
        print(build_cpp_test())
    return 0


if __name__ == "__main__":
    sys.exit(main())
