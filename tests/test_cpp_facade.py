"""The C++ façade (include/*.h) used from plain g++ code, as a reference user would."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "tests" / "cpp" / "tri_query_test"


@pytest.mark.parametrize("hdr", ["gary.h", "Defines.h", "DeviceGraph.h", "ops/Search.h",
                                 "ops/Intersection.h", "kernels/GraphQueryHom.h"])
def test_headers_compile_standalone(tmp_path, hdr):
    src = tmp_path / "t.cpp"
    src.write_text(f'#include "{hdr}"\nint main() {{ return 0; }}\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        f"-I{ROOT / 'include'}", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    if hdr == "gary.h":  # the C ABI header is plain C as well
        csrc = tmp_path / "t.c"
        csrc.write_text('#include "gary.h"\nint main(void) { return 0; }\n')
        r = subprocess.run(["gcc", "-std=c11", "-fsyntax-only", "-Wall", "-Werror",
                            f"-I{ROOT / 'include'}", str(csrc)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_facade_binary_built():
    assert BIN.exists(), "tests/cpp/tri_query_test is built by __graft_entry__.build()"


@pytest.mark.gpu
@pytest.mark.parametrize("scale", [0, 16])
def test_facade_tri_query(scale):
    r = subprocess.run([str(BIN), str(scale)], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout


CAPTURE = ROOT / "tests" / "cpp" / "graph_capture_test"


@pytest.mark.gpu
def test_graph_capture_replays():
    """gary_tc_count (learned, heuristic, bitmap) captured in a HIP graph by a C++ caller:
    every replay adds the eager count, with eager counts between replays."""
    r = subprocess.run([str(CAPTURE), "14"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASSED" in r.stdout
