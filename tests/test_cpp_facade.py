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


def test_index_cache_smem_takes_a_block_shape(tmp_path):
    """getIndexCacheSMem accepts a dim3-like block shape as the reference's does
    (src/GraphQueryHom.cu:11-13): waves of x*y*z threads times the cache size."""
    src = tmp_path / "t.cpp"
    src.write_text('''#include "kernels/GraphQueryHom.h"
#include <cstdio>
struct Shape { unsigned x, y, z; };
int main() {
  const size_t a = graph_query::getIndexCacheSMem(Shape{256, 1, 1}, 32);
  const size_t b = graph_query::getIndexCacheSMem(Shape{64, 2, 2}, 32);
  const size_t c = graph_query::getIndexCacheSMem(256, 32);
  std::printf("%zu %zu %zu\\n", a, b, c);
  return (a == 128 && b == 128 && c == 128) ? 0 : 1;
}
''')
    exe = tmp_path / "t"
    lib = ROOT / "gary_amd" / "libgary.so"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src),
                        "-o", str(exe), str(lib), f"-Wl,-rpath,{lib.parent}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr


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
