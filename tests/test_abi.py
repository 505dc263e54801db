"""The drop-in boundary without a GPU: libbgpq.so exports every function the
headers in include/ declare, and the headers compile as C (bgpq*.h) and as
C++ (the bgpq_heap.hpp facade)."""
import ctypes
import glob
import os
import re
import subprocess

import pytest

from gpuheap_amd._lib import LIB_PATH, SIGNATURES

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")
DECL = re.compile(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(bgpq_\w+)\s*\(", re.M)


def _declared():
    names = set()
    for h in glob.glob(os.path.join(INCLUDE, "*.h")):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names.update(DECL.findall(text))
    return sorted(names)


def test_headers_declare_the_entry_points():
    names = _declared()
    for must in ("bgpq_create", "bgpq_insert", "bgpq_delete", "bgpq_heapsort", "bgpq_check", "bgpq_sssp",
                 "bgpq_astar", "bgpq_knapsack", "bgpq_server_submit", "bgpq_batch_run"):
        assert must in names


def test_library_exports_every_declared_function():
    assert os.path.exists(LIB_PATH), "libbgpq.so is not built"
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_every_declared_function():
    assert set(_declared()) <= set(SIGNATURES), sorted(set(_declared()) - set(SIGNATURES))


@pytest.mark.parametrize("header", ["bgpq.h", "bgpq_apps.h", "bgpq_serve.h"])
def test_header_compiles_as_c(header, tmp_path):
    src = tmp_path / "t.c"
    src.write_text(f'#include "{header}"\nint main(void) {{ return 0; }}\n')
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-fsyntax-only", f"-I{INCLUDE}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_cpp_facade_compiles(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text('#include "bgpq_heap.hpp"\n#include "bgpq_serve.h"\n'
                   "int main() { bgpq::Heap<unsigned> *p = nullptr; (void)p; return 0; }\n")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", f"-I{INCLUDE}", str(src)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_device_api_is_public():
    # the workgroup-collective heap and its primitives are headers in include/
    for h in ("bgpq_device.hpp", "bgpq_prims.hpp"):
        assert os.path.exists(os.path.join(INCLUDE, h)), h
    text = open(os.path.join(INCLUDE, "bgpq_device.hpp")).read()
    for sym in ("struct DevHeap", "view_from", "deleteRoot(K* items, int& size, K threshold", "ifTerminate",
                "voteBusy"):
        assert sym in text, sym


def test_external_consumer_uses_only_public_headers():
    # tests/ext builds its kernel from include/ and libbgpq.so alone
    ext = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ext")
    src = open(os.path.join(ext, "ext_knapsack.hip")).read()
    incs = re.findall(r'#include\s+"([^"]+)"', src)
    assert incs and all(os.path.exists(os.path.join(INCLUDE, i)) for i in incs), incs
    mk = open(os.path.join(ext, "Makefile")).read()
    assert "csrc" not in mk.replace("Nothing under gpuheap_amd/csrc is included", "")
