"""The drop-in C++ headers compile on their own (no GPU): the two MeshInfo
types are distinct and can be overloaded and specialised on."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_meshinfo_types_distinct_and_overloadable():
    src = os.path.join(ROOT, "tests", "cpp", "meshinfo_types_test.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                        src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_headers_compile_standalone():
    for hdr in ("LAKCatmullClark.h", "CatmullClark.h", "Mesh.h", "obj.h", "slak_c.h"):
        r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-x", "c++", "-I", os.path.join(ROOT, "include"),
                            os.path.join(ROOT, "include", hdr)], capture_output=True, text=True)
        assert r.returncode == 0, (hdr, r.stderr)
