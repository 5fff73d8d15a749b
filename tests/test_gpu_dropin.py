"""The C++ drop-in headers (include/LAKCatmullClark.h, include/CatmullClark.h)
used as the reference's own demo uses them, compiled against libslak.so."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "slak_amd", "lib", "slak_dropin_test")


@pytest.mark.gpu
def test_cpp_dropin_program():
    assert os.path.exists(BIN), "build with __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok dropin" in r.stdout


def test_cpp_dropin_compiles():
    """Header-only C++ wrappers compile with a plain host compiler (no HIP headers)."""
    out = os.path.join(ROOT, "_build", "dropin_syntax.o")
    os.makedirs(os.path.dirname(out), exist_ok=True)
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), "-c",
                        os.path.join(ROOT, "tests", "cpp", "dropin_test.cpp"), "-o", out],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
