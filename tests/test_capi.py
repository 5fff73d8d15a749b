"""The C-ABI library loads and exports every entry point include/slak_c.h
declares (no compute calls: runs without a GPU)."""
import ctypes
import os
import re

from slak_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for hdr in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if not hdr.endswith(".h") or not hdr.startswith("slak"):
            continue
        text = open(os.path.join(ROOT, "include", hdr)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(slak_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "slak_subdivide" in names and "slak_free" in names and "slak_last_error" in names


def test_library_exports_every_declared_symbol():
    lib = capi.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_table_matches_header():
    bound = {n for n, _, _ in capi.EXPORTS}
    assert bound == set(declared_functions())


def test_struct_layout():
    # slak_mesh: 2 int32 + 4 uint64 + 4 pointers
    assert ctypes.sizeof(capi.SlakMesh) == 8 + 32 + 32
    assert ctypes.sizeof(capi.SlakStats) == 8 + 8 + 4 + 4 + 8 + 8


def test_version_without_gpu():
    assert capi.lib().slak_version() >= 100
