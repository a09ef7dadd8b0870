"""The C-ABI library loads without a GPU and exports every function that
include/*.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
LIB = os.path.join(ROOT, "pamr_amrd_amd", "libpamr_mg.so")


def declared_functions():
    names = set()
    for fn in sorted(os.listdir(INC)):
        if not fn.endswith(".h"):
            continue
        src = open(os.path.join(INC, fn)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"//[^\n]*", "", src)
        src = "\n".join(l for l in src.splitlines() if not l.lstrip().startswith("#"))
        src = re.sub(r"extern\s+\"C\"\s*\{", "", src)
        for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\(([^;{}]*)\)\s*;", src):
            head, name = m.group(1), m.group(2)
            if "typedef" in head or "(" in head or name in ("sizeof",):
                continue
            if re.search(r"\(\s*\*", m.group(0)):  # function pointer variable
                continue
            names.add(name)
    return sorted(names)


def test_headers_declare_something():
    names = declared_functions()
    assert "pmg_dmfwr3d" in names and "pmg_create" in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="libpamr_mg.so not built")
def test_library_exports_all_declared_functions():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
