"""The C-ABI library loads and exports every symbol include/llfem.h declares
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

from conftest import ROOT
from lowlevelfem import _lib


def declared():
    src = open(os.path.join(ROOT, "include", "llfem.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(llfem_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "llfem_assemble" in names and "llfem_mesh_create" in names and "llfem_solve" in names
    assert set(names) == set(_lib.EXPORTED)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_no_internal_symbols_leak():
    out = os.popen(f"nm -D --defined-only {_lib.LIB_PATH}").read()
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert {n for n in exported if n.startswith("llfem_")} == set(declared())
