"""CPU checks of the C-ABI boundary: the library loads (no GPU call is made)
and exports every function declared in include/*.h, with the ctypes binding
covering each of them."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text):
            name = m.group(1)
            prefix = text[max(0, m.start() - 40):m.start()]
            if name.startswith("monolis") and not re.search(r"typedef|#define", prefix.split("\n")[-1]):
                names.append(name)
    return sorted(set(names))


def test_headers_declare_reference_entry_points():
    names = declared_functions()
    for ref in ["monolis_solve_R_c_main", "monolis_matvec_product_R_c_main", "monolis_inner_product_R_c_main",
                "monolis_solve_C_c_main", "monolis_matvec_product_C_c_main", "monolis_inner_product_C_c_main"]:
        assert ref in names


def test_library_exports_every_declared_symbol():
    from monolis_amd import _lib
    L = _lib.load()   # dlopen only: no HIP call
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_ctypes_binding_covers_every_declared_symbol():
    from monolis_amd import _lib
    bound = {s[0] for s in _lib.SIGNATURES}
    missing = [n for n in declared_functions() if n not in bound]
    assert not missing, missing


def test_library_is_gfx950_code():
    """The shared object carries a gfx950 code object (hipcc --offload-arch=gfx950)."""
    from monolis_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
