"""CPU checks of the C-ABI boundary: the libraries load (no GPU call is made)
and export every function their headers declare -- include/monolis_hip.h by
libmonolis_hip.so, with the ctypes binding covering each of them, and
include/monolis.h (monolis's C API) by libmonolis_front.so."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions(header="monolis_hip.h"):
    names = []
    for h in glob.glob(os.path.join(ROOT, "include", header)):
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


def test_front_library_exports_monolis_c_api():
    names = declared_functions("monolis.h")
    for ref in ["monolis_initialize", "monolis_finalize", "monolis_get_nonzero_pattern_by_simple_mesh_R",
                "monolis_add_scalar_to_sparse_matrix_R", "monolis_set_Dirichlet_bc_R", "monolis_solve_R",
                "monolis_matvec_product_R", "monolis_inner_product_R", "monolis_solve_C", "monolis_set_method"]:
        assert ref in names
    from monolis_amd import _lib
    front = ctypes.CDLL(os.path.join(os.path.dirname(_lib.LIB_PATH), "libmonolis_front.so"))
    missing = [n for n in names if not hasattr(front, n)]
    assert not missing, missing


def test_front_c_program_host_checks():
    """monolis_amd/c/test_front.c, a C program written against include/monolis.h
    and linked against libmonolis_front.so: pattern, assembly, Dirichlet and
    parameter checks (the reference's unit-test answers); no GPU call."""
    import subprocess
    exe = os.path.join(ROOT, "monolis_amd", "bin", "test_front")
    assert os.path.exists(exe), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    r = subprocess.run([exe, "host"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok (0 failed)" in r.stdout


@pytest.mark.parametrize("case,msg", [("err_pattern", "not in the nonzero pattern"),
                                      ("err_dof", "outside its block"),
                                      ("err_dbc", "without its diagonal block"),
                                      ("err_kind", "holds no complex values"),
                                      ("err_matval", "differ from the matrix's pattern")])
def test_front_c_program_stops_like_monolis(case, msg):
    """monolis stops the program on these (monolis_stop_by_matrix_assemble,
    _by_submatrix_access, _by_set_DBC); the C API does the same: exit status 1
    and a message on stderr.  The last two would write outside monolis's own
    arrays (a value kind the pattern was not built for, a value copy of the
    wrong size); they stop the same way."""
    import subprocess
    exe = os.path.join(ROOT, "monolis_amd", "bin", "test_front")
    assert os.path.exists(exe), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    r = subprocess.run([exe, case], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1, r.stdout + r.stderr
    assert msg in r.stderr and "no error raised" not in r.stdout


def test_library_is_gfx950_code():
    """The shared object carries a gfx950 code object (hipcc --offload-arch=gfx950)."""
    from monolis_amd import _lib
    data = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in data
