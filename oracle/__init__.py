"""CPU oracle for the monolis Krylov path -- TEST INFRASTRUCTURE ONLY.

This package restates monolis's FP64 CPU algorithm (SpMV, dot, AXPBY,
convergence test, diagonal / block-Jacobi preconditioner, CG, BiCGSTAB) in
plain C (``monolis_oracle.c``) and exposes it through ctypes.  Only the test
suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it, and only to check results or to time the CPU baseline.  The
HIP product path never imports or calls anything here.

Parity status of each function and what pins it is listed in the C file
header.  Nothing here is "parity unpinned": every function is checked by a
reference golden vector / closed form or by a textbook identity in
``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_c128p = np.ctypeslib.ndpointer(dtype=np.complex128, flags="C_CONTIGUOUS")

# Iarray / Rarray slots (C index k == Fortran index k), src/define/def_solver_prm.f90:90-163
I_METHOD, I_PRECOND, I_MAX_ITER, I_CUR_ITER, I_IERR = 1, 2, 3, 4, 5
I_IS_INIT_X, I_IS_ERROR_ABORT = 10, 16
R_TOL, R_CUR_RESID = 1, 2
ITER_CG, ITER_GROPCG, ITER_PIPECG, ITER_BICGSTAB, ITER_PIPEBICGSTAB, ITER_BICGSTAB_NOPREC = 1, 2, 3, 5, 6, 7
ITER_PIPEBICGSTAB_NOPREC = 8
ITER_PIPECR = 4
ITER_COCG = 13
ZERO_DIAG = -6   # status of orc_solve_C when a diagonal block has a zero pivot
PREC_NONE, PREC_DIAG = 0, 1


def build() -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_double = ctypes.c_int, ctypes.c_double
        L.orc_set_threads.argtypes = [c_int]
        L.orc_get_threads.restype = c_int
        L.orc_matvec.argtypes = [c_int, c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                 _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_inner_product.argtypes = [c_int, _f64p, _f64p]
        L.orc_inner_product.restype = c_double
        L.orc_sorted_inner_product.argtypes = [c_int, _f64p, _f64p]
        L.orc_sorted_inner_product.restype = c_double
        L.orc_vec_axpby.argtypes = [c_int, c_double, _f64p, c_double, _f64p, _f64p]
        L.orc_vec_copy.argtypes = [c_int, _f64p, _f64p]
        L.orc_check_converge.argtypes = [c_int, _f64p, c_double, c_int, _i32p, _f64p]
        L.orc_check_converge.restype = c_int
        L.orc_diag_nn_setup.argtypes = [c_int, c_int, _i32p, _i32p, _f64p, _f64p]
        L.orc_diag_33_setup.argtypes = [c_int, _i32p, _i32p, _f64p, _f64p]
        L.orc_diag_nn_apply.argtypes = [c_int, c_int, _f64p, _f64p, _f64p]
        L.orc_diag_33_apply.argtypes = [c_int, _f64p, _f64p, _f64p]
        L.orc_diag_V_setup.argtypes = [c_int, _i32p, _i32p, _i32p, _i32p, _f64p, _i64p, _f64p]
        L.orc_diag_V_apply.argtypes = [c_int, _i32p, _i32p, _i64p, _f64p, _f64p, _f64p]
        L.orc_solve.argtypes = [c_int, c_int, c_int, c_int, ctypes.c_void_p, ctypes.c_void_p,
                                ctypes.c_void_p, _f64p, _f64p, _f64p, _i32p, _i32p, _i32p, _f64p]
        L.orc_solve.restype = c_int
        L.orc_time_matvec.argtypes = [c_int, c_int, c_int, _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_time_matvec.restype = c_double
        L.orc_matvec_C.argtypes = [c_int, c_int, _i32p, _i32p, _c128p, _c128p, _c128p]
        L.orc_inner_product_C.argtypes = [c_int, _c128p, _c128p, _c128p]
        L.orc_check_converge_C.argtypes = [c_int, _c128p, _c128p, c_int, _i32p, _f64p]
        L.orc_check_converge_C.restype = c_int
        L.orc_diag_nn_setup_C.argtypes = [c_int, c_int, _i32p, _i32p, _c128p, _c128p]
        L.orc_diag_nn_setup_C.restype = c_int
        L.orc_diag_nn_apply_C.argtypes = [c_int, c_int, _c128p, _c128p, _c128p]
        L.orc_solve_C.argtypes = [c_int, c_int, _i32p, _i32p, _c128p, _c128p, _c128p, _i32p, _f64p]
        L.orc_solve_C.restype = c_int
        L.orc_matvec_V_C.argtypes = [c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _c128p, _c128p, _c128p]
        L.orc_diag_V_setup_C.argtypes = [c_int, _i32p, _i32p, _i32p, _i32p, _c128p, _i64p, _c128p]
        L.orc_diag_V_setup_C.restype = c_int
        L.orc_diag_V_apply_C.argtypes = [c_int, _i32p, _i32p, _i64p, _c128p, _c128p, _c128p]
        L.orc_solve_V_C.argtypes = [c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _c128p, _c128p, _c128p, _i32p,
                                    _f64p]
        L.orc_solve_V_C.restype = c_int
        L.orc_pattern_simple_mesh.argtypes = [c_int, c_int, c_int, _i32p, _i32p, _i32p, ctypes.c_longlong]
        L.orc_pattern_simple_mesh.restype = c_int
        L.orc_add_matrix.argtypes = [c_int, _i32p, _i32p, _f64p, c_int, _i32p, _f64p]
        L.orc_add_matrix.restype = c_int
        L.orc_add_elements.argtypes = [c_int, _i32p, _i32p, _f64p, c_int, c_int, _i32p, _f64p]
        L.orc_add_elements.restype = c_int
        L.orc_dirichlet.argtypes = [c_int, c_int, _i32p, _i32p, _f64p, _f64p, c_int, c_int, c_double]
        L.orc_dirichlet.restype = c_int
        L.orc_dirichlet_list.argtypes = [c_int, c_int, _i32p, _i32p, _f64p, _f64p, c_int, _i32p, _i32p, _f64p]
        L.orc_dirichlet_list.restype = c_int
        L.orc_dirichlet_C.argtypes = [c_int, c_int, _i32p, _i32p, _f64p, _f64p, c_int, c_int, c_double, c_double]
        L.orc_dirichlet_C.restype = c_int
        L.orc_dirichlet_list_C.argtypes = [c_int, c_int, _i32p, _i32p, _f64p, _f64p, c_int, _i32p, _i32p, _f64p]
        L.orc_dirichlet_list_C.restype = c_int
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def matvec(mat, x):
    """y = A x over the N internal block rows (src/linalg/matvec.f90:123-188)."""
    ndof = mat.ndof
    m = mat.n_dof_index[mat.n] if ndof == -1 else mat.n * ndof
    y = np.zeros(max(m, 1), dtype=np.float64)
    lib().orc_matvec(mat.n, ndof, _ptr(mat.n_dof_list), _ptr(mat.n_dof_index), _ptr(mat.n_dof_index2),
                     mat.index, mat.item, mat.A, np.ascontiguousarray(x, dtype=np.float64), y)
    return y[:m]


def inner_product(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return float(lib().orc_inner_product(len(x), x, y))


def sorted_inner_product(x, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    return float(lib().orc_sorted_inner_product(len(x), x, y))


def vec_axpby(alpha, x, beta, y):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    z = np.empty_like(x)
    lib().orc_vec_axpby(len(x), float(alpha), x, float(beta), y, z)
    return z


def check_converge(r, B2, it, tol):
    Iarray = np.zeros(100, np.int32)
    Rarray = np.zeros(100, np.float64)
    Rarray[R_TOL] = tol
    r = np.ascontiguousarray(r, np.float64)
    conv = lib().orc_check_converge(len(r), r, float(B2), int(it), Iarray, Rarray)
    return bool(conv), int(Iarray[I_CUR_ITER]), float(Rarray[R_CUR_RESID])


def diag_setup(mat):
    """Factored diagonal blocks, the layout of PREC%R%D (diag_33.f90 / diag_nn.f90;
    diag_V.f90 for variable blocks, returned as (doff, D))."""
    n, nd = mat.n, mat.ndof
    if nd == -1:
        lst = np.ascontiguousarray(mat.n_dof_list, np.int32)
        tot = int(np.sum(lst[:n].astype(np.int64) ** 2))
        doff = np.zeros(n + 1, np.int64)
        D = np.zeros(max(tot, 1), np.float64)
        lib().orc_diag_V_setup(n, lst, np.ascontiguousarray(mat.n_dof_index2, np.int32), mat.index, mat.item,
                               mat.A, doff, D)
        return doff, D
    D = np.zeros(max(nd * nd * n, 1), np.float64)
    if nd == 3:
        lib().orc_diag_33_setup(n, mat.index, mat.item, mat.A, D)
    else:
        lib().orc_diag_nn_setup(n, nd, mat.index, mat.item, mat.A, D)
    return D


def diag_apply(mat, D, x):
    n, nd = mat.n, mat.ndof
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros_like(x)
    if nd == -1:
        doff, DD = D
        lib().orc_diag_V_apply(n, np.ascontiguousarray(mat.n_dof_list, np.int32),
                               np.ascontiguousarray(mat.n_dof_index, np.int32), doff, DD, x, y)
        return y
    if nd == 3:
        lib().orc_diag_33_apply(n, D, x, y)
    else:
        lib().orc_diag_nn_apply(n, nd, D, x, y)
    return y


def default_arrays(method=ITER_CG, precond=PREC_DIAG, tol=1e-8, max_iter=1000):
    """Iarray / Rarray as monolis_prm_initialize sets them (def_solver_prm.f90:183-222)."""
    Iarray = np.zeros(100, np.int32)
    Rarray = np.zeros(100, np.float64)
    Iarray[1], Iarray[2], Iarray[3], Iarray[4], Iarray[5] = method, precond, max_iter, 0, -1
    Iarray[10], Iarray[16] = 1, 1
    Rarray[1] = tol
    return Iarray, Rarray


def solve(mat, b, x0=None, Iarray=None, Rarray=None, **kw):
    """monolis_solve_main_R on one rank.  Returns (x, status, Iarray, Rarray)."""
    if Iarray is None:
        Iarray, Rarray = default_arrays(**kw)
    ndof = mat.ndof
    m = mat.n_dof_index[mat.n] if ndof == -1 else mat.n * ndof
    x = np.zeros(max(m, 1), np.float64) if x0 is None else np.array(x0, np.float64, copy=True)
    b = np.ascontiguousarray(b, np.float64)
    st = lib().orc_solve(mat.n, mat.n, len(mat.item), ndof, _ptr(mat.n_dof_list), _ptr(mat.n_dof_index),
                         _ptr(mat.n_dof_index2), mat.A, x, b, mat.index, mat.item, Iarray, Rarray)
    return x[:m], int(st), Iarray, Rarray


def time_matvec(mat, x, reps):
    y = np.zeros(mat.n * mat.ndof, np.float64)
    x = np.ascontiguousarray(x, np.float64)
    return float(lib().orc_time_matvec(int(reps), mat.n, mat.ndof, mat.index, mat.item, mat.A, x, y))


# ---------------------------------------------------------------------------
# FP64 CPU adjoint of x = A^{-1} b (test infrastructure, like the rest).
# monolis has no adjoint; this restates the textbook reverse-mode rule
#   A^T lambda = dL/dx,  dL/db = lambda,  dL/dA_ij = -lambda_i x_j
# on monolis's BCSR layout (src/matrix/def_mat.f90 arrays index / item / A).
# Pinned in tests/test_oracle_adjoint.py against dense solves, scipy's
# transpose and central finite differences of the loss.
# ---------------------------------------------------------------------------
def _rows_of_entries(mat):
    return np.repeat(np.arange(mat.n, dtype=np.int64), np.diff(np.asarray(mat.index, np.int64)))


def transpose(mat):
    """BCSR transpose (square, one rank): block (j, i) of A^T = (block (i, j) of A)^T,
    block columns ascending per row."""
    assert mat.ndof > 0 and mat.n == mat.np_
    n, b = mat.n, mat.ndof
    rows = _rows_of_entries(mat)
    cols = np.asarray(mat.item, np.int64) - 1
    order = np.lexsort((rows, cols))               # by column, then row
    counts = np.bincount(cols, minlength=n)
    index = np.zeros(n + 1, np.int32)
    index[1:] = np.cumsum(counts)
    item = (rows[order] + 1).astype(np.int32)
    A = np.asarray(mat.A, np.float64).reshape(-1, b, b)[order].transpose(0, 2, 1).reshape(-1).copy()
    return type(mat)(n, n, b, index, item, A).ensure_dof_index()


def outer_grad(mat, lam, x):
    """dA[e*b*b + r*b + c] = -lam[i*b + r] * x[j*b + c] for every stored block e = (i, j)."""
    b = mat.ndof
    rows = _rows_of_entries(mat)
    cols = np.asarray(mat.item, np.int64) - 1
    lb = np.asarray(lam, np.float64).reshape(-1, b)[rows]
    xb = np.asarray(x, np.float64).reshape(-1, b)[cols]
    return (-(lb[:, :, None] * xb[:, None, :])).reshape(-1)


def adjoint(mat, x, g, symmetric=False, **kw):
    """lambda = A^{-T} g by the oracle's monolis solver (A itself when symmetric),
    and dL/dA on the pattern.  Returns (lambda, dA, status, Iarray)."""
    At = mat if symmetric else transpose(mat)
    lam, st, I, R = solve(At, g, **kw)
    return lam, outer_grad(mat, lam, x), st, I


# ---------------------------------------------------------------------------
# Complex path (monolis's *_C routines and COCG), monolis_oracle_cplx.c.
# Values are complex128 (Fortran complex(kdouble) layout).
# ---------------------------------------------------------------------------
def _z(a):
    return np.ascontiguousarray(a, np.complex128)


def _vl(mat):
    return (np.ascontiguousarray(mat.n_dof_list, np.int32), np.ascontiguousarray(mat.n_dof_index, np.int32),
            np.ascontiguousarray(mat.n_dof_index2, np.int32))


def _mC(mat):
    return int(mat.n_dof_index[mat.n]) if mat.ndof == -1 else mat.n * mat.ndof


def matvec_C(mat, x):
    """y = A x in complex arithmetic (src/linalg/matvec.f90:192-227; n_dof = -1:
    monolis_matvec_V_C, :293-351)."""
    m = _mC(mat)
    y = np.zeros(max(m, 1), np.complex128)
    if mat.ndof == -1:
        lst, dofi, dofi2 = _vl(mat)
        lib().orc_matvec_V_C(mat.n, lst, dofi, dofi2, mat.index, mat.item, _z(mat.A), _z(x), y)
    else:
        lib().orc_matvec_C(mat.n, mat.ndof, mat.index, mat.item, _z(mat.A), _z(x), y)
    return y[:m]


def inner_product_C(x, y):
    """Unconjugated sum_i x_i y_i (src/linalg/inner_product.f90:457-499)."""
    x, y = _z(x), _z(y)
    out = np.zeros(1, np.complex128)
    lib().orc_inner_product_C(len(x), x, y, out)
    return complex(out[0])


def check_converge_C(r, B2, it, tol):
    Iarray = np.zeros(100, np.int32)
    Rarray = np.zeros(100, np.float64)
    Rarray[R_TOL] = tol
    r = _z(r)
    b2 = np.array([B2], np.complex128)
    conv = lib().orc_check_converge_C(len(r), r, b2, int(it), Iarray, Rarray)
    return bool(conv), int(Iarray[I_CUR_ITER]), float(Rarray[R_CUR_RESID])


def diag_setup_C(mat):
    """Complex block-Jacobi factors (diag_nn.f90:90-159; n_dof = -1: diag_V.f90:88-158,
    returned as (doff, D)).  Returns (D, status); status ZERO_DIAG where the
    reference stops on a zero pivot."""
    if mat.ndof == -1:
        lst, dofi, dofi2 = _vl(mat)
        tot = int(np.sum(lst[:mat.n].astype(np.int64) ** 2))
        doff = np.zeros(mat.n + 1, np.int64)
        D = np.zeros(max(tot, 1), np.complex128)
        st = lib().orc_diag_V_setup_C(mat.n, lst, dofi2, mat.index, mat.item, _z(mat.A), doff, D)
        return (doff, D), int(st)
    D = np.zeros(max(mat.ndof * mat.ndof * mat.n, 1), np.complex128)
    st = lib().orc_diag_nn_setup_C(mat.n, mat.ndof, mat.index, mat.item, _z(mat.A), D)
    return D, int(st)


def diag_apply_C(mat, D, x):
    x = _z(x)
    y = np.zeros_like(x)
    if mat.ndof == -1:
        doff, DD = D
        lst, dofi, _ = _vl(mat)
        lib().orc_diag_V_apply_C(mat.n, lst, dofi, doff, _z(DD), x, y)
        return y
    lib().orc_diag_nn_apply_C(mat.n, mat.ndof, _z(D), x, y)
    return y


def solve_C(mat, b, x0=None, Iarray=None, Rarray=None, **kw):
    """monolis_solve_main_C (COCG) on one rank.  Returns (x, status, Iarray, Rarray)."""
    if Iarray is None:
        kw.setdefault("method", ITER_COCG)
        Iarray, Rarray = default_arrays(**kw)
    m = _mC(mat)
    x = np.zeros(max(m, 1), np.complex128) if x0 is None else np.array(x0, np.complex128, copy=True)
    if mat.ndof == -1:
        lst, dofi, dofi2 = _vl(mat)
        st = lib().orc_solve_V_C(mat.n, lst, dofi, dofi2, mat.index, mat.item, _z(mat.A), x, _z(b), Iarray, Rarray)
    else:
        st = lib().orc_solve_C(mat.n, mat.ndof, mat.index, mat.item, _z(mat.A), x, _z(b), Iarray, Rarray)
    return x[:m], int(st), Iarray, Rarray


# ---------------------------------------------------------------------------
# Assembly (monolis_oracle_asm.c): pattern of a simple mesh, element matrices,
# Dirichlet conditions.  Fixed block size, one rank, 0-based node ids.
# ---------------------------------------------------------------------------
def pattern_by_simple_mesh(n_node, elems):
    """(index, item) of monolis_get_nonzero_pattern_by_simple_mesh_R; item 1-based."""
    e = np.ascontiguousarray(elems, np.int32)
    n_elem, n_base = (e.shape[0], e.shape[1]) if e.ndim == 2 and e.size else (0, 1)
    cap = int(n_elem) * n_base * n_base + n_node
    index = np.zeros(n_node + 1, np.int32)
    item = np.zeros(max(cap, 1), np.int32)
    nz = lib().orc_pattern_simple_mesh(int(n_node), int(n_base), int(n_elem), e.reshape(-1) if e.size else
                                       np.zeros(1, np.int32), index, item, cap)
    if nz < 0:
        raise ValueError("invalid connectivity")
    return index, item[:nz].copy()


def add_elements(mat, elems, Ke):
    """A += element matrices, one monolis_add_matrix_to_sparse_matrix_R call per
    element in order (row-major (n_base*b)^2 element matrices).  In place on mat.A."""
    e = np.ascontiguousarray(elems, np.int32)
    k = np.ascontiguousarray(Ke, np.float64).reshape(-1)
    st = lib().orc_add_elements(mat.ndof, mat.index, mat.item, mat.A, int(e.shape[0]), int(e.shape[1]),
                                e.reshape(-1), k)
    if st:
        raise KeyError("element couples nodes outside the nonzero pattern")
    return mat


def dirichlet_C(index, item, A, n, ndof, B, node, dof, val):
    """Complex conditions applied one by one (monolis_set_Dirichlet_bc_C) on a complex
    BCSR matrix: A (nz*ndof^2 complex128) and B (n*ndof complex128) updated in place."""
    node = np.ascontiguousarray(node, np.int32)
    dof = np.ascontiguousarray(dof, np.int32)
    v = np.ascontiguousarray(np.asarray(val, np.complex128)).view(np.float64)
    Af = A.view(np.float64)
    Bf = B.view(np.float64)
    st = lib().orc_dirichlet_list_C(n, ndof, np.ascontiguousarray(index, np.int32),
                                    np.ascontiguousarray(item, np.int32), Af, Bf, len(node), node, dof, v)
    if st:
        raise KeyError("Dirichlet condition on a node without a diagonal block")
    return A, B


def dirichlet(mat, B, node, dof, val):
    """Conditions applied one by one (monolis_set_Dirichlet_bc_R); mat.A and B in place."""
    node = np.ascontiguousarray(node, np.int32)
    dof = np.ascontiguousarray(dof, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    st = lib().orc_dirichlet_list(mat.n, mat.ndof, mat.index, mat.item, mat.A, B, len(node), node, dof, val)
    if st:
        raise KeyError("Dirichlet condition on a node without a diagonal block")
    return mat, B
