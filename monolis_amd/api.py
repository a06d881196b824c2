"""Python mirror of monolis's solver interface over libmonolis_hip.so.

Names, argument meaning and defaults follow monolis's C API
(wrapper/solver/monolis_solver_c.c:8-60, wrapper/linalg/monolis_matvec_c.c:8-58,
wrapper/linalg/monolis_inner_product_c.c:32-44,
wrapper/define/monolis_def_solver_prm_c.c:7-45,
wrapper/define/monolis_def_solver_prm_util_c.c:7-194), so tests read like the
reference's own.  Host numpy arrays go through the flat C-ABI exactly as a C
caller's arrays would; ``DeviceMatrix`` / ``DeviceBuffer`` use the
device-resident extension.
"""
from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _lib
from .matrix import Matrix

# ids, src/define/def_solver_prm.f90:7-58 / wrapper/define/monolis_def_solver_prm_c.h
MONOLIS_ITER_CG, MONOLIS_ITER_GROPCG, MONOLIS_ITER_PIPECG, MONOLIS_ITER_PIPECR = 1, 2, 3, 4
MONOLIS_ITER_BICGSTAB, MONOLIS_ITER_PIPEBICGSTAB = 5, 6
MONOLIS_ITER_BICGSTAB_NOPREC, MONOLIS_ITER_PIPEBICGSTAB_NOPREC = 7, 8
MONOLIS_ITER_COCG = 13
MONOLIS_PREC_NONE, MONOLIS_PREC_DIAG, MONOLIS_PREC_SOR, MONOLIS_PREC_LU = 0, 1, 2, 3
MONOLIS_PRM_I_METHOD, MONOLIS_PRM_I_PRECOND, MONOLIS_PRM_I_MAX_ITER = 1, 2, 3
MONOLIS_PRM_I_CUR_ITER, MONOLIS_PRM_I_IERR = 4, 5
MONOLIS_PRM_I_IS_INIT_X, MONOLIS_PRM_I_IS_SYM_MATRIX, MONOLIS_PRM_I_IS_DEBUG = 10, 11, 12
MONOLIS_PRM_I_IS_MEASUREMENT, MONOLIS_PRM_I_IS_CHECK_DIAG = 13, 14
MONOLIS_PRM_I_IS_PREC_STORED, MONOLIS_PRM_I_IS_ERROR_ABORT = 15, 16
MONOLIS_PRM_I_SHOW_ITERLOG, MONOLIS_PRM_I_SHOW_TIME, MONOLIS_PRM_I_SHOW_SUMMARY = 20, 21, 22
MONOLIS_PRM_I_SHOW_TIME_STATISTICS, MONOLIS_PRM_I_IDRS_DIM, MONOLIS_PRM_I_SPMV_FORMAT = 23, 24, 25
MONOLIS_PRM_R_TOL, MONOLIS_PRM_R_CUR_RESID = 1, 2
MONOLIS_R_TIME_SOL, MONOLIS_R_TIME_PREP, MONOLIS_R_TIME_SPMV = 3, 4, 5
MONOLIS_R_TIME_DOTP, MONOLIS_R_TIME_PREC = 6, 7
MONOLIS_R_TIME_COMM_DOTP, MONOLIS_R_TIME_COMM_SPMV = 8, 9


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class Prm:
    """MONOLIS_PRM (monolis_def_solver_prm_c.h:136-142) with monolis_prm_initialize defaults."""

    def __init__(self):
        self.Iarray = np.zeros(100, np.int32)
        self.Rarray = np.zeros(100, np.float64)
        I = self.Iarray
        I[MONOLIS_PRM_I_METHOD] = 1
        I[MONOLIS_PRM_I_PRECOND] = 1
        I[MONOLIS_PRM_I_MAX_ITER] = 1000
        I[MONOLIS_PRM_I_CUR_ITER] = 0
        I[MONOLIS_PRM_I_IERR] = -1
        I[MONOLIS_PRM_I_IS_INIT_X] = 1
        I[MONOLIS_PRM_I_IS_ERROR_ABORT] = 1
        I[MONOLIS_PRM_I_SHOW_ITERLOG] = 1
        I[MONOLIS_PRM_I_SHOW_TIME] = 1
        I[MONOLIS_PRM_I_SHOW_SUMMARY] = 1
        I[MONOLIS_PRM_I_IDRS_DIM] = 4
        self.Rarray[MONOLIS_PRM_R_TOL] = 1.0e-8

    def quiet(self) -> "Prm":
        self.Iarray[MONOLIS_PRM_I_SHOW_ITERLOG] = 0
        self.Iarray[MONOLIS_PRM_I_SHOW_TIME] = 0
        self.Iarray[MONOLIS_PRM_I_SHOW_SUMMARY] = 0
        return self


class Monolis:
    """MONOLIS {prm, mat, prec} (monolis_def_struc_c.h:13-17)."""

    def __init__(self, mat: Matrix, prm: Optional[Prm] = None):
        self.mat = mat.ensure_dof_index()
        self.prm = prm if prm is not None else Prm()


# setters / getters, monolis_def_solver_prm_util_c.c:7-194
def monolis_set_method(m: Monolis, v: int): m.prm.Iarray[MONOLIS_PRM_I_METHOD] = v
def monolis_set_precond(m: Monolis, v: int): m.prm.Iarray[MONOLIS_PRM_I_PRECOND] = v
def monolis_set_maxiter(m: Monolis, v: int): m.prm.Iarray[MONOLIS_PRM_I_MAX_ITER] = v
def monolis_set_tolerance(m: Monolis, v: float): m.prm.Rarray[MONOLIS_PRM_R_TOL] = v
def monolis_set_init_x(m: Monolis, v: bool): m.prm.Iarray[MONOLIS_PRM_I_IS_INIT_X] = int(v)
def monolis_set_error_abort(m: Monolis, v: bool): m.prm.Iarray[MONOLIS_PRM_I_IS_ERROR_ABORT] = int(v)
def monolis_set_performance_measurement(m: Monolis, v: bool): m.prm.Iarray[MONOLIS_PRM_I_IS_MEASUREMENT] = int(v)
def monolis_show_iterlog(m: Monolis, v: bool): m.prm.Iarray[MONOLIS_PRM_I_SHOW_ITERLOG] = int(v)
def monolis_show_timelog(m: Monolis, v: bool): m.prm.Iarray[MONOLIS_PRM_I_SHOW_TIME] = int(v)
def monolis_show_summary(m: Monolis, v: bool): m.prm.Iarray[MONOLIS_PRM_I_SHOW_SUMMARY] = int(v)
def monolis_get_converge_iter(m: Monolis) -> int: return int(m.prm.Iarray[MONOLIS_PRM_I_CUR_ITER])
def monolis_get_error_tag(m: Monolis) -> int: return int(m.prm.Iarray[MONOLIS_PRM_I_IERR])
def monolis_get_converge_residual(m: Monolis) -> float: return float(m.prm.Rarray[MONOLIS_PRM_R_CUR_RESID])
def monolis_get_time_solver(m: Monolis) -> float: return float(m.prm.Rarray[MONOLIS_R_TIME_SOL])
def monolis_get_time_preparing(m: Monolis) -> float: return float(m.prm.Rarray[MONOLIS_R_TIME_PREP])


def _flat_mat_args(mat: Matrix):
    """The 14 matrix arguments of the flat entry points (monolis_solver_c.c:27-41)."""
    mat.ensure_dof_index()
    return [mat.n, mat.np_, mat.nz, mat.ndof, mat.npndof, mat.nzndof2,
            _p(mat.n_dof_list), _p(mat.n_dof_index), _p(mat.n_dof_index2), _p(mat.A)]


_EMPTY = np.zeros(1, np.int32)


def _self_comm_args():
    """monolis_com_initialize_by_self: one rank, empty halo tables."""
    e = _p(_EMPTY)
    return [0, 0, 1, 0, 0, e, e, e, 0, 0, e, e, e]


def monolis_solve_R(monolis: Monolis, b: np.ndarray, x: np.ndarray) -> np.ndarray:
    """C monolis_solve_R (monolis_solver_c.c:8-60): x is overwritten in place."""
    L = _lib.load()
    mat = monolis.mat
    assert x.dtype == np.float64 and x.flags.c_contiguous and len(x) >= mat.npndof
    b = np.ascontiguousarray(b, np.float64)
    args = _flat_mat_args(mat)
    L.monolis_solve_R_c_main(*args[:10], _p(x), _p(b), _p(mat.index), _p(mat.item),
                             *_self_comm_args(), _p(monolis.prm.Iarray), _p(monolis.prm.Rarray))
    return x


def monolis_matvec_product_R(monolis: Monolis, x: np.ndarray, y: Optional[np.ndarray] = None) -> np.ndarray:
    """C monolis_matvec_product_R (monolis_matvec_c.c:8-58)."""
    L = _lib.load()
    mat = monolis.mat
    x = np.ascontiguousarray(x, np.float64)
    if y is None:
        y = np.zeros(mat.npndof, np.float64)
    args = _flat_mat_args(mat)
    L.monolis_matvec_product_R_c_main(*args[:10], _p(x), _p(y), _p(mat.index), _p(mat.item),
                                      *_self_comm_args())
    return y


def monolis_inner_product_R(monolis: Monolis, n_dof: int, x: np.ndarray, y: np.ndarray) -> float:
    """C monolis_inner_product_R (monolis_inner_product_c.c:32-44): over N*n_dof entries."""
    L = _lib.load()
    m = monolis.mat.n * n_dof
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    s = np.zeros(1, np.float64)
    L.monolis_inner_product_R_c_main(m, _p(x), _p(y), _p(s), 0)
    return float(s[0])


# complex path (wrapper/solver/monolis_solver_c.c monolis_solve_C,
# wrapper/linalg/monolis_matvec_c.c monolis_matvec_product_C,
# wrapper/linalg/monolis_inner_product_c.c monolis_inner_product_C)
def _flat_mat_args_C(mat: Matrix):
    mat.ensure_dof_index()
    if mat.A.dtype != np.complex128:
        raise TypeError("complex entry points take complex128 matrix values")
    return _flat_mat_args(mat)


def monolis_solve_C(monolis: Monolis, b: np.ndarray, x: np.ndarray) -> np.ndarray:
    """C monolis_solve_C: x (complex128) is overwritten in place."""
    L = _lib.load()
    mat = monolis.mat
    assert x.dtype == np.complex128 and x.flags.c_contiguous and len(x) >= mat.npndof
    b = np.ascontiguousarray(b, np.complex128)
    args = _flat_mat_args_C(mat)
    L.monolis_solve_C_c_main(*args[:10], _p(x), _p(b), _p(mat.index), _p(mat.item),
                             *_self_comm_args(), _p(monolis.prm.Iarray), _p(monolis.prm.Rarray))
    return x


def monolis_matvec_product_C(monolis: Monolis, x: np.ndarray, y: Optional[np.ndarray] = None) -> np.ndarray:
    """C monolis_matvec_product_C: y = A x in complex arithmetic."""
    L = _lib.load()
    mat = monolis.mat
    x = np.ascontiguousarray(x, np.complex128)
    if y is None:
        y = np.zeros(mat.npndof, np.complex128)
    args = _flat_mat_args_C(mat)
    L.monolis_matvec_product_C_c_main(*args[:10], _p(x), _p(y), _p(mat.index), _p(mat.item),
                                      *_self_comm_args())
    return y


def monolis_inner_product_C(monolis: Monolis, n_dof: int, x: np.ndarray, y: np.ndarray) -> complex:
    """C monolis_inner_product_C: unconjugated sum over N*n_dof entries."""
    L = _lib.load()
    m = monolis.mat.n * n_dof
    x = np.ascontiguousarray(x, np.complex128)
    y = np.ascontiguousarray(y, np.complex128)
    s = np.zeros(1, np.complex128)
    L.monolis_inner_product_C_c_main(m, _p(x), _p(y), _p(s), 0)
    return complex(s[0])


# ---------------------------------------------------------------------------
# device-resident interface
# ---------------------------------------------------------------------------
class DeviceBuffer:
    """A device allocation owned by Python (freed with the object)."""

    def __init__(self, count: int, dtype=np.float64):
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        self.nbytes = max(1, self.count * self.dtype.itemsize)
        L = _lib.load()
        p = ctypes.c_void_p()
        _lib.check(L.monolis_hip_malloc(ctypes.byref(p), self.nbytes), "monolis_hip_malloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, self.dtype)
        assert a.size <= self.count
        _lib.check(_lib.load().monolis_hip_memcpy(ctypes.c_void_p(self.ptr), _p(a), a.nbytes), "upload")

    def numpy(self, count: Optional[int] = None) -> np.ndarray:
        n = self.count if count is None else count
        out = np.empty(n, self.dtype)
        _lib.check(_lib.load().monolis_hip_memcpy(_p(out), ctypes.c_void_p(self.ptr), out.nbytes), "download")
        return out

    def zero(self) -> None:
        _lib.check(_lib.load().monolis_hip_memset(ctypes.c_void_p(self.ptr), 0, self.nbytes), "memset")

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                _lib.load().monolis_hip_free(ctypes.c_void_p(self.ptr))
                self.ptr = None
        except Exception:
            pass


class DeviceMatrix:
    """A monolis BCSR matrix resident in HBM in the GPU layout."""

    def __init__(self, mat: Matrix):
        L = _lib.load()
        h = ctypes.c_void_p()
        mat.ensure_dof_index()
        if mat.ndof == -1:
            _lib.check(L.monolis_hip_matrix_create_V(
                ctypes.byref(h), mat.n, mat.np_, mat.nz, _p(np.ascontiguousarray(mat.n_dof_list, np.int32)),
                _p(np.ascontiguousarray(mat.n_dof_index, np.int32)),
                _p(np.ascontiguousarray(mat.n_dof_index2, np.int32)), _p(mat.index), _p(mat.item), _p(mat.A)),
                "monolis_hip_matrix_create_V")
        else:
            _lib.check(L.monolis_hip_matrix_create(ctypes.byref(h), mat.n, mat.np_, mat.nz, mat.ndof,
                                                   _p(mat.index), _p(mat.item), _p(mat.A)),
                       "monolis_hip_matrix_create")
        self.h = h.value
        self.mat = mat
        self.m = mat.nndof
        self.vec = mat.npndof

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().monolis_hip_matrix_destroy(ctypes.c_void_p(self.h))
                self.h = None
        except Exception:
            pass

    def info(self) -> dict:
        a = (ctypes.c_longlong * 8)()
        _lib.check(_lib.load().monolis_hip_matrix_info(ctypes.c_void_p(self.h), a), "info")
        keys = ["n", "np", "nz", "ndof", "slices", "slots", "bytes_fp64", "bytes_fp32"]
        return dict(zip(keys, list(a)))

    def set_values(self, A) -> None:
        """New block values, same pattern (host array or DeviceBuffer)."""
        if isinstance(A, DeviceBuffer):
            ptr = ctypes.c_void_p(A.ptr)
        else:
            A = np.ascontiguousarray(A, np.float64)
            ptr = _p(A)
        _lib.check(_lib.load().monolis_hip_matrix_set_values(ctypes.c_void_p(self.h), ptr), "set_values")

    def matvec(self, x: DeviceBuffer, y: DeviceBuffer) -> None:
        L = _lib.load()
        if x.dtype == np.float32:
            _lib.check(L.monolis_hip_matvec_S(ctypes.c_void_p(self.h), ctypes.c_void_p(x.ptr), ctypes.c_void_p(y.ptr)), "matvec_S")
        else:
            _lib.check(L.monolis_hip_matvec_R(ctypes.c_void_p(self.h), ctypes.c_void_p(x.ptr), ctypes.c_void_p(y.ptr)), "matvec_R")

    def solve(self, b: DeviceBuffer, x: DeviceBuffer, prm: Prm, hist_len: int = 0):
        hist = np.zeros(max(hist_len, 1), np.float64)
        rc = _lib.load().monolis_hip_solve_R(ctypes.c_void_p(self.h), ctypes.c_void_p(b.ptr), ctypes.c_void_p(x.ptr),
                                             _p(prm.Iarray), _p(prm.Rarray), _p(hist), hist_len)
        _lib.check(rc, "monolis_hip_solve_R")
        return hist[:hist_len] if hist_len else None

    def solve_refined(self, b: DeviceBuffer, x: DeviceBuffer, prm: Prm, inner_tol: float = 1e-4,
                      max_outer: int = 50) -> dict:
        st = np.zeros(4, np.float64)
        rc = _lib.load().monolis_hip_solve_refined_R(ctypes.c_void_p(self.h), ctypes.c_void_p(b.ptr),
                                                     ctypes.c_void_p(x.ptr), _p(prm.Iarray), _p(prm.Rarray),
                                                     float(inner_tol), int(max_outer), _p(st))
        _lib.check(rc, "monolis_hip_solve_refined_R")
        return {"outer": int(st[0]), "inner_total": int(st[1]), "resid": float(st[2]), "converged": bool(st[3])}

    def solve_multi(self, b: DeviceBuffer, x: DeviceBuffer, nrhs: int, prm: Prm) -> dict:
        """nrhs CG solves sharing each SpMM; b, x hold NP x nrhs doubles, row-major."""
        it = np.zeros(nrhs, np.int32)
        res = np.zeros(nrhs, np.float64)
        st = np.zeros(nrhs, np.int32)
        rc = _lib.load().monolis_hip_solve_multi_R(ctypes.c_void_p(self.h), int(nrhs), ctypes.c_void_p(b.ptr),
                                                   ctypes.c_void_p(x.ptr), _p(prm.Iarray), _p(prm.Rarray),
                                                   _p(it), _p(res), _p(st))
        _lib.check(rc, "monolis_hip_solve_multi_R")
        return {"iters": it, "resid": res, "status": st}

    def autotune_multi(self, nrhs: int, fp32: bool = False, reps: int = 5) -> dict:
        """Times every multi-RHS SpMM configuration and keeps the fastest."""
        L = _lib.load()
        nc = L.monolis_hip_multi_num_configs()
        best = ctypes.c_int()
        ms = (ctypes.c_double * nc)()
        _lib.check(L.monolis_hip_multi_autotune(ctypes.c_void_p(self.h), int(nrhs), int(fp32), int(reps),
                                                ctypes.byref(best), ms), "multi_autotune")
        cfgs = []
        for c in range(nc):
            v = [ctypes.c_int() for _ in range(5)]  # steps per batch, wg/CU, chunk, threads, lds
            L.monolis_hip_multi_config(c, *[ctypes.byref(a) for a in v])
            cfgs.append(tuple(a.value for a in v))
        return {"best": best.value, "ms": list(ms), "configs": cfgs}

    # device-side scalar access (monolis_{set,add,get}_scalar_*_sparse_matrix_R, 0-based)
    def values(self) -> np.ndarray:
        """The device values in monolis's A layout."""
        n = self.mat.nzndof2
        A = np.zeros(max(n, 1), np.float64)
        _lib.check(_lib.load().monolis_hip_matrix_get_values_R(ctypes.c_void_p(self.h), _p(A)), "get_values")
        return A[:n]

    def add_elements(self, elems, Ke) -> None:
        """monolis_add_matrix_to_sparse_matrix_R for every element (rows of `elems`,
        0-based nodes); Ke: (n_elem, n_base*b, n_base*b) element matrices."""
        e = np.ascontiguousarray(elems, np.int32)
        k = np.ascontiguousarray(Ke, np.float64)
        _lib.check(_lib.load().monolis_hip_matrix_add_elements_R(ctypes.c_void_p(self.h), int(e.shape[0]),
                                                                  int(e.shape[1]), _p(e), _p(k)), "add_elements")

    def set_dirichlet(self, b: np.ndarray, node_id, n_dof_bc, val) -> np.ndarray:
        """monolis_set_Dirichlet_bc_R for every listed condition; b (N*b) updated in place."""
        nd = np.ascontiguousarray(node_id, np.int32)
        dd = np.ascontiguousarray(n_dof_bc, np.int32)
        vv = np.ascontiguousarray(val, np.float64)
        assert b.dtype == np.float64 and b.flags.c_contiguous
        _lib.check(_lib.load().monolis_hip_set_Dirichlet_bc_R(ctypes.c_void_p(self.h), _p(b), int(len(nd)), _p(nd),
                                                               _p(dd), _p(vv)), "set_Dirichlet_bc")
        return b

    def set_scalar(self, i: int, j: int, sub_i: int, sub_j: int, val: float) -> None:
        _lib.check(_lib.load().monolis_hip_matrix_set_scalar_R(ctypes.c_void_p(self.h), i, j, sub_i, sub_j,
                                                               float(val)), "set_scalar")

    def add_scalar(self, i: int, j: int, sub_i: int, sub_j: int, val: float) -> None:
        _lib.check(_lib.load().monolis_hip_matrix_add_scalar_R(ctypes.c_void_p(self.h), i, j, sub_i, sub_j,
                                                               float(val)), "add_scalar")

    def get_scalar(self, i: int, j: int, sub_i: int, sub_j: int):
        v, f = ctypes.c_double(), ctypes.c_int()
        _lib.check(_lib.load().monolis_hip_matrix_get_scalar_R(ctypes.c_void_p(self.h), i, j, sub_i, sub_j,
                                                               ctypes.byref(v), ctypes.byref(f)), "get_scalar")
        return v.value, bool(f.value)

    def solve_transpose(self, b: DeviceBuffer, x: DeviceBuffer, prm: Prm) -> None:
        """A^T x = b (device transpose built on first use)."""
        rc = _lib.load().monolis_hip_solve_transpose_R(ctypes.c_void_p(self.h), ctypes.c_void_p(b.ptr),
                                                       ctypes.c_void_p(x.ptr), _p(prm.Iarray), _p(prm.Rarray))
        _lib.check(rc, "monolis_hip_solve_transpose_R")

    def adjoint(self, x: DeviceBuffer, g: DeviceBuffer, lam: DeviceBuffer, dA: Optional[DeviceBuffer], prm: Prm,
                symmetric: bool = False) -> None:
        """lam = A^{-T} g, dA = -lam x^T on the pattern (if dA is given)."""
        rc = _lib.load().monolis_hip_adjoint_R(ctypes.c_void_p(self.h), ctypes.c_void_p(x.ptr),
                                               ctypes.c_void_p(g.ptr), ctypes.c_void_p(lam.ptr),
                                               ctypes.c_void_p(dA.ptr if dA is not None else None),
                                               int(symmetric), _p(prm.Iarray), _p(prm.Rarray))
        _lib.check(rc, "monolis_hip_adjoint_R")

    def matvec_multi(self, x: DeviceBuffer, y: DeviceBuffer, nrhs: int) -> None:
        """Y = A X for nrhs columns (NP x nrhs row-major device buffers)."""
        L = _lib.load()
        fn = L.monolis_hip_matvec_multi_S if x.dtype == np.float32 else L.monolis_hip_matvec_multi_R
        _lib.check(fn(ctypes.c_void_p(self.h), int(nrhs), ctypes.c_void_p(x.ptr), ctypes.c_void_p(y.ptr)),
                   "matvec_multi")

    def set_multi_config(self, cfg: int, fp32: bool = False) -> None:
        _lib.check(_lib.load().monolis_hip_multi_set_config(ctypes.c_void_p(self.h), int(fp32), int(cfg)),
                   "multi_set_config")

    def solve_multi_refined(self, b: DeviceBuffer, x: DeviceBuffer, nrhs: int, prm: Prm,
                            inner_tol: float = 1e-4, max_outer: int = 50) -> dict:
        it = np.zeros(nrhs, np.int32)
        res = np.zeros(nrhs, np.float64)
        st = np.zeros(nrhs, np.int32)
        stats = np.zeros(4, np.float64)
        rc = _lib.load().monolis_hip_solve_multi_refined_R(
            ctypes.c_void_p(self.h), int(nrhs), ctypes.c_void_p(b.ptr), ctypes.c_void_p(x.ptr), _p(prm.Iarray),
            _p(prm.Rarray), float(inner_tol), int(max_outer), _p(it), _p(res), _p(st), _p(stats))
        _lib.check(rc, "monolis_hip_solve_multi_refined_R")
        return {"inner_iters": it, "resid": res, "status": st, "outer": int(stats[0])}

    def autotune(self, fp32: bool = True, reps: int = 20) -> dict:
        L = _lib.load()
        nv = L.monolis_hip_num_variants()
        best = (ctypes.c_int * 2)()
        ms = (ctypes.c_double * (2 * nv))()
        _lib.check(L.monolis_hip_matrix_autotune(ctypes.c_void_p(self.h), int(fp32), int(reps), best, ms), "autotune")
        return {"best64": best[0], "best32": best[1] if fp32 else None,
                "ms64": list(ms)[:nv], "ms32": list(ms)[nv:] if fp32 else None}

    def autotune_solver(self, fp32: bool = False, method: int = 1, precond: int = 1, iters: int = 20,
                        reps: int = 2) -> dict:
        """Times `reps` solves of `iters` iterations per SpMV variant and keeps the fastest
        (ms per iteration in the result)."""
        L = _lib.load()
        nv = L.monolis_hip_num_variants()
        best = ctypes.c_int()
        ms = (ctypes.c_double * nv)()
        _lib.check(L.monolis_hip_matrix_autotune_solver(ctypes.c_void_p(self.h), int(fp32), int(method), int(precond),
                                                        int(iters), int(reps), ctypes.byref(best), ms),
                   "autotune_solver")
        return {"best": best.value, "ms_per_iter": list(ms)}

    def set_variant(self, variant: int, fp32: bool = False) -> None:
        _lib.check(_lib.load().monolis_hip_matrix_set_variant(ctypes.c_void_p(self.h), int(fp32), int(variant)),
                   "set_variant")

    def time_matvec(self, x: DeviceBuffer, y: DeviceBuffer, warmup: int = 5, reps: int = 50) -> float:
        ms = ctypes.c_double()
        fp32 = 1 if x.dtype == np.float32 else 0
        _lib.check(_lib.load().monolis_hip_time_matvec(ctypes.c_void_p(self.h), fp32, ctypes.c_void_p(x.ptr),
                                                       ctypes.c_void_p(y.ptr), warmup, reps, ctypes.byref(ms)),
                   "time_matvec")
        return ms.value

    def time_matvec_each(self, x: DeviceBuffer, y: DeviceBuffer, warmup: int = 5, reps: int = 50) -> np.ndarray:
        """Milliseconds of each of `reps` SpMVs (one event pair around each)."""
        out = np.zeros(max(reps, 1), np.float64)
        fp32 = 1 if x.dtype == np.float32 else 0
        _lib.check(_lib.load().monolis_hip_time_matvec_each(ctypes.c_void_p(self.h), fp32, ctypes.c_void_p(x.ptr),
                                                            ctypes.c_void_p(y.ptr), warmup, reps, _p(out)),
                   "time_matvec_each")
        return out[:reps]


def device_matrix_by_simple_mesh(n_node: int, n_dof: int, elems) -> "DeviceMatrix":
    """monolis_get_nonzero_pattern_by_simple_mesh_R on the GPU: a DeviceMatrix
    with the mesh's pattern and zero values (elems: 0-based node ids, one row
    per element)."""
    L = _lib.load()
    e = np.ascontiguousarray(elems, np.int32)
    n_elem, n_base = (e.shape[0], e.shape[1]) if e.ndim == 2 else (0, 1)
    h = ctypes.c_void_p()
    _lib.check(L.monolis_hip_matrix_create_by_simple_mesh(ctypes.byref(h), int(n_node), int(n_base), int(n_dof),
                                                          int(n_elem), _p(e)), "create_by_simple_mesh")
    info = (ctypes.c_longlong * 8)()
    _lib.check(L.monolis_hip_matrix_info(h, info), "info")
    nz = int(info[2])
    index = np.zeros(n_node + 1, np.int32)
    item = np.zeros(max(nz, 1), np.int32)
    _lib.check(L.monolis_hip_matrix_get_pattern(h, _p(index), _p(item)), "get_pattern")
    mat = Matrix(n_node, n_node, n_dof, index, item[:nz], np.zeros(nz * n_dof * n_dof)).ensure_dof_index()
    dm = DeviceMatrix.__new__(DeviceMatrix)
    dm.h, dm.mat, dm.m, dm.vec = h.value, mat, mat.nndof, mat.npndof
    return dm


class DeviceMatrixC:
    """A complex monolis BCSR matrix (complex128 values) resident in HBM."""

    def __init__(self, mat: Matrix):
        L = _lib.load()
        h = ctypes.c_void_p()
        mat.ensure_dof_index()
        A = np.ascontiguousarray(mat.A, np.complex128)
        if mat.ndof == -1:
            _lib.check(L.monolis_hip_matrix_create_V_C(
                ctypes.byref(h), mat.n, mat.np_, mat.nz, _p(np.ascontiguousarray(mat.n_dof_list, np.int32)),
                _p(np.ascontiguousarray(mat.n_dof_index, np.int32)),
                _p(np.ascontiguousarray(mat.n_dof_index2, np.int32)), _p(mat.index), _p(mat.item), _p(A)),
                "monolis_hip_matrix_create_V_C")
        else:
            _lib.check(L.monolis_hip_matrix_create_C(ctypes.byref(h), mat.n, mat.np_, mat.nz, mat.ndof,
                                                     _p(mat.index), _p(mat.item), _p(A)),
                       "monolis_hip_matrix_create_C")
        self.h = h.value
        self.mat = mat
        self.m = mat.nndof
        self.vec = mat.npndof

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().monolis_hip_matrix_destroy_C(ctypes.c_void_p(self.h))
                self.h = None
        except Exception:
            pass

    def info(self) -> dict:
        a = (ctypes.c_longlong * 8)()
        _lib.check(_lib.load().monolis_hip_matrix_info_C(ctypes.c_void_p(self.h), a), "info_C")
        keys = ["n", "np", "nz", "ndof", "slices", "slots", "bytes", "unused"]
        return dict(zip(keys, list(a)))

    def set_values(self, A) -> None:
        if isinstance(A, DeviceBuffer):
            ptr = ctypes.c_void_p(A.ptr)
        else:
            A = np.ascontiguousarray(A, np.complex128)
            ptr = _p(A)
        _lib.check(_lib.load().monolis_hip_matrix_set_values_C(ctypes.c_void_p(self.h), ptr), "set_values_C")

    def matvec(self, x: DeviceBuffer, y: DeviceBuffer) -> None:
        _lib.check(_lib.load().monolis_hip_matvec_C(ctypes.c_void_p(self.h), ctypes.c_void_p(x.ptr),
                                                    ctypes.c_void_p(y.ptr)), "matvec_C")

    def solve(self, b: DeviceBuffer, x: DeviceBuffer, prm: Prm, hist_len: int = 0):
        hist = np.zeros(max(hist_len, 1), np.float64)
        rc = _lib.load().monolis_hip_solve_C(ctypes.c_void_p(self.h), ctypes.c_void_p(b.ptr), ctypes.c_void_p(x.ptr),
                                             _p(prm.Iarray), _p(prm.Rarray), _p(hist), hist_len)
        _lib.check(rc, "monolis_hip_solve_C")
        return hist[:hist_len] if hist_len else None

    def time_matvec(self, x: DeviceBuffer, y: DeviceBuffer, warmup: int = 5, reps: int = 50) -> float:
        ms = ctypes.c_double()
        _lib.check(_lib.load().monolis_hip_time_matvec_C(ctypes.c_void_p(self.h), ctypes.c_void_p(x.ptr),
                                                         ctypes.c_void_p(y.ptr), warmup, reps, ctypes.byref(ms)),
                   "time_matvec_C")
        return ms.value


    # device assembly (monolis_add_matrix_to_sparse_matrix_C, monolis_set_Dirichlet_bc_C, 0-based)
    def values(self) -> np.ndarray:
        """The device values in monolis's A layout (complex128)."""
        n = self.mat.nzndof2
        A = np.zeros(max(n, 1), np.complex128)
        _lib.check(_lib.load().monolis_hip_matrix_get_values_C(ctypes.c_void_p(self.h), _p(A)), "get_values_C")
        return A[:n]

    def add_elements(self, elems, Ke) -> None:
        """monolis_add_matrix_to_sparse_matrix_C for every element; Ke: complex
        (n_elem, n_base*b, n_base*b) element matrices."""
        e = np.ascontiguousarray(elems, np.int32)
        k = np.ascontiguousarray(Ke, np.complex128)
        _lib.check(_lib.load().monolis_hip_matrix_add_elements_C(ctypes.c_void_p(self.h), int(e.shape[0]),
                                                                  int(e.shape[1]), _p(e), _p(k)), "add_elements_C")

    def set_dirichlet(self, b: np.ndarray, node_id, n_dof_bc, val) -> np.ndarray:
        """monolis_set_Dirichlet_bc_C for every listed condition; b (complex128, N*b) in place."""
        nd = np.ascontiguousarray(node_id, np.int32)
        dd = np.ascontiguousarray(n_dof_bc, np.int32)
        vv = np.ascontiguousarray(val, np.complex128)
        assert b.dtype == np.complex128 and b.flags.c_contiguous
        _lib.check(_lib.load().monolis_hip_set_Dirichlet_bc_C(ctypes.c_void_p(self.h), _p(b), int(len(nd)), _p(nd),
                                                               _p(dd), _p(vv)), "set_Dirichlet_bc_C")
        return b


def device_matrix_by_simple_mesh_V(n_node: int, n_dof_list, elems, complex_values: bool = False):
    """monolis_get_nonzero_pattern_by_simple_mesh_V(_C) on the GPU: a DeviceMatrix
    (DeviceMatrixC with complex_values) with the mesh's pattern, a block size per
    node and zero values."""
    from .matrix import nonzero_pattern_by_simple_mesh_V
    L = _lib.load()
    e = np.ascontiguousarray(elems, np.int32)
    n_elem, n_base = (e.shape[0], e.shape[1]) if e.ndim == 2 else (0, 1)
    lst = np.ascontiguousarray(n_dof_list, np.int32)
    h = ctypes.c_void_p()
    fn = L.monolis_hip_matrix_create_by_simple_mesh_V_C if complex_values else L.monolis_hip_matrix_create_by_simple_mesh_V
    _lib.check(fn(ctypes.byref(h), int(n_node), int(n_base), _p(lst), int(n_elem), _p(e)), "create_by_simple_mesh_V")
    mat = nonzero_pattern_by_simple_mesh_V(n_node, lst, e)   # host restatement: index, item, dof arrays
    if complex_values:
        mat.A = mat.A.astype(np.complex128)
        dm = DeviceMatrixC.__new__(DeviceMatrixC)
    else:
        dm = DeviceMatrix.__new__(DeviceMatrix)
    dm.h, dm.mat, dm.m, dm.vec = h.value, mat, mat.nndof, mat.npndof
    return dm


def device_matrix_by_simple_mesh_C(n_node: int, n_dof: int, elems) -> "DeviceMatrixC":
    """monolis_get_nonzero_pattern_by_simple_mesh_C on the GPU: a DeviceMatrixC with
    the mesh's pattern and zero values."""
    L = _lib.load()
    e = np.ascontiguousarray(elems, np.int32)
    n_elem, n_base = (e.shape[0], e.shape[1]) if e.ndim == 2 else (0, 1)
    h = ctypes.c_void_p()
    _lib.check(L.monolis_hip_matrix_create_by_simple_mesh_C(ctypes.byref(h), int(n_node), int(n_base), int(n_dof),
                                                            int(n_elem), _p(e)), "create_by_simple_mesh_C")
    info = (ctypes.c_longlong * 8)()
    _lib.check(L.monolis_hip_matrix_info_C(h, info), "info_C")
    nz = int(info[2])
    index = np.zeros(n_node + 1, np.int32)
    item = np.zeros(max(nz, 1), np.int32)
    _lib.check(L.monolis_hip_matrix_get_pattern_C(h, _p(index), _p(item)), "get_pattern_C")
    mat = Matrix(n_node, n_node, n_dof, index, item[:nz], np.zeros(nz * n_dof * n_dof, np.complex128))
    mat.ensure_dof_index()
    dm = DeviceMatrixC.__new__(DeviceMatrixC)
    dm.h, dm.mat, dm.m, dm.vec = h.value, mat, mat.nndof, mat.npndof
    return dm


def inner_product_device_C(x: DeviceBuffer, y: DeviceBuffer, m: Optional[int] = None) -> complex:
    """Unconjugated (x, y) of two complex device buffers."""
    n = x.count if m is None else m
    s = np.zeros(1, np.complex128)
    _lib.check(_lib.load().monolis_hip_inner_product_C(n, ctypes.c_void_p(x.ptr), ctypes.c_void_p(y.ptr), _p(s)),
               "inner_product_C")
    return complex(s[0])


def inner_product_device(x: DeviceBuffer, y: DeviceBuffer, m: Optional[int] = None) -> float:
    s = ctypes.c_double()
    n = x.count if m is None else m
    _lib.check(_lib.load().monolis_hip_inner_product_R(n, ctypes.c_void_p(x.ptr), ctypes.c_void_p(y.ptr),
                                                       ctypes.byref(s)), "inner_product")
    return s.value


class DeviceBatch:
    """Many independent systems (block-diagonal scalar CSR, system k owns rows
    sys_row[k]..sys_row[k+1]-1) solved by one persistent, graph-captured launch."""

    def __init__(self, mat: Matrix, sys_row):
        L = _lib.load()
        h = ctypes.c_void_p()
        self.sys_row = np.ascontiguousarray(sys_row, np.int32)
        self.nsys = len(self.sys_row) - 1
        _lib.check(L.monolis_hip_batch_create(ctypes.byref(h), self.nsys, _p(self.sys_row), mat.n, mat.nz,
                                              _p(mat.index), _p(mat.item), _p(mat.A)), "monolis_hip_batch_create")
        self.h = h.value
        self.n = mat.n

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _lib.load().monolis_hip_batch_destroy(ctypes.c_void_p(self.h))
                self.h = None
        except Exception:
            pass

    def solve(self, b: DeviceBuffer, x: DeviceBuffer, prm: Prm):
        iters = np.zeros(self.nsys, np.int32)
        resid = np.zeros(self.nsys, np.float64)
        status = np.zeros(self.nsys, np.int32)
        _lib.check(_lib.load().monolis_hip_batch_solve(ctypes.c_void_p(self.h), ctypes.c_void_p(b.ptr),
                                                       ctypes.c_void_p(x.ptr), _p(prm.Iarray), _p(prm.Rarray),
                                                       _p(iters), _p(resid), _p(status)), "batch_solve")
        return iters, resid, status

    def autotune(self, b: DeviceBuffer, x: DeviceBuffer, prm: Prm, reps: int = 2) -> dict:
        L = _lib.load()
        nc = L.monolis_hip_batch_num_configs()
        best = ctypes.c_int()
        ms = (ctypes.c_double * nc)()
        _lib.check(L.monolis_hip_batch_autotune(ctypes.c_void_p(self.h), ctypes.c_void_p(b.ptr), ctypes.c_void_p(x.ptr),
                                                _p(prm.Iarray), _p(prm.Rarray), int(reps), ctypes.byref(best), ms),
                   "batch_autotune")
        cfgs = []
        for c in range(nc):
            t, u = ctypes.c_int(), ctypes.c_int()
            L.monolis_hip_batch_config(c, ctypes.byref(t), ctypes.byref(u))
            cfgs.append({"threads": t.value, "unroll": u.value, **self.schedule(c)})
        return {"best": best.value, "ms": list(ms), "configs": cfgs}

    def set_config(self, cfg: int):
        _lib.check(_lib.load().monolis_hip_batch_set_config(ctypes.c_void_p(self.h), int(cfg)), "batch_set_config")

    def schedule(self, cfg: int) -> dict:
        """Team split of configuration cfg and its schedule's slot / team counts."""
        sp, ns, nt = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _lib.check(_lib.load().monolis_hip_batch_schedule(ctypes.c_void_p(self.h), int(cfg), ctypes.byref(sp),
                                                          ctypes.byref(ns), ctypes.byref(nt)), "batch_schedule")
        return {"split": sp.value, "slots": ns.value, "teams": nt.value}

    def set_wait_limit(self, seconds: float):
        _lib.check(_lib.load().monolis_hip_batch_set_wait_limit(ctypes.c_void_p(self.h), float(seconds)),
                   "batch_set_wait_limit")

    def time(self, b: DeviceBuffer, x: DeviceBuffer, prm: Prm, reps: int = 5) -> float:
        ms = ctypes.c_double()
        _lib.check(_lib.load().monolis_hip_batch_time(ctypes.c_void_p(self.h), ctypes.c_void_p(b.ptr),
                                                      ctypes.c_void_p(x.ptr), _p(prm.Iarray), _p(prm.Rarray),
                                                      int(reps), ctypes.byref(ms)), "batch_time")
        return ms.value
