"""Host-side BCSR container and assembly, mirroring monolis's matrix layer.

The container holds exactly the arrays monolis's flat C entry points take
(``monolis_solve_R_c_main``, wrapper/solver/monolis_solver_c.h:51-80):

* ``index``  int32, NP+1, 0-based row offsets (src/define/def_mat.f90:185-216);
* ``item``   int32, NZ, **1-based** block columns, diagonal included, ascending
  per row (src/matrix/spmat_nzpattern_util.f90:34-48; the C API stores item+1,
  wrapper/matrix/monolis_spmat_handler_c.c:232);
* ``A``      float64, NZ*b*b (fixed b) or n_dof_index2[NZ] (variable b), block
  j row k col l at ``b*b*j + b*k + l`` -- row-major inside a block
  (src/linalg/matvec.f90:392);
* ``n_dof_list`` / ``n_dof_index`` / ``n_dof_index2`` prefix sums
  (src/matrix/spmat_nzpattern_util.f90:62-90).

Assembly follows ``monolis_get_nonzero_pattern_by_simple_mesh_R``
(src/matrix/spmat_nzpattern.f90:15-49): two nodes are coupled when they share
an element, every row holds its diagonal, columns are sorted.  Scalar add/set
use the same block addressing as src/matrix/spmat_handler_util.f90:306-348.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional

import numpy as np


@dataclass
class Matrix:
    n: int                      # internal block rows (N)
    np_: int                    # all block rows incl. halo (NP); == n on one rank
    ndof: int                   # block size b, or -1 for variable blocks
    index: np.ndarray           # int32 [NP+1]
    item: np.ndarray            # int32 [NZ], 1-based
    A: np.ndarray               # float64 (complex128 for the complex path)
    n_dof_list: Optional[np.ndarray] = None
    n_dof_index: Optional[np.ndarray] = None
    n_dof_index2: Optional[np.ndarray] = None
    meta: dict = field(default_factory=dict)

    @property
    def nz(self) -> int:
        return int(self.index[self.np_])

    @property
    def nndof(self) -> int:
        """Vector length over internal rows (monolis_get_vec_size, def_mat.f90:699-721)."""
        if self.ndof == -1:
            return int(self.n_dof_index[self.n])
        return self.n * self.ndof

    @property
    def npndof(self) -> int:
        if self.ndof == -1:
            return int(self.n_dof_index[self.np_])
        return self.np_ * self.ndof

    @property
    def nzndof2(self) -> int:
        if self.ndof == -1:
            return int(self.n_dof_index2[self.nz])
        return self.nz * self.ndof * self.ndof

    def ensure_dof_index(self) -> "Matrix":
        """Fill n_dof_list / n_dof_index / n_dof_index2 for a fixed block size
        (monolis_set_n_dof_index, src/matrix/spmat_nzpattern_util.f90:62-90)."""
        if self.n_dof_list is None:
            b = self.ndof
            self.n_dof_list = np.full(self.np_, b, np.int32)
            self.n_dof_index = (np.arange(self.np_ + 1, dtype=np.int64) * b).astype(np.int32)
            nz = self.nz
            if nz * b * b < 2**31:
                self.n_dof_index2 = (np.arange(nz + 1, dtype=np.int64) * b * b).astype(np.int32)
            else:
                # int32 prefix sums overflow at this size; a fixed block size
                # never reads n_dof_index2, so keep a placeholder.
                self.n_dof_index2 = np.zeros(nz + 1, np.int32)
        return self

    def to_scipy(self):
        """scipy.sparse view of the internal rows (workload synthesis only)."""
        import scipy.sparse as sp
        cols = self.item.astype(np.int64) - 1
        if self.ndof == 1:
            return sp.csr_matrix((self.A, cols, self.index.astype(np.int64)), shape=(self.n, self.np_))
        if self.ndof > 1:
            b = self.ndof
            return sp.bsr_matrix((self.A.reshape(-1, b, b), cols, self.index.astype(np.int64)),
                                 shape=(self.n * b, self.np_ * b))
        raise ValueError("variable-block matrices have no scipy view")

    def dense(self) -> np.ndarray:
        """Dense copy (tiny matrices only: brute-force checks)."""
        if self.ndof == -1:
            m = int(self.n_dof_index[self.np_])
            D = np.zeros((int(self.n_dof_index[self.n]), m), self.A.dtype)
            for i in range(self.n):
                n1 = self.n_dof_list[i]
                r0 = self.n_dof_index[i]
                for j in range(self.index[i], self.index[i + 1]):
                    c = self.item[j] - 1
                    n2 = self.n_dof_list[c]
                    c0 = self.n_dof_index[c]
                    a0 = self.n_dof_index2[j]
                    D[r0:r0 + n1, c0:c0 + n2] = self.A[a0:a0 + n1 * n2].reshape(n1, n2)
            return D
        return self.to_scipy().toarray()

    def astype_values(self, dtype) -> np.ndarray:
        return self.A.astype(dtype)


def nonzero_pattern_by_simple_mesh(n_node: int, ndof: int, elems, dtype=np.float64) -> Matrix:
    """monolis_get_nonzero_pattern_by_simple_mesh_R for elements given as
    0-based node lists (the C API's convention; node ids become 1-based
    columns in ``item``).  ``dtype=np.complex128`` gives the _C variant
    (src/matrix/spmat_nzpattern.f90:53-87)."""
    nbr = [set() for _ in range(n_node)]
    for e in elems:
        e = [int(v) for v in e]
        for a in e:
            for b in e:
                if a != b:
                    nbr[a].add(b)
    rows = [sorted(nbr[i] | {i}) for i in range(n_node)]
    index = np.zeros(n_node + 1, np.int32)
    index[1:] = np.cumsum([len(r) for r in rows])
    item = np.array([c + 1 for r in rows for c in r], np.int32)
    A = np.zeros(int(index[-1]) * ndof * ndof, dtype)
    return Matrix(n_node, n_node, ndof, index, item, A).ensure_dof_index()


def nonzero_pattern_by_simple_mesh_V(n_node: int, n_dof_list, elems) -> Matrix:
    """monolis_get_nonzero_pattern_by_simple_mesh_V_R: the pattern of
    nonzero_pattern_by_simple_mesh with a block size per node; n_dof_index and
    n_dof_index2 are the prefix sums of monolis_set_n_dof_index
    (src/matrix/spmat_nzpattern_util.f90:62-90)."""
    m = nonzero_pattern_by_simple_mesh(n_node, 1, elems)
    lst = np.ascontiguousarray(n_dof_list, np.int32)
    dofi = np.zeros(n_node + 1, np.int32)
    dofi[1:] = np.cumsum(lst)
    rows = np.repeat(np.arange(n_node), np.diff(m.index))
    sizes = lst[rows].astype(np.int64) * lst[m.item - 1]
    idx2 = np.zeros(len(m.item) + 1, np.int32)
    idx2[1:] = np.cumsum(sizes)
    return Matrix(n_node, n_node, -1, m.index, m.item, np.zeros(int(idx2[-1]), np.float64),
                  n_dof_list=lst, n_dof_index=dofi, n_dof_index2=idx2)


def _find_block(mat: Matrix, i: int, j: int) -> int:
    """Binary search of 1-based column j+1 inside row i (spmat_handler_util.f90:339)."""
    lo, hi = int(mat.index[i]), int(mat.index[i + 1])
    k = lo + int(np.searchsorted(mat.item[lo:hi], j + 1))
    if k >= hi or mat.item[k] != j + 1:
        raise KeyError(f"block ({i},{j}) is not in the nonzero pattern")
    return k


def add_scalar(mat: Matrix, i: int, j: int, si: int, sj: int, val: float) -> None:
    """monolis_add_scalar_to_sparse_matrix_R with 0-based (i, j, si, sj) as in the C API."""
    k = _find_block(mat, i, j)
    b = mat.ndof
    mat.A[b * b * k + b * si + sj] += val


def set_scalar(mat: Matrix, i: int, j: int, si: int, sj: int, val: float) -> None:
    k = _find_block(mat, i, j)
    b = mat.ndof
    mat.A[b * b * k + b * si + sj] = val


def get_scalar(mat: Matrix, i: int, j: int, si: int, sj: int) -> float:
    k = _find_block(mat, i, j)
    b = mat.ndof
    return float(mat.A[b * b * k + b * si + sj])


def from_bcsr(n: int, ndof: int, index, item_1based, A) -> Matrix:
    """Wrap caller arrays (monolis_set_matrix_BCSR_R, src/matrix/spmat_handler.f90:265-333)."""
    return Matrix(n, n, ndof, np.ascontiguousarray(index, np.int32),
                  np.ascontiguousarray(item_1based, np.int32),
                  np.ascontiguousarray(A, np.float64)).ensure_dof_index()
