"""Synthetic structured-grid systems for parity tests and the benchmark.

All generators are deterministic functions of their seed (numpy PCG64 via
``default_rng``) and vectorised, so the 100^3-node BCSR3 system builds in a few
seconds on the host.  Node ids are z-major (id = z*n*n + y*n + x); a stencil
offset list in lexicographic (dz, dy, dx) order is also ascending column order,
which gives the sorted rows monolis requires
(src/matrix/spmat_nzpattern_util.f90:44-48).  Out-of-grid neighbours are
dropped (homogeneous Dirichlet boundary).

Workloads (sizes as used by the tests and bench.py):

* ``poisson7``         3-D 7-point Laplacian, diagonal 6, off-diagonal -1, CSR.
                       32^3 = 32,768 DOF is the CPU-runnable parity case.
* ``stencil27_bcsr3``  27-point, 3 DOF per node, BCSR 3x3.  Off-diagonal blocks
                       U[-1,0] with A_ji = A_ij^T; the diagonal block has
                       symmetric in-block off-diagonals U[-1,0] and diagonal
                       entries = sum of |off-diagonals| in that scalar row + 1
                       (strictly diagonally dominant => SPD; block-Jacobi
                       differs from point-Jacobi).  100^3 nodes = 3.0 M DOF,
                       nnzb = (3n-2)^3 = 26,463,592.
* ``convdiff7``        7-point convection-diffusion, flow along +x, central
                       differences: west -1 - Pe/2, east -1 + Pe/2, other
                       neighbours -1, diagonal 6 (nonsymmetric).
* ``scale_symmetric``  D A D with D log-uniform over [1, dmax] per DOF.
* ``poisson_batch``    many independent 7-point systems with random edges.
* ``helmholtz7``       complex symmetric 7-point damped Helmholtz operator
                       L - s (1 + i eta) I (L the 7-point Laplacian), CSR.
* ``stencil27_bcsr3_complex``  K + i g D on the stencil27_bcsr3 pattern: K the
                       real SPD system, D its diagonal blocks (a damped
                       dynamic stiffness; complex symmetric).
"""
from __future__ import annotations

import numpy as np

from .matrix import Matrix

_OFF7 = [(-1, 0, 0), (0, -1, 0), (0, 0, -1), (0, 0, 0), (0, 0, 1), (0, 1, 0), (1, 0, 0)]
_OFF27 = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def _grid_pattern(n: int, offsets):
    """CSR pattern of a stencil on an n^3 grid.  Returns (index, item, pos, valid, lin)
    where pos[p, k] is the storage slot of neighbour k of node p."""
    N = n * n * n
    ids = np.arange(N, dtype=np.int64)
    z, rem = np.divmod(ids, n * n)
    y, x = np.divmod(rem, n)
    K = len(offsets)
    valid = np.empty((N, K), dtype=bool)
    lin = np.empty(K, dtype=np.int64)
    for k, (dz, dy, dx) in enumerate(offsets):
        valid[:, k] = ((z + dz >= 0) & (z + dz < n) & (y + dy >= 0) & (y + dy < n)
                       & (x + dx >= 0) & (x + dx < n))
        lin[k] = dz * n * n + dy * n + dx
    rowlen = valid.sum(axis=1)
    index = np.zeros(N + 1, dtype=np.int64)
    np.cumsum(rowlen, out=index[1:])
    pos = np.cumsum(valid, axis=1, dtype=np.int64)
    pos += index[:-1, None] - 1
    nnz = int(index[-1])
    item = np.empty(nnz, dtype=np.int32)
    for k in range(K):
        v = valid[:, k]
        item[pos[v, k]] = (ids[v] + lin[k] + 1).astype(np.int32)
    return index.astype(np.int32), item, pos, valid, lin


def poisson7(n: int) -> Matrix:
    """Scalar CSR 7-point Laplacian (diag 6, off -1) on n^3 nodes."""
    index, item, pos, valid, _ = _grid_pattern(n, _OFF7)
    A = np.where(item == np.repeat(np.arange(1, n ** 3 + 1, dtype=np.int32), np.diff(index)), 6.0, -1.0)
    m = Matrix(n ** 3, n ** 3, 1, index, item, A.astype(np.float64))
    m.meta = {"workload": f"poisson7_{n}^3", "grid": n}
    return m.ensure_dof_index()


def convdiff7(n: int, peclet: float = 0.5) -> Matrix:
    """Nonsymmetric 7-point convection-diffusion (flow along +x), scalar CSR."""
    index, item, pos, valid, _ = _grid_pattern(n, _OFF7)
    coef = [-1.0, -1.0, -1.0 - 0.5 * peclet, 6.0, -1.0 + 0.5 * peclet, -1.0, -1.0]
    A = np.empty(len(item), dtype=np.float64)
    for k, c in enumerate(coef):
        A[pos[valid[:, k], k]] = c
    m = Matrix(n ** 3, n ** 3, 1, index, item, A)
    m.meta = {"workload": f"convdiff7_{n}^3_pe{peclet}", "grid": n, "peclet": peclet}
    return m.ensure_dof_index()


def stencil27_bcsr3(n: int, seed: int = 0) -> Matrix:
    """27-point BCSR3 SPD system on n^3 nodes (3 DOF per node)."""
    index, item, pos, valid, lin = _grid_pattern(n, _OFF27)
    N = n ** 3
    nnz = len(item)
    A = np.zeros((nnz, 3, 3), dtype=np.float64)
    rowsum = np.zeros((N, 3), dtype=np.float64)
    rng = np.random.default_rng(seed)
    ids = np.arange(N, dtype=np.int64)
    for k in range(14, 27):             # offsets with positive linear distance
        kr = 26 - k                     # the mirrored offset
        v = valid[:, k]
        p = ids[v]
        q = p + lin[k]
        R = rng.uniform(-1.0, 0.0, size=(len(p), 3, 3))
        A[pos[p, k]] = R
        A[pos[q, kr]] = R.transpose(0, 2, 1)
        aR = np.abs(R)
        rowsum[p] += aR.sum(axis=2)
        rowsum[q] += aR.sum(axis=1)
        del R, aR, p, q, v
    od = rng.uniform(-1.0, 0.0, size=(N, 3))
    Dg = np.zeros((N, 3, 3), dtype=np.float64)
    Dg[:, 0, 1] = Dg[:, 1, 0] = od[:, 0]
    Dg[:, 0, 2] = Dg[:, 2, 0] = od[:, 1]
    Dg[:, 1, 2] = Dg[:, 2, 1] = od[:, 2]
    rowsum += np.abs(Dg).sum(axis=2)
    for r in range(3):
        Dg[:, r, r] = rowsum[:, r] + 1.0
    A[pos[ids, 13]] = Dg
    m = Matrix(N, N, 3, index, item, A.reshape(-1))
    m.meta = {"workload": f"stencil27_bcsr3_{n}^3", "grid": n, "seed": seed}
    return m.ensure_dof_index()


def scale_symmetric(mat: Matrix, dmax: float = 1e3, seed: int = 1) -> Matrix:
    """Return D A D with D = diag(10^U[0, log10 dmax]) per DOF (ill-conditioned twin)."""
    rng = np.random.default_rng(seed)
    b = mat.ndof
    d = 10.0 ** rng.uniform(0.0, np.log10(dmax), size=mat.n * b)
    rows = np.repeat(np.arange(mat.n, dtype=np.int64), np.diff(mat.index))
    cols = mat.item.astype(np.int64) - 1
    if b == 1:
        A = mat.A * d[rows] * d[cols]
    else:
        blk = mat.A.reshape(-1, b, b)
        dr = d.reshape(-1, b)[rows]
        dc = d.reshape(-1, b)[cols]
        A = (blk * dr[:, :, None] * dc[:, None, :]).reshape(-1)
    out = Matrix(mat.n, mat.np_, b, mat.index, mat.item, A)
    out.meta = dict(mat.meta, workload=mat.meta.get("workload", "") + f"_DAD{dmax:g}", scale_seed=seed)
    out.meta["scale"] = d
    return out.ensure_dof_index()


def uniform_vector(m: int, seed: int, lo: float = -1.0, hi: float = 1.0) -> np.ndarray:
    return np.random.default_rng(seed).uniform(lo, hi, size=m)


def rhs_from_solution(mat: Matrix, x_true: np.ndarray) -> np.ndarray:
    """b = A x_true on the host (scipy), for building synthetic systems."""
    return np.asarray(mat.to_scipy() @ x_true, dtype=np.float64)


def poisson_batch(count: int = 512, lo: int = 10, hi: int = 40, seed: int = 0, edges=None):
    """`count` independent 7-point Poisson cubes with edges U{lo..hi} (or the
    given `edges`), packed as one block-diagonal CSR.  Returns (Matrix, row_ptr)
    where system s owns rows row_ptr[s]..row_ptr[s+1]-1."""
    if edges is None:
        rng = np.random.default_rng(seed)
        edges = rng.integers(lo, hi + 1, size=count)
    else:
        edges = np.asarray(edges, dtype=np.int64)
        count, lo, hi = len(edges), int(edges.min()), int(edges.max())
    cache = {}
    idx_parts, item_parts, a_parts = [], [], []
    row_ptr = np.zeros(count + 1, dtype=np.int64)
    nnz_off = 0
    row_off = 0
    for s, e in enumerate(edges):
        e = int(e)
        if e not in cache:
            cache[e] = poisson7(e)
        sub = cache[e]
        idx_parts.append(sub.index[:-1].astype(np.int64) + nnz_off)
        item_parts.append(sub.item.astype(np.int64) + row_off)
        a_parts.append(sub.A)
        nnz_off += len(sub.item)
        row_off += sub.n
        row_ptr[s + 1] = row_off
    index = np.concatenate(idx_parts + [np.array([nnz_off])]).astype(np.int32)
    item = np.concatenate(item_parts).astype(np.int32)
    A = np.concatenate(a_parts)
    m = Matrix(row_off, row_off, 1, index, item, A)
    m.meta = {"workload": f"poisson_batch_{count}x[{lo}..{hi}]^3", "edges": edges, "seed": seed}
    return m.ensure_dof_index(), row_ptr


def helmholtz7(n: int, shift: float = 0.5, damping: float = 0.2) -> Matrix:
    """Complex symmetric damped Helmholtz on n^3 nodes: 7-point Laplacian with
    diagonal 6 - shift (1 + i damping), off-diagonals -1 (complex128 values)."""
    m = poisson7(n)
    diag = m.item == np.repeat(np.arange(1, n ** 3 + 1, dtype=np.int32), np.diff(m.index))
    A = m.A.astype(np.complex128)
    A[diag] -= shift * (1.0 + 1j * damping)
    out = Matrix(m.n, m.np_, 1, m.index, m.item, A)
    out.meta = {"workload": f"helmholtz7_{n}^3_s{shift:g}_eta{damping:g}", "grid": n}
    return out.ensure_dof_index()


def stencil27_bcsr3_complex(n: int, seed: int = 0, damping: float = 0.5) -> Matrix:
    """K + i damping D with K = stencil27_bcsr3(n, seed) and D its diagonal blocks."""
    m = stencil27_bcsr3(n, seed)
    N = n ** 3
    A = m.A.astype(np.complex128)
    rows = np.repeat(np.arange(1, N + 1, dtype=np.int32), np.diff(m.index))
    dpos = np.nonzero(m.item == rows)[0]
    del rows
    blk = A.reshape(-1, 9)
    blk[dpos] += 1j * damping * m.A.reshape(-1, 9)[dpos]
    out = Matrix(N, N, 3, m.index, m.item, A)
    out.meta = {"workload": f"stencil27_bcsr3_complex_{n}^3", "grid": n, "seed": seed, "damping": damping}
    return out.ensure_dof_index()
