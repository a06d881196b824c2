"""Shared test helpers: build reference test matrices from the golden fixtures."""
import json
import os

import numpy as np

from monolis_amd import matrix as mm

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "monolis_kat.json")


def kat():
    with open(GOLDEN) as f:
        return json.load(f)


def from_mesh_case(case):
    m = mm.nonzero_pattern_by_simple_mesh(case["n_node"], case["ndof"], case["elems"])
    for i, j, si, sj, v in case.get("add", []):
        mm.add_scalar(m, i, j, si, sj, v)
    for i, j, si, sj, v in case.get("set", []):
        mm.set_scalar(m, i, j, si, sj, v)
    return m


def from_mesh_case_complex(case):
    """Complex matrix of a fixture case whose "set_complex" rows are (i, j, si, sj, re, im)."""
    m = mm.nonzero_pattern_by_simple_mesh(case["n_node"], case["ndof"], case["elems"], dtype=np.complex128)
    for i, j, si, sj, re, im in case["set_complex"]:
        mm.set_scalar(m, i, j, si, sj, complex(re, im))
    return m


def mtx_matrix(case):
    """Symmetric matrix of mtx.dat as main.f90:176-190 assembles it."""
    n = case["n_node"]
    elems = [[i, j] for i, j, _ in case["entries"] if i != j]
    m = mm.nonzero_pattern_by_simple_mesh(n, 1, elems)
    for i, j, v in case["entries"]:
        mm.add_scalar(m, i, j, 0, 0, v)
        if i != j:
            mm.add_scalar(m, j, i, 0, 0, v)
    return m


def chain_matrix(n_node, ndof, seed):
    """Shape of src_test/iterative/CG_test.f90:40-70: 10-node chain, diagonal
    U[0,1)+2 on the block diagonals, symmetric U[0,1) off-diagonal blocks."""
    rng = np.random.default_rng(seed)
    elems = [[i, i + 1] for i in range(n_node - 1)]
    m = mm.nonzero_pattern_by_simple_mesh(n_node, ndof, elems)
    for i in range(n_node):
        for d in range(ndof):
            mm.add_scalar(m, i, i, d, d, rng.random() + 2.0)
    for a, b in elems:
        for i2 in range(ndof):
            for j2 in range(ndof):
                v = rng.random()
                mm.add_scalar(m, a, b, i2, j2, v)
                mm.add_scalar(m, b, a, j2, i2, v)
    return m


def random_bcsr(n, ndof, density, seed, spd=False, diag_missing=()):
    """Random block pattern (diagonal included unless listed) with random values."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        cols = set(np.nonzero(rng.random(n) < density)[0].tolist())
        if i not in diag_missing:
            cols.add(i)
        else:
            cols.discard(i)
        if spd:
            pass
        rows.append(sorted(cols))
    if spd:  # symmetrise the pattern
        sets = [set(r) for r in rows]
        for i, r in enumerate(rows):
            for c in r:
                sets[c].add(i)
        rows = [sorted(s) for s in sets]
    index = np.zeros(n + 1, np.int32)
    index[1:] = np.cumsum([len(r) for r in rows])
    item = np.array([c + 1 for r in rows for c in r], np.int32)
    b2 = ndof * ndof
    A = rng.uniform(-1, 1, size=len(item) * b2)
    m = mm.Matrix(n, n, ndof, index, item, A).ensure_dof_index()
    if spd:
        D = m.dense()
        D = 0.5 * (D + D.T)
        D += np.eye(len(D)) * (np.abs(D).sum(1).max() + 1.0)
        for i in range(n):
            for k in range(index[i], index[i + 1]):
                c = item[k] - 1
                m.A[k * b2:(k + 1) * b2] = D[i * ndof:(i + 1) * ndof, c * ndof:(c + 1) * ndof].reshape(-1)
    return m


def random_vblock(n, density, seed, maxb=4, spd=True, sizes=None):
    """Random variable-block matrix (monolis n_dof = -1): block sizes 1..maxb,
    symmetric pattern with the diagonal, values uniform; SPD (symmetric,
    diagonally dominant) when spd."""
    rng = np.random.default_rng(seed)
    lst = np.asarray(sizes if sizes is not None else rng.integers(1, maxb + 1, size=n), np.int32)
    elems = []
    for i in range(n):
        for j in np.nonzero(rng.random(n) < density)[0]:
            if j != i:
                elems.append((i, int(j)))
    m = mm.nonzero_pattern_by_simple_mesh_V(n, lst, elems)
    m.A = rng.uniform(-1, 1, len(m.A))
    if spd:
        D = m.dense()
        D = 0.5 * (D + D.T)
        D += np.eye(len(D)) * (np.abs(D).sum(1).max() + 1.0)
        for i in range(n):
            r0, n1 = int(m.n_dof_index[i]), int(lst[i])
            for e in range(m.index[i], m.index[i + 1]):
                j = int(m.item[e]) - 1
                c0, n2 = int(m.n_dof_index[j]), int(lst[j])
                a0 = int(m.n_dof_index2[e])
                m.A[a0:a0 + n1 * n2] = D[r0:r0 + n1, c0:c0 + n2].reshape(-1)
    return m


def chain_matrix_C(n_node, ndof, seed):
    """Shape of src_test/iterative/COCG_test.f90:41-70: 10-node chain, diagonal
    (U[0,1) + i U[0,1)) + (2 + 2i) on the block diagonals, complex symmetric
    (A_ij = A_ji, not conjugated) U[0,1) + i U[0,1) off-diagonal blocks."""
    rng = np.random.default_rng(seed)
    elems = [[i, i + 1] for i in range(n_node - 1)]
    m = mm.nonzero_pattern_by_simple_mesh(n_node, ndof, elems, dtype=np.complex128)
    for i in range(n_node):
        for d in range(ndof):
            mm.add_scalar(m, i, i, d, d, complex(rng.random(), rng.random()) + (2.0 + 2.0j))
    for a, b in elems:
        for i2 in range(ndof):
            for j2 in range(ndof):
                v = complex(rng.random(), rng.random())
                mm.add_scalar(m, a, b, i2, j2, v)
                mm.add_scalar(m, b, a, j2, i2, v)
    return m


def chain_matrix_C_wrapper(seed, n_node=10, ndof=2):
    """Shape of wrapper_test/solver/monolis_solver_c_test.c:83-141: diagonal
    v + 0.1 v i with v in 5000..6000, off-diagonal blocks v + 0.1 v i with
    v in 0..1000 set at (e0, e1, j, k) and (e1, e0, j, k) (the same block on
    both sides, so not symmetric for n_dof > 1)."""
    rng = np.random.default_rng(seed)
    elems = [[i, i + 1] for i in range(n_node - 1)]
    m = mm.nonzero_pattern_by_simple_mesh(n_node, ndof, elems, dtype=np.complex128)
    for i in range(n_node):
        for j in range(ndof):
            v = float(rng.integers(0, 1001)) + 5000.0
            mm.set_scalar(m, i, i, j, j, v + 0.1 * v * 1j)
    for a, b in elems:
        for j in range(ndof):
            for k in range(ndof):
                v = float(rng.integers(0, 1001))
                mm.set_scalar(m, a, b, j, k, v + 0.1 * v * 1j)
                mm.set_scalar(m, b, a, j, k, v + 0.1 * v * 1j)
    return m


def random_bcsr_C(n, ndof, density, seed, symmetric=True, diag_missing=()):
    """Random complex block matrix; with `symmetric` it is complex symmetric
    (A = A^T, no conjugation), diagonally dominant with a positive imaginary
    shift on the diagonal -- the damped-Helmholtz class COCG is meant for."""
    mr = random_bcsr(n, ndof, density, seed, spd=False, diag_missing=diag_missing)
    rng = np.random.default_rng(seed + 1000)
    b2 = ndof * ndof
    A = mr.A + 1j * rng.uniform(-1, 1, size=mr.A.size)
    m = mm.Matrix(n, n, ndof, mr.index.copy(), mr.item.copy(), A).ensure_dof_index()
    if symmetric:
        sets = [set((m.item[m.index[i]:m.index[i + 1]] - 1).tolist()) for i in range(n)]
        for i in range(n):
            for c in list(sets[i]):
                sets[c].add(i)
        rows = [sorted(s) for s in sets]
        index = np.zeros(n + 1, np.int32)
        index[1:] = np.cumsum([len(r) for r in rows])
        item = np.array([c + 1 for r in rows for c in r], np.int32)
        D = np.zeros((n * ndof, n * ndof), np.complex128)
        for i in range(n):
            for k in range(m.index[i], m.index[i + 1]):
                c = m.item[k] - 1
                D[i * ndof:(i + 1) * ndof, c * ndof:(c + 1) * ndof] = m.A[k * b2:(k + 1) * b2].reshape(ndof, ndof)
        D = 0.5 * (D + D.T)
        D += np.eye(len(D)) * (np.abs(D).sum(1).max() + 1.0) * (1.0 + 0.5j)
        A = np.zeros(len(item) * b2, np.complex128)
        for i in range(n):
            for k in range(index[i], index[i + 1]):
                c = item[k] - 1
                A[k * b2:(k + 1) * b2] = D[i * ndof:(i + 1) * ndof, c * ndof:(c + 1) * ndof].reshape(-1)
        m = mm.Matrix(n, n, ndof, index, item, A).ensure_dof_index()
    return m


def dense_C(m):
    """Dense complex copy of a fixed-block matrix (tiny sizes only)."""
    b = m.ndof
    D = np.zeros((m.n * b, m.np_ * b), np.complex128)
    for i in range(m.n):
        for k in range(m.index[i], m.index[i + 1]):
            c = m.item[k] - 1
            D[i * b:(i + 1) * b, c * b:(c + 1) * b] = m.A[k * b * b:(k + 1) * b * b].reshape(b, b)
    return D


def random_vblock_C(n, density, seed, maxb=4, sizes=None):
    """Complex symmetric variable-block matrix (A = A^T, no conjugation; monolis
    n_dof = -1): the pattern of random_vblock, complex values, diagonally
    dominant with a positive imaginary shift on the diagonal."""
    m = random_vblock(n, density, seed, maxb=maxb, spd=False, sizes=sizes)
    rng = np.random.default_rng(seed + 2000)
    m.A = m.A + 1j * rng.uniform(-1, 1, len(m.A))
    D = m.dense()
    D = 0.5 * (D + D.T)
    D += np.eye(len(D)) * (np.abs(D).sum(1).max() + 1.0) * (1.0 + 0.5j)
    _store_vblock_dense(m, D)
    return m


def _store_vblock_dense(m, D):
    for i in range(m.n):
        r0, n1 = int(m.n_dof_index[i]), int(m.n_dof_list[i])
        for e in range(m.index[i], m.index[i + 1]):
            j = int(m.item[e]) - 1
            c0, n2 = int(m.n_dof_index[j]), int(m.n_dof_list[j])
            a0 = int(m.n_dof_index2[e])
            m.A[a0:a0 + n1 * n2] = D[r0:r0 + n1, c0:c0 + n2].reshape(-1)


def as_vblock(m):
    """The fixed-block matrix m restated with n_dof = -1 (every block m.ndof):
    same index / item / values, n_dof_list = ndof, n_dof_index = i*ndof,
    n_dof_index2 = e*ndof^2."""
    b = m.ndof
    v = mm.Matrix(m.n, m.np_, -1, m.index.copy(), m.item.copy(), m.A.copy())
    v.n_dof_list = np.full(m.np_, b, np.int32)
    v.n_dof_index = (np.arange(m.np_ + 1, dtype=np.int64) * b).astype(np.int32)
    v.n_dof_index2 = (np.arange(len(m.item) + 1, dtype=np.int64) * b * b).astype(np.int32)
    return v


def hex_mesh(nx, ny=None, nz=None):
    """Structured hexahedral mesh: (nx+1)(ny+1)(nz+1) nodes, nx*ny*nz 8-node
    elements, 0-based node ids, element-major (n_elem, 8)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    ex, ey, ez = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ex, ey, ez = ex.ravel(), ey.ravel(), ez.ravel()
    nid = lambda x, y, z: (z * (ny + 1) + y) * (nx + 1) + x
    corners = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
    elems = np.stack([nid(ex + dx, ey + dy, ez + dz) for dx, dy, dz in corners], axis=1).astype(np.int32)
    return (nx + 1) * (ny + 1) * (nz + 1), elems


def element_matrices(n_elem, nzm, seed, spd=True):
    """Random element matrices (n_elem, nzm, nzm); symmetric positive
    semi-definite plus a small diagonal shift when spd."""
    rng = np.random.default_rng(seed)
    R = rng.uniform(-1, 1, size=(n_elem, nzm, nzm))
    if not spd:
        return R
    return np.matmul(R, R.transpose(0, 2, 1)) / nzm + 0.01 * np.eye(nzm)
