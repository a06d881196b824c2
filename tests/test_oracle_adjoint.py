"""The oracle's FP64 adjoint (transpose, lambda = A^{-T} g, dL/dA = -lambda x^T on
the pattern), pinned on the CPU against:
  * scipy's sparse transpose (structure and values, block transposes);
  * dense numpy solves of A^T lambda = g;
  * central finite differences (h = 1e-6) of L = ||A(theta)^{-1} b - x_t||^2
    with dense solves, for SPD block and nonsymmetric scalar systems;
  * a planted mistake (A instead of A^T for a nonsymmetric matrix) that the
    finite-difference check must reject.
"""
import numpy as np
import pytest

import oracle
from monolis_amd import workloads as W

from _helpers import random_bcsr


def _dense_solve(mat, b):
    return np.linalg.solve(mat.dense(), b)


def _loss(mat, b, xt):
    x = _dense_solve(mat, b)
    return float(np.sum((x - xt) ** 2))


@pytest.mark.parametrize("ndof", [1, 2, 3])
def test_transpose_matches_scipy(ndof):
    mat = random_bcsr(40, ndof, 0.1, seed=ndof)
    T = oracle.transpose(mat)
    assert (T.to_scipy() != mat.to_scipy().T).nnz == 0
    # block columns ascending per row, diagonal kept
    for i in range(T.n):
        cols = T.item[T.index[i]:T.index[i + 1]]
        assert np.all(np.diff(cols) > 0)
    TT = oracle.transpose(T)
    assert np.array_equal(TT.index, mat.index) and np.array_equal(TT.item, mat.item)
    assert np.array_equal(TT.A, mat.A)


@pytest.mark.parametrize("ndof", [1, 3])
def test_adjoint_lambda_matches_dense(ndof):
    mat = random_bcsr(30, ndof, 0.15, seed=11, spd=False)
    D = mat.dense()
    D += np.eye(len(D)) * (np.abs(D).sum(1).max() + 1.0)     # diagonally dominant: BiCGSTAB converges
    mat = _with_dense(mat, D)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, mat.n * ndof)
    g = rng.uniform(-1, 1, mat.n * ndof)
    lam, dA, st, I = oracle.adjoint(mat, x, g, method=oracle.ITER_BICGSTAB, tol=1e-12)
    assert st == 0
    lam_d = np.linalg.solve(mat.dense().T, g)
    assert np.linalg.norm(lam - lam_d) / np.linalg.norm(lam_d) < 1e-10
    # dA equals the dense -lambda x^T at the stored entries and nowhere else
    full = -np.outer(lam, x)
    G = _with_values(mat, dA).dense()
    mask = _with_values(mat, np.ones_like(dA)).dense() != 0
    assert np.allclose(G[mask], full[mask], rtol=0, atol=1e-14)


def _with_values(mat, A):
    return type(mat)(mat.n, mat.np_, mat.ndof, mat.index.copy(), mat.item.copy(),
                     np.asarray(A, np.float64).copy()).ensure_dof_index()


def _with_dense(mat, D):
    """Same pattern, values taken from the dense matrix D."""
    b = mat.ndof
    rows = np.repeat(np.arange(mat.n), np.diff(mat.index))
    cols = mat.item - 1
    A = np.stack([D[i * b:(i + 1) * b, j * b:(j + 1) * b] for i, j in zip(rows, cols)]).reshape(-1)
    return _with_values(mat, A)


def _fd_check(mat, b, xt, symmetric, method, n_check=40, seed=0, planted=False):
    x = _dense_solve(mat, b)
    g = 2.0 * (x - xt)
    if planted:   # the mistake: solve with A instead of A^T
        lam = oracle.solve(mat, g, method=method, tol=1e-13)[0]
        dA = oracle.outer_grad(mat, lam, x)
    else:
        lam, dA, st, I = oracle.adjoint(mat, x, g, symmetric=symmetric, method=method, tol=1e-13)
        assert st == 0
    rng = np.random.default_rng(seed)
    h = 1e-6
    worst = 0.0
    for p in rng.choice(len(mat.A), size=n_check, replace=False):
        Ap, Am = mat.A.copy(), mat.A.copy()
        Ap[p] += h
        Am[p] -= h
        fd = (_loss(_with_values(mat, Ap), b, xt) - _loss(_with_values(mat, Am), b, xt)) / (2 * h)
        worst = max(worst, abs(fd - dA[p]) / (1e-5 * abs(fd) + 1e-7))
    for r in rng.choice(len(b), size=10, replace=False):
        bp, bm = b.copy(), b.copy()
        bp[r] += h
        bm[r] -= h
        fd = (_loss(mat, bp, xt) - _loss(mat, bm, xt)) / (2 * h)
        worst = max(worst, abs(fd - lam[r]) / (1e-5 * abs(fd) + 1e-7))
    return worst   # <= 1: every entry within 1e-5 relative + 1e-7 absolute (FD roundoff of L)


def test_adjoint_fd_spd_bcsr3():
    mat = W.stencil27_bcsr3(4, seed=5)
    rng = np.random.default_rng(6)
    b = rng.uniform(-1, 1, mat.n * 3)
    xt = rng.uniform(-1, 1, mat.n * 3)
    assert _fd_check(mat, b, xt, symmetric=True, method=oracle.ITER_CG) <= 1.0
    # the same gradient through the explicit transpose (A^T = A here)
    assert _fd_check(mat, b, xt, symmetric=False, method=oracle.ITER_CG) <= 1.0


def test_adjoint_fd_nonsymmetric_convdiff():
    mat = W.convdiff7(5, peclet=1.5)
    rng = np.random.default_rng(7)
    b = rng.uniform(-1, 1, mat.n)
    xt = rng.uniform(-1, 1, mat.n)
    assert _fd_check(mat, b, xt, symmetric=False, method=oracle.ITER_BICGSTAB) <= 1.0
    # planted mistake (A in place of A^T) must be caught
    assert _fd_check(mat, b, xt, symmetric=False, method=oracle.ITER_BICGSTAB, planted=True) > 100.0
