"""The oracle's PipeCR (monolis method 4, src/iterative/PipeCR.f90), pinned on
the CPU against:
  * the reference's test (src_test/iterative/PipeCR_test.f90:21-90: the
    10-node chain with n_dof = 1..3, b = A 1, tol 1e-10, solution 1 for NONE
    and DIAG);
  * brute force: preconditioned conjugate residuals minimise r^T M^-1 r over
    x0 + K_k(M^-1 A, M^-1 r0) (M = I: the 2-norm of r, as MINRES).  The
    minimum over an orthonormal Krylov basis by least squares must equal the
    oracle's residual after k updates; a wrong scalar recurrence breaks that;
  * the stopping rule: the test runs before the updates of an iteration on
    ||M^-1 r|| / ||M^-1 r0|| <= tol (PipeCR.f90:78-97), so cur_iter is one
    more than the first k whose brute-force minimiser passes it;
  * dense solves, with the 50-iteration residual replacement.
"""
import numpy as np
import pytest

import oracle
from monolis_amd import workloads as W

from _helpers import chain_matrix


@pytest.fixture(scope="module", autouse=True)
def _threads():
    oracle.set_threads(1)
    yield


def _block_diag_inv(m):
    """M^-1 of monolis's DIAG preconditioner (the inverse of the diagonal blocks)."""
    Ad = m.dense()
    b = m.ndof
    Minv = np.zeros_like(Ad)
    for i in range(m.n):
        s = slice(i * b, (i + 1) * b)
        Minv[s, s] = np.linalg.inv(Ad[s, s])
    return Minv


def _krylov_minimisers(A, Minv, b, kmax):
    """x_k minimising r^T Minv r over K_k(Minv A, Minv b) (x0 = 0), k = 1..kmax."""
    C = np.linalg.cholesky(Minv)             # Minv = C C^T, r^T Minv r = ||C^T r||^2
    n = len(b)
    V = np.zeros((n, kmax))
    w = Minv @ b
    V[:, 0] = w / np.linalg.norm(w)
    out = []
    for k in range(1, kmax + 1):
        Vk = V[:, :k]
        y, *_ = np.linalg.lstsq(C.T @ (A @ Vk), C.T @ b, rcond=None)
        out.append(Vk @ y)
        if k < kmax:                          # Arnoldi step, modified Gram-Schmidt twice
            w = Minv @ (A @ V[:, k - 1])
            for _ in range(2):
                for j in range(k):
                    w -= (V[:, j] @ w) * V[:, j]
            V[:, k] = w / np.linalg.norm(w)
    return out, C


@pytest.mark.parametrize("ndof", [1, 2, 3])
@pytest.mark.parametrize("precond", [oracle.PREC_NONE, oracle.PREC_DIAG])
def test_pipecr_chain_x_is_one(ndof, precond):
    m = chain_matrix(10, ndof, seed=ndof)
    b = oracle.matvec(m, np.ones(m.npndof))
    x, st, I, R = oracle.solve(m, b, method=oracle.ITER_PIPECR, precond=precond, tol=1e-10)
    assert st == 0 and I[4] >= 1 and R[2] <= 1e-10
    assert np.allclose(x, 1.0, rtol=0, atol=1e-7)


@pytest.mark.parametrize("case", ["poisson5_none", "poisson5_diag", "bcsr3_3_diag"])
def test_pipecr_minimises_preconditioned_residual(case):
    if case.startswith("poisson5"):
        m = W.poisson7(5)
    else:
        m = W.stencil27_bcsr3(3, seed=4)
    precond = oracle.PREC_DIAG if case.endswith("diag") else oracle.PREC_NONE
    A = m.dense()
    Minv = _block_diag_inv(m) if precond == oracle.PREC_DIAG else np.eye(m.nndof)
    b = np.random.default_rng(7).uniform(-1, 1, m.nndof)
    kmax = 8
    xs, C = _krylov_minimisers(A, Minv, b, kmax)
    for k in range(1, kmax + 1):
        Io, Ro = oracle.default_arrays(method=oracle.ITER_PIPECR, precond=precond, tol=1e-300, max_iter=k)
        Io[16] = 0
        x, st, I, R = oracle.solve(m, b, Iarray=Io, Rarray=Ro)
        assert st == 1 and I[4] == k
        got = np.linalg.norm(C.T @ (b - A @ x))
        best = np.linalg.norm(C.T @ (b - A @ xs[k - 1]))
        assert got <= best * (1 + 1e-7) + 1e-13, (k, got, best)
        assert np.linalg.norm(x - xs[k - 1]) <= 1e-6 * np.linalg.norm(xs[k - 1])


@pytest.mark.parametrize("precond", [oracle.PREC_NONE, oracle.PREC_DIAG])
def test_pipecr_stopping_rule_and_reported_residual(precond):
    m = W.stencil27_bcsr3(3, seed=5)
    A = m.dense()
    Minv = _block_diag_inv(m) if precond == oracle.PREC_DIAG else np.eye(m.nndof)
    b = np.random.default_rng(8).uniform(-1, 1, m.nndof)
    tol = 1e-5
    xs, _ = _krylov_minimisers(A, Minv, b, 40)
    u0 = np.linalg.norm(Minv @ b)
    ratios = [np.linalg.norm(Minv @ (b - A @ xk)) / u0 for xk in xs]
    first = next(k for k, q in enumerate(ratios, start=1) if q <= tol)
    x, st, I, R = oracle.solve(m, b, method=oracle.ITER_PIPECR, precond=precond, tol=tol)
    assert st == 0
    assert abs(int(I[4]) - (first + 1)) <= 1, (I[4], first)
    true_ratio = np.linalg.norm(Minv @ (b - A @ x)) / u0
    assert R[2] <= tol
    assert abs(R[2] - true_ratio) <= 1e-3 * true_ratio


@pytest.mark.parametrize("case", ["poisson40_none", "bcsr3_6_diag"])
def test_pipecr_matches_dense_with_residual_replacement(case):
    if case == "poisson40_none":
        m, precond = W.poisson7(40), oracle.PREC_NONE         # > 50 iterations: residual replacement
    else:
        m, precond = W.stencil27_bcsr3(6, seed=2), oracle.PREC_DIAG
    b = np.random.default_rng(3).uniform(-1, 1, m.nndof)
    x, st, I, R = oracle.solve(m, b, method=oracle.ITER_PIPECR, precond=precond, tol=1e-10)
    xc, stc, Ic, Rc = oracle.solve(m, b, method=oracle.ITER_CG, precond=precond, tol=1e-12)
    assert st == 0 and stc == 0
    if case == "poisson40_none":
        assert I[4] > 50
    assert np.linalg.norm(x - xc) / np.linalg.norm(xc) < 1e-8
    if m.nndof <= 5000:
        xd = np.linalg.solve(m.dense(), b)
        assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-8


def test_pipecr_zero_rhs_and_max_iter():
    m = W.poisson7(8)
    x, st, I, R = oracle.solve(m, np.zeros(m.nndof), method=oracle.ITER_PIPECR)
    assert st == 0 and np.all(x == 0.0)
    Io, Ro = oracle.default_arrays(method=oracle.ITER_PIPECR, max_iter=3)
    Io[16] = 0
    x, st, I, R = oracle.solve(m, np.ones(m.nndof), Iarray=Io, Rarray=Ro)
    assert st == 1 and I[4] == 3 and np.all(np.isfinite(x))
