"""The oracle's PipeCG (monolis method 3, src/iterative/PipeCG.f90) and GropCG
(method 2, src/iterative/GropCG.f90), pinned on the CPU against:
  * the reference's tests of both (src_test/iterative/PipeCG_test.f90,
    GropCG_test.f90: the 10-node chain with n_dof = 1..3, b = A 1, solution 1
    for NONE and DIAG);
  * dense solves;
  * textbook PCG: pipelined CG is the same Krylov method, so its iteration
    counts match the oracle's CG within a couple of iterations on
    well-conditioned systems, with and without the 50-iteration residual
    replacement; a wrong scalar recurrence breaks that.
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


@pytest.mark.parametrize("method", [oracle.ITER_PIPECG, oracle.ITER_GROPCG])
@pytest.mark.parametrize("ndof", [1, 2, 3])
@pytest.mark.parametrize("precond", [oracle.PREC_NONE, oracle.PREC_DIAG])
def test_pipecg_chain_x_is_one(ndof, precond, method):
    m = chain_matrix(10, ndof, seed=ndof)
    b = oracle.matvec(m, np.ones(m.npndof))
    x, st, I, R = oracle.solve(m, b, method=method, precond=precond)
    assert st == 0
    assert np.allclose(x, 1.0, rtol=0, atol=1e-7)


@pytest.mark.parametrize("method", [oracle.ITER_PIPECG, oracle.ITER_GROPCG])
@pytest.mark.parametrize("case", ["poisson16", "bcsr3_6", "poisson40_none"])
def test_pipecg_matches_pcg_and_dense(case, method):
    if case == "poisson16":
        m, precond = W.poisson7(16), oracle.PREC_DIAG
    elif case == "bcsr3_6":
        m, precond = W.stencil27_bcsr3(6, seed=2), oracle.PREC_DIAG
    else:
        m, precond = W.poisson7(40), oracle.PREC_NONE      # > 50 iterations: residual replacement
    b = np.random.default_rng(3).uniform(-1, 1, m.nndof)
    xp, stp, Ip, Rp = oracle.solve(m, b, method=method, precond=precond, tol=1e-10)
    xc, stc, Ic, Rc = oracle.solve(m, b, method=oracle.ITER_CG, precond=precond, tol=1e-10)
    assert stp == 0 and stc == 0
    assert abs(int(Ip[4]) - int(Ic[4])) <= 2, (Ip[4], Ic[4])
    if case == "poisson40_none":
        assert Ip[4] > 50
    assert np.linalg.norm(xp - xc) / np.linalg.norm(xc) < 1e-8
    if m.nndof <= 5000:
        xd = np.linalg.solve(m.dense(), b)
        assert np.linalg.norm(xp - xd) / np.linalg.norm(xd) < 1e-8


@pytest.mark.parametrize("method", [oracle.ITER_PIPECG, oracle.ITER_GROPCG])
def test_pipecg_zero_rhs_and_max_iter(method):
    m = W.poisson7(8)
    x, st, I, R = oracle.solve(m, np.zeros(m.nndof), method=method)
    assert st == 0 and np.all(x == 0.0)
    Io, Ro = oracle.default_arrays(method=method, max_iter=3)
    Io[16] = 0
    x, st, I, R = oracle.solve(m, np.ones(m.nndof), Iarray=Io, Rarray=Ro)
    assert st == 1 and I[4] == 3


@pytest.mark.parametrize("ndof", [1, 2, 3])
def test_bicgstab_noprec_chain_x_is_one(ndof):
    # src_test/iterative/BiCGSTAB_noprec_test.f90: the chain, b = A 1
    m = chain_matrix(10, ndof, seed=ndof)
    b = oracle.matvec(m, np.ones(m.npndof))
    x, st, I, R = oracle.solve(m, b, method=oracle.ITER_BICGSTAB_NOPREC, precond=oracle.PREC_DIAG, tol=1e-10)
    assert st == 0
    assert np.allclose(x, 1.0, rtol=0, atol=1e-8)


def test_bicgstab_noprec_is_unpreconditioned_bicgstab():
    # below 50 iterations the two methods do the same arithmetic
    m = W.convdiff7(8, peclet=1.0)
    b = np.random.default_rng(5).uniform(-1, 1, m.nndof)
    x7, st7, I7, R7 = oracle.solve(m, b, method=oracle.ITER_BICGSTAB_NOPREC, precond=oracle.PREC_DIAG)
    x5, st5, I5, R5 = oracle.solve(m, b, method=oracle.ITER_BICGSTAB, precond=oracle.PREC_NONE)
    assert st7 == 0 and st5 == 0 and I7[4] < 50
    assert I7[4] == I5[4] and np.array_equal(x7, x5)
    # beyond 50 iterations the true residual comes every 50 iterations; the answer is the same
    m = W.convdiff7(24, peclet=0.5)
    b = np.random.default_rng(6).uniform(-1, 1, m.nndof)
    x7, st7, I7, R7 = oracle.solve(m, b, method=oracle.ITER_BICGSTAB_NOPREC, tol=1e-10)
    x5, st5, I5, R5 = oracle.solve(m, b, method=oracle.ITER_BICGSTAB, precond=oracle.PREC_NONE, tol=1e-10)
    assert st7 == 0 and I7[4] > 50 and abs(int(I7[4]) - int(I5[4])) <= 3
    assert np.linalg.norm(x7 - x5) / np.linalg.norm(x5) < 1e-8


@pytest.mark.parametrize("method", [oracle.ITER_PIPEBICGSTAB, oracle.ITER_PIPEBICGSTAB_NOPREC])
@pytest.mark.parametrize("ndof", [1, 2, 3])
def test_pipebicgstab_chain_x_is_one(method, ndof):
    # src_test/iterative/PipeBiCGSTAB_test.f90, PipeBiCGSTAB_noprec_test.f90: the chain, b = A 1
    m = chain_matrix(10, ndof, seed=ndof)
    b = oracle.matvec(m, np.ones(m.npndof))
    x, st, I, R = oracle.solve(m, b, method=method, precond=oracle.PREC_DIAG, tol=1e-10)
    assert st == 0
    assert np.allclose(x, 1.0, rtol=0, atol=1e-8)


@pytest.mark.parametrize("case", ["convdiff10_diag", "convdiff16_none", "bcsr3_5_diag"])
def test_pipebicgstab_matches_bicgstab_and_dense(case):
    """Pipelined BiCGSTAB is BiCGSTAB with the recurrences rearranged: the same
    iteration count within a few, the same solution; the noprec variant equals
    the preconditioned one with M = I."""
    if case == "convdiff10_diag":
        m, precond = W.convdiff7(10, peclet=1.0), oracle.PREC_DIAG
    elif case == "convdiff16_none":
        m, precond = W.convdiff7(16, peclet=0.5), oracle.PREC_NONE
    else:
        m, precond = W.stencil27_bcsr3(5, seed=4), oracle.PREC_DIAG
    b = np.random.default_rng(8).uniform(-1, 1, m.nndof)
    xp, stp, Ip, Rp = oracle.solve(m, b, method=oracle.ITER_PIPEBICGSTAB, precond=precond, tol=1e-10)
    xb, stb, Ib, Rb = oracle.solve(m, b, method=oracle.ITER_BICGSTAB, precond=precond, tol=1e-10)
    assert stp == 0 and stb == 0
    assert abs(int(Ip[4]) - int(Ib[4])) <= max(3, int(0.1 * Ib[4])), (Ip[4], Ib[4])
    assert np.linalg.norm(xp - xb) / np.linalg.norm(xb) < 1e-8
    if m.nndof <= 5000:
        xd = np.linalg.solve(m.dense(), b)
        assert np.linalg.norm(xp - xd) / np.linalg.norm(xd) < 1e-8
    if precond == oracle.PREC_NONE:
        xn, stn, In, Rn = oracle.solve(m, b, method=oracle.ITER_PIPEBICGSTAB_NOPREC, precond=oracle.PREC_DIAG,
                                       tol=1e-10)
        assert stn == 0 and In[4] == Ip[4] and np.array_equal(xn, xp)
