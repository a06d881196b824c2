"""Pins the oracle's complex path (monolis_oracle_cplx.c) before it is used to
check the GPU: brute force, closed forms, and the reduction of COCG to the
already pinned real CG on real-valued data."""
import numpy as np
import pytest

import oracle
from _helpers import (as_vblock, chain_matrix, chain_matrix_C, chain_matrix_C_wrapper, dense_C, random_bcsr,
                      random_bcsr_C, random_vblock, random_vblock_C)
from monolis_amd import matrix as mm


@pytest.mark.parametrize("ndof", [1, 2, 3, 4])
def test_matvec_C_matches_dense_brute_force(ndof):
    m = random_bcsr_C(37, ndof, 0.15, seed=ndof, symmetric=False)
    rng = np.random.default_rng(3)
    x = rng.standard_normal(m.n * ndof) + 1j * rng.standard_normal(m.n * ndof)
    y = oracle.matvec_C(m, x)
    np.testing.assert_allclose(y, dense_C(m) @ x, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("ndof", [1, 3])
def test_matvec_C_real_data_equals_real_oracle_and_is_complex_linear(ndof):
    mr = random_bcsr(29, ndof, 0.2, seed=7)
    mc = mm.Matrix(mr.n, mr.n, ndof, mr.index, mr.item, mr.A.astype(np.complex128)).ensure_dof_index()
    x = np.random.default_rng(1).standard_normal(mr.n * ndof)
    yc = oracle.matvec_C(mc, x.astype(np.complex128))
    assert np.array_equal(yc.real, oracle.matvec(mr, x)) and not yc.imag.any()
    # i * (A x) == A (i x): a swapped real/imaginary part or a dropped cross term fails this
    mz = random_bcsr_C(29, ndof, 0.2, seed=8, symmetric=False)
    z = x + 1j * np.random.default_rng(2).standard_normal(x.size)
    np.testing.assert_allclose(oracle.matvec_C(mz, 1j * z), 1j * oracle.matvec_C(mz, z), rtol=1e-14, atol=1e-14)


def test_inner_product_C_is_unconjugated():
    m = 10
    # (i, i) = -1 per entry; a conjugating (Hermitian) product would give +1
    assert oracle.inner_product_C(np.full(m, 1j), np.full(m, 1j)) == -10
    # (1 + i)^2 = 2i
    assert oracle.inner_product_C(np.full(m, 1 + 1j), np.full(m, 1 + 1j)) == 20j
    rng = np.random.default_rng(0)
    x = rng.standard_normal(100) + 1j * rng.standard_normal(100)
    y = rng.standard_normal(100) + 1j * rng.standard_normal(100)
    assert abs(oracle.inner_product_C(x, y) - np.sum(x * y)) < 1e-12


def test_check_converge_C_closed_forms():
    # R2 = 1, B2 = 4: resid = sqrt(1/4) = 0.5
    conv, it, res = oracle.check_converge_C(np.array([1.0 + 0j, 0.0]), 4.0, 3, 0.6)
    assert conv and it == 3 and res == 0.5
    # R2 = (i)^2 = -1, B2 = 1: principal sqrt(-1) = i, so the stored (real) residual
    # is 0 while |resid| = 1 decides convergence
    conv, it, res = oracle.check_converge_C(np.array([1j]), 1.0, 5, 0.5)
    assert not conv and it == 5 and res == 0.0
    # R2 = 2i, B2 = 2: sqrt(i) = (1 + i)/sqrt(2): real part 1/sqrt(2), modulus 1
    conv, _, res = oracle.check_converge_C(np.array([1 + 1j]), 2.0, 1, 1.0 + 1e-12)
    assert conv and abs(res - 2 ** -0.5) < 1e-15


@pytest.mark.parametrize("ndof", [1, 2, 3, 4])
def test_diag_C_inverts_each_diagonal_block(ndof):
    m = random_bcsr_C(23, ndof, 0.2, seed=11)
    D, st = oracle.diag_setup_C(m)
    assert st == 0
    full = dense_C(m)
    rng = np.random.default_rng(5)
    v = rng.standard_normal(m.n * ndof) + 1j * rng.standard_normal(m.n * ndof)
    y = oracle.diag_apply_C(m, D, v)
    for i in range(m.n):
        blk = full[i * ndof:(i + 1) * ndof, i * ndof:(i + 1) * ndof]
        np.testing.assert_allclose(y[i * ndof:(i + 1) * ndof], np.linalg.solve(blk, v[i * ndof:(i + 1) * ndof]),
                                   rtol=1e-12, atol=1e-12)


def test_diag_C_zero_pivot_and_missing_diagonal():
    m = chain_matrix_C(4, 2, seed=0)
    mm.set_scalar(m, 2, 2, 0, 0, 0.0)
    _, st = oracle.diag_setup_C(m)
    assert st == oracle.ZERO_DIAG
    # a row without a diagonal block keeps zero factors (diag_nn.f90:110-116 only factors found blocks)
    m2 = random_bcsr_C(8, 2, 0.3, seed=3, symmetric=False, diag_missing=(5,))
    D, st = oracle.diag_setup_C(m2)
    assert st == 0 and not D[5 * 4:6 * 4].any()


@pytest.mark.parametrize("ndof", [1, 2, 3])
@pytest.mark.parametrize("precond", [oracle.PREC_NONE, oracle.PREC_DIAG])
def test_cocg_chain_x_is_one_plus_i(ndof, precond):
    """src_test/iterative/COCG_test.f90:71-85: b = A (1 + i), solve at 1e-12, x = 1 + i."""
    m = chain_matrix_C(10, ndof, seed=ndof)
    b = oracle.matvec_C(m, np.full(m.n * ndof, 1 + 1j))
    x, st, I, R = oracle.solve_C(m, b, precond=precond, tol=1e-12)
    assert st == 0 and 0 < I[oracle.I_CUR_ITER] <= 10 * ndof + 5
    np.testing.assert_allclose(x, 1 + 1j, rtol=1e-9, atol=1e-9)


def test_cocg_wrapper_test_shape():
    """wrapper_test/solver/monolis_solver_c_test.c:83-158 (n_dof 2, tol 1e-10)."""
    m = chain_matrix_C_wrapper(seed=4)
    b = oracle.matvec_C(m, np.full(20, 1 + 1j))
    x, st, I, R = oracle.solve_C(m, b, tol=1e-10)
    assert st == 0
    np.testing.assert_allclose(x, 1 + 1j, rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize("ndof", [1, 2, 4])
def test_cocg_matches_dense_solve(ndof):
    m = random_bcsr_C(60, ndof, 0.08, seed=20 + ndof)
    rng = np.random.default_rng(ndof)
    b = rng.standard_normal(m.n * ndof) + 1j * rng.standard_normal(m.n * ndof)
    x, st, I, R = oracle.solve_C(m, b, tol=1e-12)
    assert st == 0
    np.testing.assert_allclose(x, np.linalg.solve(dense_C(m), b), rtol=1e-9, atol=1e-9)
    # Rarray[2] holds the real part of the complex relative residual, below tol in modulus
    assert abs(R[oracle.R_CUR_RESID]) < 1e-12


@pytest.mark.parametrize("ndof", [1, 2])
@pytest.mark.parametrize("precond", [oracle.PREC_NONE, oracle.PREC_DIAG])
def test_cocg_on_real_data_is_the_real_cg(ndof, precond):
    """With real A and b every imaginary part stays exactly zero and COCG is the
    CG recurrence of CG.f90 (B = 1, 2 use the same diag_nn factors in both
    paths): same iterations, same iterate, bit for bit on one thread."""
    oracle.set_threads(1)
    mr = chain_matrix(10, ndof, seed=ndof)
    mc = mm.Matrix(mr.n, mr.n, ndof, mr.index, mr.item, mr.A.astype(np.complex128)).ensure_dof_index()
    b = oracle.matvec(mr, np.ones(mr.n * ndof))
    xr, str_, Ir, Rr = oracle.solve(mr, b, method=oracle.ITER_CG, precond=precond, tol=1e-12)
    xc, stc, Ic, Rc = oracle.solve_C(mc, b.astype(np.complex128), precond=precond, tol=1e-12)
    assert str_ == stc == 0
    assert Ir[oracle.I_CUR_ITER] == Ic[oracle.I_CUR_ITER]
    assert np.array_equal(xc.real, xr) and not xc.imag.any()
    assert Rc[oracle.R_CUR_RESID] == Rr[oracle.R_CUR_RESID]


def test_cocg_zero_rhs_and_bad_arguments():
    m = chain_matrix_C(10, 2, seed=1)
    x, st, I, R = oracle.solve_C(m, np.zeros(20, np.complex128), x0=np.full(20, 3 + 1j), precond=0)
    assert st == 0 and not x.any()
    _, st, _, _ = oracle.solve_C(m, np.ones(20, np.complex128), method=oracle.ITER_CG)
    assert st == -2
    _, st, _, _ = oracle.solve_C(m, np.ones(20, np.complex128), precond=2)
    assert st == -3
    mm.set_scalar(m, 3, 3, 1, 1, 0.0)
    _, st, I, _ = oracle.solve_C(m, np.ones(20, np.complex128), precond=oracle.PREC_DIAG)
    assert st == oracle.ZERO_DIAG and I[oracle.I_IERR] == oracle.ZERO_DIAG


# ---- variable blocks (n_dof = -1): monolis_matvec_V_C, diag_V_*_C -------------------
def _dense_V(m):
    return m.dense()


def test_matvec_V_C_matches_dense_brute_force():
    m = random_vblock_C(40, 0.12, seed=3, maxb=5)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(m.npndof) + 1j * rng.standard_normal(m.npndof)
    np.testing.assert_allclose(oracle.matvec_C(m, x), _dense_V(m) @ x, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("ndof", [1, 2, 3])
def test_V_C_with_one_block_size_is_the_fixed_block_path(ndof):
    """Every block of size b: the same sums in the same order as matvec_nn_C /
    diag_nn_*_C / COCG on the fixed-b matrix, so bit for bit on one thread."""
    oracle.set_threads(1)
    m = random_bcsr_C(30, ndof, 0.15, seed=30 + ndof)
    v = as_vblock(m)
    rng = np.random.default_rng(ndof)
    x = rng.standard_normal(m.n * ndof) + 1j * rng.standard_normal(m.n * ndof)
    assert np.array_equal(oracle.matvec_C(v, x), oracle.matvec_C(m, x))
    (doff, Dv), stv = oracle.diag_setup_C(v)
    Dm, stm = oracle.diag_setup_C(m)
    assert stv == stm == 0 and np.array_equal(Dv, Dm)
    assert np.array_equal(oracle.diag_apply_C(v, (doff, Dv), x), oracle.diag_apply_C(m, Dm, x))
    for precond in (oracle.PREC_NONE, oracle.PREC_DIAG):
        xv, sv, Iv, Rv = oracle.solve_C(v, x, precond=precond, tol=1e-12)
        xm, sm, Im, Rm = oracle.solve_C(m, x, precond=precond, tol=1e-12)
        assert sv == sm == 0 and Iv[oracle.I_CUR_ITER] == Im[oracle.I_CUR_ITER]
        assert np.array_equal(xv, xm)


def test_diag_V_C_inverts_each_diagonal_block_and_reports_zero_pivot():
    m = random_vblock_C(30, 0.15, seed=6, maxb=6)
    (doff, D), st = oracle.diag_setup_C(m)
    assert st == 0
    full = _dense_V(m)
    rng = np.random.default_rng(7)
    v = rng.standard_normal(m.npndof) + 1j * rng.standard_normal(m.npndof)
    y = oracle.diag_apply_C(m, (doff, D), v)
    for i in range(m.n):
        r0, n1 = int(m.n_dof_index[i]), int(m.n_dof_list[i])
        blk = full[r0:r0 + n1, r0:r0 + n1]
        np.testing.assert_allclose(y[r0:r0 + n1], np.linalg.solve(blk, v[r0:r0 + n1]), rtol=1e-12, atol=1e-12)
    # a zero pivot stops the reference (diag_V.f90:123); the oracle reports it
    big = int(np.argmax(m.n_dof_list[:m.n]))
    e = next(e for e in range(m.index[big], m.index[big + 1]) if m.item[e] - 1 == big)
    m.A[int(m.n_dof_index2[e])] = 0.0
    _, st = oracle.diag_setup_C(m)
    assert st == oracle.ZERO_DIAG


def test_cocg_V_matches_dense_solve_and_real_V_cg():
    m = random_vblock_C(80, 0.06, seed=9, maxb=4)
    rng = np.random.default_rng(10)
    b = rng.standard_normal(m.nndof) + 1j * rng.standard_normal(m.nndof)
    for precond in (oracle.PREC_NONE, oracle.PREC_DIAG):
        x, st, I, R = oracle.solve_C(m, b, precond=precond, tol=1e-12)
        assert st == 0
        np.testing.assert_allclose(x, np.linalg.solve(_dense_V(m), b), rtol=1e-9, atol=1e-9)
    # real-valued data: COCG on the complex V matrix is the real CG on the V matrix
    oracle.set_threads(1)
    mr = random_vblock(50, 0.08, seed=12, maxb=4, spd=True)
    mc = random_vblock(50, 0.08, seed=12, maxb=4, spd=True)
    mc.A = mr.A.astype(np.complex128)
    br = oracle.matvec(mr, np.ones(mr.npndof))
    xr, sr, Ir, Rr = oracle.solve(mr, br, method=oracle.ITER_CG, precond=oracle.PREC_DIAG, tol=1e-12)
    xc, sc, Ic, Rc = oracle.solve_C(mc, br.astype(np.complex128), precond=oracle.PREC_DIAG, tol=1e-12)
    assert sr == sc == 0 and Ir[oracle.I_CUR_ITER] == Ic[oracle.I_CUR_ITER]
    assert np.array_equal(xc.real, xr) and not xc.imag.any()
