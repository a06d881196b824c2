"""Complex path on the GPU (monolis's *_C routines and COCG, method 13) against
the oracle (monolis_oracle_cplx.c).  Bars: SpMV and dot within 1e-12
relative; solves with the same status, iterations within +-2 % (+-1 below 50)
of the oracle's and ||x - x_o|| / ||x_o|| <= 1e-6; the reference's own chain
tests to x = 1 + i."""
import ctypes

import numpy as np
import pytest

import oracle
from monolis_amd import _lib, api, workloads as W
from monolis_amd import matrix as mm

from _helpers import (as_vblock, chain_matrix, chain_matrix_C, chain_matrix_C_wrapper, random_bcsr_C, random_vblock,
                      random_vblock_C)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    oracle.set_threads(1)
    yield


def _iter_ok(got, want):
    return abs(int(got) - int(want)) <= max(1, int(0.02 * want))


def _crand(m, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(m) + 1j * rng.standard_normal(m)


def _flat_solve_C(m, b, precond, tol=1e-8, max_iter=1000, method=13, abort=True):
    mo = api.Monolis(m)
    mo.prm.quiet()
    api.monolis_set_method(mo, method)
    api.monolis_set_precond(mo, precond)
    api.monolis_set_tolerance(mo, tol)
    api.monolis_set_maxiter(mo, max_iter)
    api.monolis_set_error_abort(mo, abort)
    x = np.zeros(m.npndof, np.complex128)
    api.monolis_solve_C(mo, b, x)
    return x, mo.prm


@pytest.mark.parametrize("ndof", [1, 2, 3, 4, 5, 6, 7, 8])
def test_matvec_C_matches_oracle(ndof):
    # 300 block rows: five 64-row slices, the last one partial; ragged rows
    m = random_bcsr_C(300, ndof, 0.03, seed=ndof, symmetric=False)
    x = _crand(m.n * ndof, ndof)
    y = api.monolis_matvec_product_C(api.Monolis(m), x)[: m.n * ndof]
    yo = oracle.matvec_C(m, x)
    assert np.linalg.norm(y - yo) <= 1e-12 * np.linalg.norm(yo)


def test_matvec_C_device_handle_and_set_values():
    m = W.helmholtz7(20, shift=1.0, damping=0.3)
    dm = api.DeviceMatrixC(m)
    x = _crand(m.n, 4)
    dx, dy = api.DeviceBuffer.from_numpy(x), api.DeviceBuffer(m.n, np.complex128)
    dm.matvec(dx, dy)
    yo = oracle.matvec_C(m, x)
    assert np.linalg.norm(dy.numpy() - yo) <= 1e-13 * np.linalg.norm(yo)
    # new values, same pattern
    m2 = W.helmholtz7(20, shift=3.0, damping=0.1)
    dm.set_values(m2.A)
    dm.matvec(dx, dy)
    yo2 = oracle.matvec_C(m2, x)
    assert np.linalg.norm(dy.numpy() - yo2) <= 1e-13 * np.linalg.norm(yo2)
    info = dm.info()
    assert info["n"] == m.n and info["ndof"] == 1 and info["bytes"] > 0


@pytest.mark.parametrize("m_len", [1, 63, 64, 1000, 100_003])
def test_inner_product_C_is_unconjugated(m_len):
    x, y = _crand(m_len, 1), _crand(m_len, 2)
    # the flat entry point sums over N * n_dof entries of the handle's matrix
    mo = api.Monolis(mm.Matrix(m_len, m_len, 1, np.zeros(m_len + 1, np.int32), np.zeros(0, np.int32),
                               np.zeros(0, np.complex128)))
    got = api.monolis_inner_product_C(mo, 1, x, y)
    want = oracle.inner_product_C(x, y)
    assert abs(got - want) <= 1e-12 * np.sum(np.abs(x * y))
    # device buffers, and the (i, i) = -1 closed form
    d = api.DeviceBuffer.from_numpy(np.full(m_len, 1j))
    assert api.inner_product_device_C(d, d) == -m_len


@pytest.mark.parametrize("ndof", [1, 2, 3])
@pytest.mark.parametrize("precond", [0, 1])
def test_cocg_chain_x_is_one_plus_i(ndof, precond):
    """src_test/iterative/COCG_test.f90:71-85 through the flat C-ABI."""
    m = chain_matrix_C(10, ndof, seed=ndof)
    b = oracle.matvec_C(m, np.full(m.n * ndof, 1 + 1j))
    x, prm = _flat_solve_C(m, b, precond, tol=1e-12)
    assert prm.Iarray[5] == 0
    np.testing.assert_allclose(x[: m.n * ndof], 1 + 1j, rtol=1e-9, atol=1e-9)
    xo, st, I, R = oracle.solve_C(m, b, precond=precond, tol=1e-12)
    assert st == 0 and _iter_ok(prm.Iarray[4], I[4])


def test_cocg_wrapper_test_shape():
    """wrapper_test/solver/monolis_solver_c_test.c:83-158 (n_dof 2, tol 1e-10)."""
    m = chain_matrix_C_wrapper(seed=4)
    b = oracle.matvec_C(m, np.full(20, 1 + 1j))
    x, prm = _flat_solve_C(m, b, 1, tol=1e-10)
    assert prm.Iarray[5] == 0
    np.testing.assert_allclose(x, 1 + 1j, rtol=1e-8, atol=1e-8)


@pytest.mark.parametrize("ndof", [1, 2, 3, 4, 8])
@pytest.mark.parametrize("precond", [0, 1])
def test_cocg_random_matches_oracle(ndof, precond):
    m = random_bcsr_C(400, ndof, 0.01, seed=30 + ndof)
    b = _crand(m.n * ndof, ndof)
    x, prm = _flat_solve_C(m, b, precond, tol=1e-10)
    xo, st, I, R = oracle.solve_C(m, b, precond=precond, tol=1e-10)
    assert prm.Iarray[5] == st == 0
    assert _iter_ok(prm.Iarray[4], I[4]), (prm.Iarray[4], I[4])
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    # Rarray[2]: real part of the complex residual, as monolis stores it
    if prm.Iarray[4] == I[4]:
        assert abs(prm.Rarray[2] - R[2]) <= 1e-14 + 0.05 * abs(R[2])
    assert 0.0 <= prm.Rarray[2] < 1e-10


@pytest.mark.parametrize("precond", [0, 1])
def test_cocg_helmholtz_past_residual_replacement(precond):
    """About 400 iterations: crosses the true-residual step at 200 (COCG.f90:88-92)."""
    m = W.helmholtz7(32, shift=2.0, damping=0.1)
    b = _crand(m.n, 9)
    dm = api.DeviceMatrixC(m)
    db, dx = api.DeviceBuffer.from_numpy(b), api.DeviceBuffer(m.n, np.complex128)
    prm = api.Prm().quiet()
    prm.Iarray[1], prm.Iarray[2], prm.Iarray[3] = 13, precond, 2000
    prm.Rarray[1] = 1e-8
    hist = dm.solve(db, dx, prm, hist_len=2001)
    xo, st, I, R = oracle.solve_C(m, b, precond=precond, tol=1e-8, max_iter=2000)
    assert st == 0 and prm.Iarray[5] == 0 and I[4] > 200
    assert _iter_ok(prm.Iarray[4], I[4]), (prm.Iarray[4], I[4])
    x = dx.numpy()
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    # the final true residual meets the tolerance (|.| of the complex relative residual)
    r = b - oracle.matvec_C(m, x)
    assert np.sqrt(abs(np.sum(r * r)) / abs(np.sum(b * b))) < 2e-8
    it = int(prm.Iarray[4])
    assert hist[it] < 1e-8 <= hist[it - 1]


def test_cocg_real_data_matches_real_cg():
    """COCG on real-valued data is CG: the GPU COCG and the GPU CG agree."""
    mr = chain_matrix(10, 2, seed=2)
    mc = mm.Matrix(mr.n, mr.n, 2, mr.index, mr.item, mr.A.astype(np.complex128)).ensure_dof_index()
    b = oracle.matvec(mr, np.ones(20))
    xc, prmc = _flat_solve_C(mc, b.astype(np.complex128), 1, tol=1e-12)
    mo = api.Monolis(mr)
    mo.prm.quiet()
    api.monolis_set_tolerance(mo, 1e-12)
    xr = np.zeros(20)
    api.monolis_solve_R(mo, b, xr)
    assert _iter_ok(prmc.Iarray[4], mo.prm.Iarray[4])
    assert np.abs(xc.imag).max() == 0.0
    np.testing.assert_allclose(xc.real, xr, rtol=1e-10, atol=1e-10)


def test_cocg_zero_rhs_and_errors():
    m = chain_matrix_C(10, 2, seed=1)
    x, prm = _flat_solve_C(m, np.zeros(20, np.complex128), 1)
    assert prm.Iarray[5] == 0 and not x.any()
    # not COCG: monolis_solver_select_C rejects it
    _, prm = _flat_solve_C(m, np.ones(20, np.complex128), 1, method=1, abort=False)
    assert prm.Iarray[5] == -4
    _, prm = _flat_solve_C(m, np.ones(20, np.complex128), 2, abort=False)
    assert prm.Iarray[5] == -3
    # zero pivot in a diagonal block: monolis stops; here status -6
    mm.set_scalar(m, 3, 3, 1, 1, 0.0)
    _, prm = _flat_solve_C(m, np.ones(20, np.complex128), 1, abort=False)
    assert prm.Iarray[5] == -6
    # the same matrix without the preconditioner still solves
    x, prm = _flat_solve_C(m, oracle.matvec_C(m, np.full(20, 1 + 1j)), 0, tol=1e-12)
    assert prm.Iarray[5] == 0
    np.testing.assert_allclose(x, 1 + 1j, rtol=1e-8, atol=1e-8)


def test_cocg_not_converged_status_and_fixed_handle_rejects_variable_blocks():
    m = W.helmholtz7(16, shift=2.0, damping=0.05)
    b = _crand(m.n, 1)
    x, prm = _flat_solve_C(m, b, 1, tol=1e-12, max_iter=5, abort=False)
    assert prm.Iarray[5] == 1 and prm.Iarray[4] == 5
    mv = mm.nonzero_pattern_by_simple_mesh_V(3, [1, 2, 1], [[0, 1], [1, 2]])
    mv.A = mv.A.astype(np.complex128)
    h = ctypes.c_void_p()
    rc = _lib.load().monolis_hip_matrix_create_C(ctypes.byref(h), 3, 3, mv.nz, -1, api._p(mv.index),
                                                 api._p(mv.item), api._p(mv.A))
    assert rc == -2 and b"variable block" in _lib.load().monolis_hip_last_error()


def test_cocg_bcsr3_complex_workload():
    m = W.stencil27_bcsr3_complex(24)
    b = _crand(m.n * 3, 5)
    x, prm = _flat_solve_C(m, b, 1, tol=1e-10)
    xo, st, I, R = oracle.solve_C(m, b, precond=1, tol=1e-10)
    assert prm.Iarray[5] == st == 0 and _iter_ok(prm.Iarray[4], I[4])
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)


# ---- variable blocks (n_dof = -1): monolis_matvec_V_C, diag_V_*_C ------------------
@pytest.mark.parametrize("maxb", [3, 6])
def test_matvec_V_C_matches_oracle(maxb):
    m = random_vblock_C(300, 0.02, seed=maxb, maxb=maxb)
    x = _crand(m.npndof, maxb)
    y = api.monolis_matvec_product_C(api.Monolis(m), x)[: m.nndof]
    yo = oracle.matvec_C(m, x)
    assert np.linalg.norm(y - yo) <= 1e-12 * np.linalg.norm(yo)
    dm = api.DeviceMatrixC(m)
    dx, dy = api.DeviceBuffer.from_numpy(x), api.DeviceBuffer(m.npndof, np.complex128)
    dm.matvec(dx, dy)
    assert np.linalg.norm(dy.numpy()[: m.nndof] - yo) <= 1e-12 * np.linalg.norm(yo)


@pytest.mark.parametrize("precond", [0, 1])
def test_cocg_V_matches_oracle(precond):
    m = random_vblock_C(400, 0.012, seed=17, maxb=5)
    b = _crand(m.nndof, 18)
    x, prm = _flat_solve_C(m, b, precond, tol=1e-10)
    xo, st, I, R = oracle.solve_C(m, b, precond=precond, tol=1e-10)
    assert prm.Iarray[5] == st == 0 and _iter_ok(prm.Iarray[4], I[4]), (prm.Iarray[4], I[4])
    assert np.linalg.norm(x[: m.nndof] - xo) <= 1e-6 * np.linalg.norm(xo)


def test_cocg_V_with_one_block_size_is_the_fixed_block_path():
    m = random_bcsr_C(500, 3, 0.01, seed=19)
    b = _crand(m.n * 3, 20)
    xf, pf = _flat_solve_C(m, b, 1, tol=1e-10)
    xv, pv = _flat_solve_C(as_vblock(m), b, 1, tol=1e-10)
    assert pf.Iarray[5] == pv.Iarray[5] == 0 and abs(int(pf.Iarray[4]) - int(pv.Iarray[4])) <= 1
    assert np.linalg.norm(xv - xf) <= 1e-9 * np.linalg.norm(xf)


def test_cocg_V_real_data_matches_real_V_cg():
    mr = random_vblock(300, 0.015, seed=21, maxb=4, spd=True)
    mc = random_vblock(300, 0.015, seed=21, maxb=4, spd=True)
    mc.A = mr.A.astype(np.complex128)
    b = oracle.matvec(mr, np.ones(mr.npndof))
    xc, pc = _flat_solve_C(mc, b.astype(np.complex128), 1, tol=1e-10)
    mo = api.Monolis(mr)
    mo.prm.quiet()
    api.monolis_set_tolerance(mo, 1e-10)
    xr = np.zeros(mr.npndof)
    api.monolis_solve_R(mo, b, xr)
    assert pc.Iarray[5] == 0 and _iter_ok(pc.Iarray[4], mo.prm.Iarray[4])
    assert np.abs(xc.imag).max() == 0.0
    np.testing.assert_allclose(xc.real, xr, rtol=1e-10, atol=1e-10)


def test_cocg_V_zero_pivot_reported_then_cleared_by_set_values():
    m = random_vblock_C(120, 0.03, seed=23, maxb=4)
    b = _crand(m.nndof, 24)
    big = int(np.argmax(m.n_dof_list[: m.n]))
    e = next(e for e in range(m.index[big], m.index[big + 1]) if m.item[e] - 1 == big)
    good = m.A.copy()
    m.A[int(m.n_dof_index2[e])] = 0.0
    # the reference stops in diag_V_setup_C (diag_V.f90:123); here status -6
    _, prm = _flat_solve_C(m, b, 1, abort=False)
    assert prm.Iarray[5] == -6
    dm = api.DeviceMatrixC(m)
    db, dx = api.DeviceBuffer.from_numpy(b), api.DeviceBuffer(m.npndof, np.complex128)
    prm = api.Prm().quiet()
    prm.Iarray[1], prm.Iarray[2] = 13, 1
    with pytest.raises(_lib.MonolisHipError, match="zero diag"):
        dm.solve(db, dx, prm)
    prm.Iarray[2] = 0                                   # without the preconditioner it solves
    dm.solve(db, dx, prm)
    assert prm.Iarray[5] == 0
    dm.set_values(good)                                  # new values: the factors are rebuilt
    prm = api.Prm().quiet()
    prm.Iarray[1], prm.Iarray[2] = 13, 1
    prm.Rarray[1] = 1e-10
    dm.solve(db, dx, prm)
    m.A = good
    xo, st, I, R = oracle.solve_C(m, b, precond=1, tol=1e-10)
    assert prm.Iarray[5] == st == 0 and _iter_ok(prm.Iarray[4], I[4])
    assert np.linalg.norm(dx.numpy()[: m.nndof] - xo) <= 1e-6 * np.linalg.norm(xo)


@pytest.mark.parametrize("ndof", [10, 16])
def test_cocg_fixed_blocks_above_8(ndof):
    m = random_bcsr_C(80, ndof, 0.04, seed=ndof)
    x = _crand(m.n * ndof, ndof)
    y = api.monolis_matvec_product_C(api.Monolis(m), x)[: m.n * ndof]
    yo = oracle.matvec_C(m, x)
    assert np.linalg.norm(y - yo) <= 1e-12 * np.linalg.norm(yo)
    b = _crand(m.n * ndof, ndof + 1)
    xs, prm = _flat_solve_C(m, b, 1, tol=1e-10)
    xo, st, I, R = oracle.solve_C(m, b, precond=1, tol=1e-10)
    assert prm.Iarray[5] == st == 0 and _iter_ok(prm.Iarray[4], I[4])
    assert np.linalg.norm(xs[: m.nndof] - xo) <= 1e-6 * np.linalg.norm(xo)
