"""PipeCR (monolis method 4, src/iterative/PipeCR.f90) on the GPU against the
oracle's.  Bars as for CG: same status, iterations within +-2 % (+-1 below
50), ||x - x_o|| / ||x_o|| <= 1e-6; the reported residual (the preconditioned
residual ratio ||M^-1 r|| / ||M^-1 r0||) below tol."""
import numpy as np
import pytest

import oracle
from monolis_amd import api, workloads as W

from _helpers import chain_matrix, random_vblock

pytestmark = pytest.mark.gpu

PIPECR = api.MONOLIS_ITER_PIPECR


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    oracle.set_threads(1)
    yield


def _iter_ok(got, want):
    return abs(int(got) - int(want)) <= max(1, int(0.02 * want))


def _flat_solve(m, b, precond, tol=1e-8, max_iter=1000):
    mo = api.Monolis(m)
    mo.prm.quiet()
    api.monolis_set_method(mo, PIPECR)
    api.monolis_set_precond(mo, precond)
    api.monolis_set_tolerance(mo, tol)
    api.monolis_set_maxiter(mo, max_iter)
    x = np.zeros(m.npndof)
    api.monolis_solve_R(mo, b, x)
    return x, mo.prm


@pytest.mark.parametrize("ndof", [1, 2, 3])
@pytest.mark.parametrize("precond", [0, 1])
def test_pipecr_chain_x_is_one(ndof, precond):
    """src_test/iterative/PipeCR_test.f90:21-90 through the flat C ABI."""
    m = chain_matrix(10, ndof, seed=ndof)
    b = oracle.matvec(m, np.ones(m.npndof))
    x, prm = _flat_solve(m, b, precond, tol=1e-10)
    xo, st, I, R = oracle.solve(m, b, method=PIPECR, precond=precond, tol=1e-10)
    assert prm.Iarray[5] == 0 and st == 0
    assert _iter_ok(prm.Iarray[4], I[4]), (prm.Iarray[4], I[4])
    assert 0.0 < prm.Rarray[2] <= 1e-10
    assert np.allclose(x[:m.nndof], 1.0, rtol=0, atol=1e-7)


@pytest.mark.parametrize("case", ["poisson16", "bcsr3_8", "poisson40_none", "vblock"])
def test_pipecr_matches_oracle(case):
    if case == "poisson16":
        m, precond = W.poisson7(16), 1
    elif case == "bcsr3_8":
        m, precond = W.stencil27_bcsr3(8, seed=2), 1
    elif case == "poisson40_none":
        m, precond = W.poisson7(40), 0        # > 50 iterations: residual replacement runs
    else:
        m, precond = random_vblock(400, 0.015, 3, maxb=4, spd=True), 1
    b = np.random.default_rng(3).uniform(-1, 1, m.nndof)
    x, prm = _flat_solve(m, b, precond, tol=1e-10)
    xo, st, I, R = oracle.solve(m, b, method=PIPECR, precond=precond, tol=1e-10)
    assert prm.Iarray[5] == 0 and st == 0
    assert _iter_ok(prm.Iarray[4], I[4]), (prm.Iarray[4], I[4])
    if case == "poisson40_none":
        assert prm.Iarray[4] > 50
    assert 0.0 < prm.Rarray[2] <= 1e-10
    assert np.linalg.norm(x[:m.nndof] - xo) / np.linalg.norm(xo) <= 1e-6


def test_pipecr_refined_zero_rhs_and_max_iter():
    m = W.stencil27_bcsr3(10, seed=4)
    b = np.random.default_rng(5).uniform(-1, 1, m.nndof)
    dm = api.DeviceMatrix(m)
    db, dx = api.DeviceBuffer.from_numpy(b), api.DeviceBuffer(m.npndof)
    # FP32 PipeCR inner solves refined in FP64
    prm = api.Prm().quiet()
    prm.Iarray[1] = PIPECR
    prm.Rarray[1] = 1e-10
    st = dm.solve_refined(db, dx, prm, inner_tol=1e-4)
    assert st["converged"]
    xo, sto, I, R = oracle.solve(m, b, method=oracle.ITER_CG, tol=1e-12)
    assert np.linalg.norm(dx.numpy()[:m.nndof] - xo) / np.linalg.norm(xo) <= 1e-8
    # zero right-hand side: x = 0, converged
    z = api.DeviceBuffer.from_numpy(np.zeros(m.nndof))
    prm = api.Prm().quiet()
    prm.Iarray[1] = PIPECR
    dm.solve(z, dx, prm)
    assert prm.Iarray[5] == 0 and np.all(dx.numpy() == 0.0)
    # max_iter reached: status 1, four updates, as the oracle
    prm = api.Prm().quiet()
    prm.Iarray[1], prm.Iarray[3] = PIPECR, 4
    dm.solve(db, dx, prm)
    Io, Ro = oracle.default_arrays(method=PIPECR, max_iter=4)
    Io[16] = 0
    xo, sto, I, R = oracle.solve(m, b, Iarray=Io, Rarray=Ro)
    assert prm.Iarray[5] == 1 and sto == 1 and prm.Iarray[4] == 4
    assert np.linalg.norm(dx.numpy()[:m.nndof] - xo) / np.linalg.norm(xo) <= 1e-10


def test_pipecr_flat_abi_does_not_stop_at_max_iter():
    """PipeCR.f90:81-123 ends the loop silently at max_iter: the flat entry
    point returns (status 1) even with error_abort set."""
    m = W.poisson7(12)
    b = np.random.default_rng(6).uniform(-1, 1, m.nndof)
    mo = api.Monolis(m)
    mo.prm.quiet()
    mo.prm.Iarray[16] = 1
    api.monolis_set_method(mo, PIPECR)
    api.monolis_set_precond(mo, 1)
    api.monolis_set_maxiter(mo, 3)
    x = np.zeros(m.npndof)
    api.monolis_solve_R(mo, b, x)
    assert mo.prm.Iarray[5] == 1 and mo.prm.Iarray[4] == 3
    Io, Ro = oracle.default_arrays(method=PIPECR, precond=1, max_iter=3)
    Io[16] = 0
    xo, sto, I, R = oracle.solve(m, b, Iarray=Io, Rarray=Ro)
    assert np.linalg.norm(x[:m.nndof] - xo) / np.linalg.norm(xo) <= 1e-10
