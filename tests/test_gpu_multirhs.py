"""Multi-right-hand-side CG (one SpMM per iteration for all columns) against one
oracle solve per column.

Bars: status identical, iteration count within +-2 % (+-1 below 50) of the
oracle's solve of that column alone, ||x_j - x_j,oracle|| / ||x_j,oracle|| <= 1e-6
for FP64 (both stop at the same 1e-8 relative residual); the FP32-refined solve
reaches its FP64 relative residual and agrees with a tight FP64 oracle solve.
"""
import numpy as np
import pytest

import oracle
from monolis_amd import api, workloads as W
from monolis_amd._lib import MonolisHipError

from _helpers import random_bcsr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    oracle.set_threads(1)
    yield


def _prm(tol=1e-8, precond=1, max_iter=1000):
    prm = api.Prm().quiet()
    prm.Iarray[2] = precond
    prm.Iarray[3] = max_iter
    prm.Rarray[1] = tol
    return prm


def _iter_ok(got, want):
    return abs(int(got) - int(want)) <= max(1, int(0.02 * want))


def _check_columns(mat, B, X, res, precond, tol=1e-8):
    k = B.shape[1]
    for j in range(k):
        if not np.any(B[:, j]):
            assert res["status"][j] == 2 and np.all(X[:, j] == 0.0)
            continue
        xo, st, I, R = oracle.solve(mat, B[:, j], precond=precond, tol=tol)
        assert res["status"][j] == 0 and st == 0, (j, res["status"][j], st)
        assert _iter_ok(res["iters"][j], I[4]), (j, res["iters"][j], I[4])
        assert res["resid"][j] < tol
        assert np.linalg.norm(X[:, j] - xo) / np.linalg.norm(xo) <= 1e-6, j


@pytest.mark.parametrize("precond", [0, 1])
@pytest.mark.parametrize("k", [1, 7, 64])
def test_multi_matches_columnwise_oracle(precond, k):
    mat = W.poisson7(12)
    rng = np.random.default_rng(10 + k)
    B = rng.uniform(-1, 1, size=(mat.n, k))
    if k > 1:
        B[:, 1] = 0.0                                   # zero right-hand side -> x = 0, status 2
    if k > 3:
        B[:, 3] = 0.0
        B[mat.n // 2, 3] = 1.0                          # point source: different iteration count
    dm = api.DeviceMatrix(mat)
    db = api.DeviceBuffer.from_numpy(B)
    dx = api.DeviceBuffer(mat.n * k)
    res = dm.solve_multi(db, dx, k, _prm(precond=precond))
    X = dx.numpy().reshape(mat.n, k)
    _check_columns(mat, B, X, res, precond)


def test_multi_ragged_pattern_diag():
    mat = random_bcsr(300, 1, 0.02, seed=5, spd=True)
    k = 5
    B = np.random.default_rng(6).uniform(-1, 1, size=(mat.n, k))
    dm = api.DeviceMatrix(mat)
    db = api.DeviceBuffer.from_numpy(B)
    dx = api.DeviceBuffer(mat.n * k)
    res = dm.solve_multi(db, dx, k, _prm(precond=1))
    _check_columns(mat, B, dx.numpy().reshape(mat.n, k), res, 1)


def test_multi_residual_replacement_beyond_200():
    # unpreconditioned 48^3 Poisson needs more than 200 iterations at 1e-10,
    # so the true-residual replacement step runs.
    mat = W.poisson7(48)
    k = 3
    B = np.random.default_rng(2).uniform(-1, 1, size=(mat.n, k))
    dm = api.DeviceMatrix(mat)
    db = api.DeviceBuffer.from_numpy(B)
    dx = api.DeviceBuffer(mat.n * k)
    res = dm.solve_multi(db, dx, k, _prm(precond=0, tol=1e-10))
    assert np.all(res["iters"] > 200)
    _check_columns(mat, B, dx.numpy().reshape(mat.n, k), res, 0, tol=1e-10)


def test_multi_max_iter_and_initial_guess():
    mat = W.poisson7(10)
    k = 4
    B = np.random.default_rng(3).uniform(-1, 1, size=(mat.n, k))
    dm = api.DeviceMatrix(mat)
    db = api.DeviceBuffer.from_numpy(B)
    dx = api.DeviceBuffer(mat.n * k)
    res = dm.solve_multi(db, dx, k, _prm(max_iter=5))
    assert np.all(res["status"] == 1) and np.all(res["iters"] == 5)
    X5 = dx.numpy().reshape(mat.n, k)
    for j in range(k):
        Io, Ro = oracle.default_arrays(max_iter=5)
        Io[16] = 0
        xo, st, I, R = oracle.solve(mat, B[:, j], Iarray=Io, Rarray=Ro)
        assert st == 1 and np.linalg.norm(X5[:, j] - xo) <= 1e-10 * np.linalg.norm(xo)
    # continue from the current x (is_init_x = 0) to convergence
    prm = _prm()
    prm.Iarray[10] = 0
    res = dm.solve_multi(db, dx, k, prm)
    assert np.all(res["status"] == 0)
    X = dx.numpy().reshape(mat.n, k)
    for j in range(k):
        Io, Ro = oracle.default_arrays()
        Io[10] = 0
        xo, st, I, R = oracle.solve(mat, B[:, j], x0=X5[:, j], Iarray=Io, Rarray=Ro)
        assert _iter_ok(res["iters"][j], I[4])
        assert np.linalg.norm(X[:, j] - xo) / np.linalg.norm(xo) <= 1e-6


def test_multi_refined_reaches_fp64_accuracy():
    mat = W.poisson7(20)
    k = 16
    B = np.random.default_rng(4).uniform(-1, 1, size=(mat.n, k))
    B[:, 2] = 0.0
    dm = api.DeviceMatrix(mat)
    db = api.DeviceBuffer.from_numpy(B)
    dx = api.DeviceBuffer(mat.n * k)
    res = dm.solve_multi_refined(db, dx, k, _prm(tol=1e-10), inner_tol=1e-4)
    X = dx.numpy().reshape(mat.n, k)
    S = mat.to_scipy()
    assert res["outer"] >= 2
    for j in range(k):
        if j == 2:
            assert res["status"][j] == 2 and np.all(X[:, j] == 0.0)
            continue
        assert res["status"][j] == 0 and res["resid"][j] < 1e-10
        # FP64 relative residual measured on the host
        assert np.linalg.norm(B[:, j] - S @ X[:, j]) / np.linalg.norm(B[:, j]) < 1.5e-10
        xo, st, I, R = oracle.solve(mat, B[:, j], tol=1e-12)
        assert np.linalg.norm(X[:, j] - xo) / np.linalg.norm(xo) <= 1e-8


def test_multi_rejects_bad_arguments():
    mat = W.poisson7(6)
    dm = api.DeviceMatrix(mat)
    db = api.DeviceBuffer(mat.n * 65)
    dx = api.DeviceBuffer(mat.n * 65)
    with pytest.raises(MonolisHipError):
        dm.solve_multi(db, dx, 65, _prm())
    with pytest.raises(MonolisHipError):
        dm.solve_multi(db, dx, 0, _prm())
    prm = _prm()
    prm.Iarray[1] = 5                                    # BiCGSTAB is not a multi-RHS method
    with pytest.raises(MonolisHipError):
        dm.solve_multi(db, dx, 2, prm)
    blk = random_bcsr(20, 3, 0.2, seed=1, spd=True)     # block matrices are rejected
    dmb = api.DeviceMatrix(blk)
    with pytest.raises(MonolisHipError):
        dmb.solve_multi(db, dx, 2, _prm())


@pytest.mark.parametrize("k", [5, 16, 33, 64])
def test_multi_every_spmm_config(k):
    """Every SpMM launch configuration gives the same solution (within the
    solve tolerance) and the same iteration counts (+-1: reduction grouping
    follows the grid); the autotuner picks one of them."""
    mat = random_bcsr(700, 1, 0.01, seed=k, spd=True)
    B = np.random.default_rng(k).uniform(-1, 1, size=(mat.n, k))
    dm = api.DeviceMatrix(mat)
    db = api.DeviceBuffer.from_numpy(B)
    dx = api.DeviceBuffer(mat.n * k)
    tune = dm.autotune_multi(k, fp32=False, reps=2)
    assert 0 <= tune["best"] < len(tune["configs"])
    ref = None
    for cfg in range(len(tune["configs"])):
        dm.set_multi_config(cfg)
        res = dm.solve_multi(db, dx, k, _prm())
        X = dx.numpy().reshape(mat.n, k)
        assert np.all(res["status"] == 0)
        if ref is None:
            ref = (X, res["iters"])
            _check_columns(mat, B, X, res, 1)
        else:
            assert np.all(np.abs(res["iters"] - ref[1]) <= 1)
            assert np.linalg.norm(X - ref[0]) / np.linalg.norm(ref[0]) <= 1e-6


@pytest.mark.parametrize("k", [7, 20, 64])
@pytest.mark.parametrize("fp32", [False, True])
def test_matvec_multi_every_config_bit_identical(k, fp32):
    """Every SpMM configuration -- including the ones that take the right-hand
    sides in chunks of 8, 16 or 32 per pass over the matrix, with a ragged last
    chunk at k = 7 and 20 -- forms each entry with the same multiply-add chain,
    so all give bit-identical Y; config 0 agrees with the column-wise oracle."""
    import oracle
    mat = random_bcsr(411, 1, 0.03, seed=k + 11, spd=False)
    X = np.random.default_rng(k + 2).uniform(-1, 1, size=(mat.n, k))
    dt = np.float32 if fp32 else np.float64
    dm = api.DeviceMatrix(mat)
    dx = api.DeviceBuffer.from_numpy(X.astype(dt))
    dy = api.DeviceBuffer(mat.n * k, dt)
    ncfg = len(dm.autotune_multi(k, fp32=fp32, reps=1)["configs"])
    ref = None
    for cfg in range(ncfg):
        dm.set_multi_config(cfg, fp32=fp32)
        dm.matvec_multi(dx, dy, k)
        Y = dy.numpy().reshape(mat.n, k)
        if ref is None:
            ref = Y.copy()
            Yo = np.stack([oracle.matvec(mat, X.astype(dt).astype(np.float64)[:, j]) for j in range(k)], axis=1)
            tol = 1e-5 if fp32 else 1e-12
            assert np.max(np.abs(Y - Yo)) <= tol * np.max(np.abs(Yo))
        else:
            assert np.array_equal(Y, ref), f"config {cfg} differs"


@pytest.mark.parametrize("k", [1, 3, 8, 20, 64])
@pytest.mark.parametrize("fp32", [False, True])
def test_matvec_multi_matches_columnwise_oracle(k, fp32):
    """SpMM column j == oracle matvec of column j: FP64 within 1e-12 relative
    (a different but fixed summation order), FP32 within 1e-5."""
    mat = random_bcsr(333, 1, 0.03, seed=k, spd=False)
    X = np.random.default_rng(k + 1).uniform(-1, 1, size=(mat.n, k))
    dt = np.float32 if fp32 else np.float64
    dm = api.DeviceMatrix(mat)
    dx = api.DeviceBuffer.from_numpy(X.astype(dt))
    dy = api.DeviceBuffer(mat.n * k, dt)
    dm.matvec_multi(dx, dy, k)
    Y = dy.numpy().reshape(mat.n, k)
    tol = 1e-5 if fp32 else 1e-12
    for j in range(k):
        yo = oracle.matvec(mat, X[:, j])
        assert np.linalg.norm(Y[:, j] - yo) / np.linalg.norm(yo) <= tol, j
