"""Variable block sizes (monolis n_dof = -1) on the GPU, through the flat ABI
and the device API, against the oracle's matvec_V / diag_V / solvers.

Bars: the reference's matvec_V test inputs exactly; random SpMV within
1e-12 x max |y|; solves with the same status, iterations within +-2 % (+-1
below 50) of the oracle's, ||x - x_o|| / ||x_o|| <= 1e-6.
"""
import numpy as np
import pytest

import oracle
from monolis_amd import api
from monolis_amd import matrix as mm
from monolis_amd._lib import MonolisHipError

from _helpers import random_bcsr, random_vblock
from test_oracle_vblock import reference_matvec_v_case

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    oracle.set_threads(1)
    yield


def _iter_ok(got, want):
    return abs(int(got) - int(want)) <= max(1, int(0.02 * want))


def _grid_vblock(n, seed):
    """7-point grid graph on n^3 nodes with block sizes cycling 1, 2, 3 (SPD)."""
    N = n ** 3
    ids = np.arange(N).reshape(n, n, n)
    pairs = []
    for ax in range(3):
        a = np.take(ids, range(n - 1), axis=ax).reshape(-1)
        b = np.take(ids, range(1, n), axis=ax).reshape(-1)
        pairs += list(zip(a.tolist(), b.tolist()))
    lst = (np.arange(N) % 3 + 1).astype(np.int32)
    m = mm.nonzero_pattern_by_simple_mesh_V(N, lst, pairs)
    rng = np.random.default_rng(seed)
    # symmetric off-diagonal blocks U[-1, 0]; diagonal blocks: symmetric, dominant
    A = np.zeros(len(m.A))
    rows = np.repeat(np.arange(N), np.diff(m.index))
    for e in range(len(m.item)):
        i, j = rows[e], m.item[e] - 1
        if j < i:
            continue
        n1, n2 = lst[i], lst[j]
        blk = rng.uniform(-1, 0, (n1, n2))
        if i == j:
            blk = 0.5 * (blk + blk.T)
        A[m.n_dof_index2[e]:m.n_dof_index2[e] + n1 * n2] = blk.reshape(-1)
        if i != j:
            et = m.index[j] + np.searchsorted(m.item[m.index[j]:m.index[j + 1]], i + 1)
            A[m.n_dof_index2[et]:m.n_dof_index2[et] + n1 * n2] = blk.T.reshape(-1)
    m.A = A
    rowsum = np.abs(m.dense()).sum(axis=1)
    for i in range(N):
        e = m.index[i] + np.searchsorted(m.item[m.index[i]:m.index[i + 1]], i + 1)
        a0, nd = m.n_dof_index2[e], lst[i]
        blk = m.A[a0:a0 + nd * nd].reshape(nd, nd)
        for k in range(nd):
            r = m.n_dof_index[i] + k
            blk[k, k] = rowsum[r] - abs(blk[k, k]) + 1.0
        m.A[a0:a0 + nd * nd] = blk.reshape(-1)
    return m


def test_matvec_v_reference_inputs_exact():
    m, x = reference_matvec_v_case()
    mo = api.Monolis(m)
    y = api.monolis_matvec_product_R(mo, x)
    assert np.array_equal(y[:m.nndof], m.dense() @ x)


@pytest.mark.parametrize("seed", [1, 2])
def test_matvec_v_random(seed):
    m = random_vblock(400, 0.02, seed, maxb=6, spd=False)
    x = np.random.default_rng(seed).uniform(-1, 1, m.npndof)
    yo = oracle.matvec(m, x)
    y = api.monolis_matvec_product_R(api.Monolis(m), x)
    assert np.max(np.abs(y[:m.nndof] - yo[:m.nndof])) <= 1e-12 * np.max(np.abs(yo))
    dm = api.DeviceMatrix(m)
    assert dm.info()["ndof"] == -1
    dx, dy = api.DeviceBuffer.from_numpy(x), api.DeviceBuffer(m.npndof)
    dm.matvec(dx, dy)
    assert np.max(np.abs(dy.numpy()[:m.nndof] - yo[:m.nndof])) <= 1e-12 * np.max(np.abs(yo))


@pytest.mark.parametrize("method,precond", [(1, 1), (1, 0), (5, 1), (5, 0)])
def test_solve_v_flat_abi(method, precond):
    m = _grid_vblock(10, seed=3)
    b = np.random.default_rng(4).uniform(-1, 1, m.nndof)
    mo = api.Monolis(m)
    mo.prm.quiet()
    api.monolis_set_method(mo, method)
    api.monolis_set_precond(mo, precond)
    x = np.zeros(m.npndof)
    api.monolis_solve_R(mo, b, x)
    xo, st, I, R = oracle.solve(m, b, method=method, precond=precond)
    assert mo.prm.Iarray[5] == 0 and st == 0
    assert _iter_ok(mo.prm.Iarray[4], I[4]), (mo.prm.Iarray[4], I[4])
    assert np.linalg.norm(x[:m.nndof] - xo) / np.linalg.norm(xo) <= 1e-6


def test_device_v_set_values_refined_and_limits():
    m = random_vblock(300, 0.02, 5, maxb=5, spd=True)
    dm = api.DeviceMatrix(m)
    b = np.random.default_rng(6).uniform(-1, 1, m.nndof)
    db, dx = api.DeviceBuffer.from_numpy(b), api.DeviceBuffer(m.npndof)
    prm = api.Prm().quiet()
    dm.solve(db, dx, prm)
    xo, st, I, R = oracle.solve(m, b)
    assert prm.Iarray[5] == 0 and _iter_ok(prm.Iarray[4], I[4])
    assert np.linalg.norm(dx.numpy()[:m.nndof] - xo) / np.linalg.norm(xo) <= 1e-6
    # new values, same pattern
    m.A = m.A * 1.5
    dm.set_values(m.A)
    prm = api.Prm().quiet()
    dm.solve(db, dx, prm)
    xo, st, I, R = oracle.solve(m, b)
    assert prm.Iarray[5] == 0 and _iter_ok(prm.Iarray[4], I[4])
    assert np.linalg.norm(dx.numpy()[:m.nndof] - xo) / np.linalg.norm(xo) <= 1e-6
    # FP32-inner refinement with FP32 diag_V factors
    prm = api.Prm().quiet()
    prm.Rarray[1] = 1e-10
    stt = dm.solve_refined(db, dx, prm, inner_tol=1e-4)
    assert stt["converged"]
    xd = np.linalg.solve(m.dense(), b)
    assert np.linalg.norm(dx.numpy()[:m.nndof] - xd) / np.linalg.norm(xd) <= 1e-8
    # paths that need fixed blocks refuse variable blocks
    with pytest.raises(MonolisHipError):
        dm.solve_multi(db, dx, 1, api.Prm().quiet())
    with pytest.raises(MonolisHipError):
        dm.set_scalar(0, 0, 0, 0, 1.0)


def test_v_rejects_oversized_and_inconsistent_blocks():
    m = random_vblock(5, 0.5, 1, sizes=[1, 17, 2, 1, 1], spd=False)
    with pytest.raises(MonolisHipError):
        api.DeviceMatrix(m)
    m = random_vblock(5, 0.5, 1, sizes=[1, 2, 2, 1, 1], spd=False)
    m.n_dof_index = m.n_dof_index.copy()
    m.n_dof_index[2] += 1                      # disagrees with n_dof_list
    with pytest.raises(MonolisHipError):
        api.DeviceMatrix(m)


@pytest.mark.parametrize("ndof", [9, 12, 16])
def test_fixed_blocks_above_8_go_through_the_expansion(ndof):
    """monolis takes any fixed n_dof (matvec_nn, diag_nn); blocks above the
    fixed-block kernels' 8 are stored as the variable-block expansion."""
    m = random_bcsr(120, ndof, 0.03, seed=ndof, spd=True)
    x = np.random.default_rng(ndof).uniform(-1, 1, m.npndof)
    y = api.monolis_matvec_product_R(api.Monolis(m), x)[: m.nndof]
    yo = oracle.matvec(m, x)
    assert np.max(np.abs(y - yo)) <= 1e-12 * np.max(np.abs(yo))
    b = np.random.default_rng(ndof + 1).uniform(-1, 1, m.nndof)
    for method in (1, 5):
        mo = api.Monolis(m)
        mo.prm.quiet()
        api.monolis_set_method(mo, method)
        api.monolis_set_precond(mo, 1)
        api.monolis_set_tolerance(mo, 1e-10)
        xs = np.zeros(m.npndof)
        api.monolis_solve_R(mo, b, xs)
        xo, st, I, R = oracle.solve(m, b, method=method, precond=1, tol=1e-10)
        assert mo.prm.Iarray[5] == 0 and st == 0
        assert _iter_ok(mo.prm.Iarray[4], I[4]), (method, mo.prm.Iarray[4], I[4])
        assert np.linalg.norm(xs[: m.nndof] - xo) / np.linalg.norm(xo) <= 1e-6


def test_fixed_block_above_16_is_rejected():
    m = random_bcsr(10, 17, 0.3, seed=1, spd=True)
    with pytest.raises(MonolisHipError, match="16"):
        api.DeviceMatrix(m)
