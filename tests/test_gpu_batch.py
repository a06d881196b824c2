"""Batched solves (many independent systems, one persistent graph-captured
launch) against one oracle solve per system.

Bars: status identical, iteration count within +-2 % (+-1 below 50) of the
oracle's solve of that system alone, ||x_k - x_k,oracle|| / ||x_k,oracle|| <= 1e-6.
The team configurations (the largest systems split over several workgroups)
are held to the same bars, on batches whose largest system gets a team, one
with a zero right-hand side and one past the residual replacement at
iteration 200.
"""
import numpy as np
import pytest

import oracle
from monolis_amd import api, workloads as W
from monolis_amd.matrix import Matrix

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    oracle.set_threads(1)
    yield


def _system(mat, sys_row, k):
    r0, r1 = int(sys_row[k]), int(sys_row[k + 1])
    j0, j1 = int(mat.index[r0]), int(mat.index[r1])
    index = (mat.index[r0:r1 + 1] - j0).astype(np.int32)
    item = (mat.item[j0:j1] - r0).astype(np.int32)
    return Matrix(r1 - r0, r1 - r0, 1, index, item, mat.A[j0:j1].copy()).ensure_dof_index()


def _prm(tol=1e-8, precond=1, max_iter=1000):
    prm = api.Prm().quiet()
    prm.Iarray[2] = precond
    prm.Iarray[3] = max_iter
    prm.Rarray[1] = tol
    return prm


@pytest.mark.parametrize("precond", [0, 1])
def test_batch_matches_individual_solves(precond):
    mat, sys_row = W.poisson_batch(count=40, lo=3, hi=13, seed=7)
    b = W.uniform_vector(mat.n, seed=8)
    b[sys_row[5]:sys_row[6]] = 0.0                 # one zero right-hand side
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(mat.n)
    B = api.DeviceBatch(mat, sys_row)
    iters, resid, status = B.solve(db, dx, _prm(precond=precond))
    x = dx.numpy()
    for k in range(B.nsys):
        sk = _system(mat, sys_row, k)
        bk = b[sys_row[k]:sys_row[k + 1]]
        xo, st, I, R = oracle.solve(sk, bk, precond=precond)
        xk = x[sys_row[k]:sys_row[k + 1]]
        if k == 5:
            assert status[k] == 2 and np.all(xk == 0.0)
            continue
        assert status[k] == 0 and st == 0
        assert abs(int(iters[k]) - int(I[4])) <= max(1, int(0.02 * I[4])), (k, iters[k], I[4])
        assert np.linalg.norm(xk - xo) / np.linalg.norm(xo) <= 1e-6
        assert resid[k] < 1e-8


def test_batch_configs_graph_replay_and_autotune():
    mat, sys_row = W.poisson_batch(count=24, lo=4, hi=20, seed=3)
    b = W.uniform_vector(mat.n, seed=4)
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(mat.n)
    B = api.DeviceBatch(mat, sys_row)
    prm = _prm()
    it0, r0, s0 = B.solve(db, dx, prm)
    x0 = dx.numpy()
    it1, r1, s1 = B.solve(db, dx, prm)          # graph replay: same bits
    assert np.array_equal(x0, dx.numpy()) and np.array_equal(it0, it1)
    tune = B.autotune(db, dx, prm, reps=1)
    assert 0 <= tune["best"] < len(tune["ms"])
    it2, r2, s2 = B.solve(db, dx, prm)          # tuned configuration
    assert np.all(s2 == 0) and np.all(np.abs(it2 - it0) <= 1)
    assert np.linalg.norm(dx.numpy() - x0) / np.linalg.norm(x0) <= 1e-9


def _check_against_oracle(mat, sys_row, b, x, iters, resid, status, precond=1, zero=()):
    for k in range(len(sys_row) - 1):
        xk = x[sys_row[k]:sys_row[k + 1]]
        if k in zero:
            assert status[k] == 2 and np.all(xk == 0.0), k
            continue
        sk = _system(mat, sys_row, k)
        xo, st, I, R = oracle.solve(sk, b[sys_row[k]:sys_row[k + 1]], precond=precond)
        assert status[k] == 0 and st == 0, (k, status[k])
        assert abs(int(iters[k]) - int(I[4])) <= max(1, int(0.02 * I[4])), (k, iters[k], I[4])
        assert np.linalg.norm(xk - xo) / np.linalg.norm(xo) <= 1e-6, k
        assert resid[k] < 1e-8


@pytest.mark.parametrize("cfg", [3, 5, 13])
def test_batch_one_workgroup_configs_match_individual_solves(cfg):
    # 1024-thread workgroups, plain and non-temporal vector streams
    mat, sys_row = W.poisson_batch(count=24, lo=4, hi=20, seed=3)
    b = W.uniform_vector(mat.n, seed=4)
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(mat.n)
    B = api.DeviceBatch(mat, sys_row)
    assert B.schedule(cfg)["split"] == 0
    B.set_config(cfg)
    iters, resid, status = B.solve(db, dx, _prm())
    _check_against_oracle(mat, sys_row, b, dx.numpy(), iters, resid, status)


@pytest.mark.parametrize("cfg", [7, 8, 9, 10, 11, 12, 14])
def test_batch_team_configs_match_individual_solves(cfg):
    mat, sys_row = W.poisson_batch(count=24, lo=4, hi=20, seed=3)
    sizes = np.diff(sys_row)
    big = int(np.argmax(sizes))
    zero = int(np.argsort(sizes)[-2])               # the second largest system: zero right-hand side
    b = W.uniform_vector(mat.n, seed=4)
    b[sys_row[zero]:sys_row[zero + 1]] = 0.0
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(mat.n)
    B = api.DeviceBatch(mat, sys_row)
    sch = B.schedule(cfg)
    assert sch["split"] > 0 and sch["teams"] >= 1 and sch["slots"] > B.nsys, sch
    B.set_config(cfg)
    iters, resid, status = B.solve(db, dx, _prm())
    x = dx.numpy()
    print(f"cfg {cfg}: {sch}, largest system {sizes[big]} rows, {iters[big]} iterations", flush=True)
    _check_against_oracle(mat, sys_row, b, x, iters, resid, status, zero=(zero,))
    it1, r1, s1 = B.solve(db, dx, _prm())        # graph replay: same bits
    assert np.array_equal(x, dx.numpy()) and np.array_equal(iters, it1) and np.array_equal(status, s1)


@pytest.mark.parametrize("cfg", [7, 10])
def test_batch_team_residual_replacement(cfg):
    # one 56^3 system (about 210 DIAG-CG iterations) among small ones: its team
    # passes the true-residual replacement at iteration 200
    edges = [56] + [5, 9, 12, 7, 14, 6, 11, 8]
    mat, sys_row = W.poisson_batch(edges=edges)
    b = W.uniform_vector(mat.n, seed=21)
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(mat.n)
    B = api.DeviceBatch(mat, sys_row)
    sch = B.schedule(cfg)
    assert sch["teams"] >= 1, sch
    B.set_config(cfg)
    iters, resid, status = B.solve(db, dx, _prm())
    assert iters[0] > 200, iters[0]
    _check_against_oracle(mat, sys_row, b, dx.numpy(), iters, resid, status)


def test_batch_team_wait_limit_gives_up_cleanly():
    # a zero wait limit makes any team member that arrives before its teammates
    # give up at once (every wait still ends): the call reports an error or, if
    # no member ever had to wait, succeeds; afterwards the default limit solves
    # the batch correctly (the abort flag and counters are reset per launch)
    mat, sys_row = W.poisson_batch(count=24, lo=4, hi=20, seed=3)
    b = W.uniform_vector(mat.n, seed=4)
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(mat.n)
    B = api.DeviceBatch(mat, sys_row)
    B.set_config(10)
    B.set_wait_limit(0.0)
    try:
        B.solve(db, dx, _prm())
        gave_up = False
    except RuntimeError as e:
        gave_up = True
        assert "gave up" in str(e)
    # timed replays report a give-up in any replay, not only the last one
    try:
        B.time(db, dx, _prm(), reps=3)
        timed_gave_up = False
    except RuntimeError as e:
        timed_gave_up = True
        assert "gave up" in str(e)
    B.set_wait_limit(5.0)
    B.time(db, dx, _prm(), reps=2)                # no give-up left over from the runs above
    iters, resid, status = B.solve(db, dx, _prm())
    print(f"zero wait limit gave up: solve {gave_up}, timed replays {timed_gave_up}", flush=True)
    _check_against_oracle(mat, sys_row, b, dx.numpy(), iters, resid, status)


def test_batch_long_range_coupling():
    # a 40,000-row system whose first and last rows couple 39,999 rows apart (a
    # SELL slice whose columns span the whole system) beside Poisson systems,
    # one workgroup per system and with teams
    n = 40000
    rows, cols, vals = [], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 4.0), (i + 1, -1.0)):
            if 0 <= j < n:
                rows.append(i); cols.append(j); vals.append(v)
    rows += [0, n - 1]; cols += [n - 1, 0]; vals += [-1.0, -1.0]
    order = np.lexsort((cols, rows))
    rows, cols, vals = np.array(rows)[order], np.array(cols)[order], np.array(vals)[order]
    pm, prow = W.poisson_batch(edges=[12, 7, 20])
    index = np.concatenate([np.searchsorted(rows, np.arange(n + 1)).astype(np.int64),
                            pm.index[1:].astype(np.int64) + len(rows)]).astype(np.int32)
    item = np.concatenate([cols + 1, pm.item.astype(np.int64) + n]).astype(np.int32)
    A = np.concatenate([vals, pm.A])
    mat = Matrix(n + pm.n, n + pm.n, 1, index, item, A).ensure_dof_index()
    sys_row = np.concatenate([[0], np.asarray(prow, np.int64) + n])
    b = W.uniform_vector(mat.n, seed=5)
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(mat.n)
    for cfg in (5, 11):
        B = api.DeviceBatch(mat, sys_row)
        B.set_config(cfg)
        iters, resid, status = B.solve(db, dx, _prm())
        _check_against_oracle(mat, sys_row, b, dx.numpy(), iters, resid, status)


@pytest.mark.parametrize("cfg", [11, 14])
def test_batch_team_no_preconditioner_and_iteration_limit(cfg):
    mat, sys_row = W.poisson_batch(count=24, lo=4, hi=20, seed=3)
    b = W.uniform_vector(mat.n, seed=4)
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(mat.n)
    B = api.DeviceBatch(mat, sys_row)
    assert B.schedule(cfg)["teams"] >= 1
    B.set_config(cfg)
    iters, resid, status = B.solve(db, dx, _prm(precond=0))
    _check_against_oracle(mat, sys_row, b, dx.numpy(), iters, resid, status, precond=0)
    # an iteration limit of 12: the systems that need more stop with status 1,
    # 12 iterations and the oracle's iterate after 12 iterations
    iters, resid, status = B.solve(db, dx, _prm(max_iter=12))
    x = dx.numpy()
    limited = 0
    for k in range(B.nsys):
        sk = _system(mat, sys_row, k)
        xo, st, I, R = oracle.solve(sk, b[sys_row[k]:sys_row[k + 1]], precond=1, max_iter=12)
        xk = x[sys_row[k]:sys_row[k + 1]]
        if st == 0:
            assert status[k] == 0, k
        else:
            limited += 1
            assert status[k] == 1 and iters[k] == 12, (k, status[k], iters[k])
        assert np.linalg.norm(xk - xo) / np.linalg.norm(xo) <= 1e-6, k
    assert limited > 0
