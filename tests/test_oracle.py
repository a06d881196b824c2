"""Pins the CPU oracle against monolis's own golden vectors, closed forms,
dense brute force and independent library implementations (numpy / scipy).

CPU only (no GPU marker).  The oracle is the checker for the HIP path, so
every function it exposes is pinned here before any GPU test relies on it.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from monolis_amd import workloads as W

from _helpers import chain_matrix, from_mesh_case, from_mesh_case_complex, kat, mtx_matrix, random_bcsr


@pytest.fixture(scope="module", autouse=True)
def one_thread():
    oracle.set_threads(1)
    yield


# --- SpMV ---------------------------------------------------------------------
@pytest.mark.parametrize("case", kat()["matvec"], ids=lambda c: c["name"])
def test_matvec_reference_kat(case):
    m = from_mesh_case(case)
    y = oracle.matvec(m, np.array(case["x"]))
    assert np.array_equal(y, np.array(case["y"]))   # integers: exact


@pytest.mark.parametrize("ndof", [1, 2, 3, 4, 5])
def test_matvec_dense_bruteforce(ndof):
    m = random_bcsr(23, ndof, 0.2, seed=ndof)
    x = np.random.default_rng(7).uniform(-1, 1, m.npndof)
    y = oracle.matvec(m, x)
    assert np.allclose(y, m.dense() @ x, rtol=1e-13, atol=1e-13)


def test_matvec_stencil_matches_scipy():
    m = W.stencil27_bcsr3(6, seed=2)
    x = np.random.default_rng(3).uniform(-1, 1, m.nndof)
    y = oracle.matvec(m, x)
    ref = m.to_scipy() @ x
    assert np.max(np.abs(y - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_matvec_catches_transposed_blocks():
    """A transposed block layout (a plausible mistake) is caught by the KAT check."""
    m = random_bcsr(9, 3, 0.3, seed=11)
    x = np.random.default_rng(1).uniform(-1, 1, m.npndof)
    good = oracle.matvec(m, x)
    mt = random_bcsr(9, 3, 0.3, seed=11)
    mt.A = mt.A.reshape(-1, 3, 3).transpose(0, 2, 1).reshape(-1).copy()
    assert not np.allclose(oracle.matvec(mt, x), good)


# --- BLAS-1 / convergence ------------------------------------------------------
def test_inner_product_closed_form():
    k = kat()["blas"]["inner_product"]
    assert oracle.inner_product(k["x"], k["y"]) == k["sum"]


def test_axpby_closed_form():
    k = kat()["blas"]["axpby"]
    assert np.array_equal(oracle.vec_axpby(k["alpha"], k["x"], k["beta"], k["y"]), np.array(k["z"]))


def test_check_converge_closed_form():
    k = kat()["blas"]["converge"]
    conv, it, resid = oracle.check_converge(np.array(k["R"]), k["B2"], k["iter"], tol=1e-8)
    assert (conv, it, resid) == (k["is_converge"], k["cur_iter"], k["cur_resid"])


def test_sorted_inner_product_bruteforce():
    rng = np.random.default_rng(5)
    x = rng.normal(size=101) * 10.0 ** rng.integers(-8, 8, size=101)
    y = rng.normal(size=101)
    prods = sorted((a * b for a, b in zip(x, y)), key=abs)
    s = 0.0
    for p in prods:
        s += p
    assert oracle.sorted_inner_product(x, y) == s


def test_sorted_inner_product_ties_keep_index_order():
    """Products of equal magnitude and opposite sign: the fallback adds them in
    index order (Python's sort is stable), so the sum depends on that order in
    the last bit -- the order monolis_qsort_R_2d would give is unpinned."""
    # 1e-16 first, then +1 and -1: (1e-16 + 1) - 1 = 0 but (1e-16 - 1) + 1 = 2^-53
    x = np.array([1.0, 0.5, -1.0, 1e-16, -0.5])
    y = np.ones_like(x)
    s = 0.0
    for p in sorted(x, key=abs):                  # stable: ties stay in index order
        s += p
    assert oracle.sorted_inner_product(x, y) == s
    s_rev = 0.0                                   # ties in reverse index order
    for p in sorted(x[::-1], key=abs):
        s_rev += p
    assert s_rev != s                             # the tie order matters for this data


# --- block-Jacobi ---------------------------------------------------------------
@pytest.mark.parametrize("ndof", [1, 2, 3, 4])
def test_diag_apply_inverts_diagonal_blocks(ndof):
    m = random_bcsr(17, ndof, 0.2, seed=20 + ndof, spd=True)
    D = oracle.diag_setup(m)
    Dense = m.dense()
    rng = np.random.default_rng(ndof)
    v = rng.uniform(-1, 1, m.nndof)
    # apply(D, A_ii v_i) == v_i for every block row
    blockdiag = np.zeros_like(v)
    for i in range(m.n):
        s = slice(i * ndof, (i + 1) * ndof)
        blockdiag[s] = Dense[s, s] @ v[s]
    assert np.allclose(oracle.diag_apply(m, D, blockdiag), v, rtol=1e-12, atol=1e-12)
    # and equals numpy.linalg.solve per block
    z = oracle.diag_apply(m, D, v)
    for i in range(m.n):
        s = slice(i * ndof, (i + 1) * ndof)
        assert np.allclose(z[s], np.linalg.solve(Dense[s, s], v[s]), rtol=1e-12, atol=1e-12)


def test_diag_zero_pivot_and_missing_diagonal():
    # zero pivot -> 1 (diag_33.f90:74, diag_nn.f90:58-60)
    m = random_bcsr(4, 3, 0.0, seed=1)
    m.A[:] = 0.0
    for i in range(4):
        m.A[9 * i:9 * i + 9] = np.eye(3).reshape(-1) * (i + 1)
    m.A[9 * 2] = 0.0        # block 2 entry (0,0) is zero
    D = oracle.diag_setup(m)
    z = oracle.diag_apply(m, D, np.ones(12))
    assert np.allclose(z[6:9], [1.0, 1.0 / 3.0, 1.0 / 3.0])
    # b = 3: a row without a diagonal block factors to the identity;
    # b != 3: it stays zero (diag_33 factors all rows, diag_nn only found ones)
    m3 = random_bcsr(6, 3, 0.5, seed=4, diag_missing=(2,))
    z3 = oracle.diag_apply(m3, oracle.diag_setup(m3), np.ones(18))
    assert np.allclose(z3[6:9], 1.0)
    m2 = random_bcsr(6, 2, 0.5, seed=4, diag_missing=(2,))
    z2 = oracle.diag_apply(m2, oracle.diag_setup(m2), np.ones(12))
    assert np.allclose(z2[4:6], 0.0)


def test_diag_catches_swapped_triangles():
    """Swapping L and U in the factors (a plausible mistake) breaks the inverse."""
    m = random_bcsr(8, 3, 0.3, seed=9, spd=True)
    D = oracle.diag_setup(m).reshape(-1, 3, 3)
    Dbad = D.transpose(0, 2, 1).copy()
    for k in range(3):
        Dbad[:, k, k] = D[:, k, k]
    v = np.random.default_rng(2).uniform(-1, 1, 24)
    good = oracle.diag_apply(m, D.reshape(-1), v)
    bad = oracle.diag_apply(m, Dbad.reshape(-1), v)
    assert not np.allclose(good, bad)


# --- solvers --------------------------------------------------------------------
@pytest.mark.parametrize("method", [oracle.ITER_CG, oracle.ITER_BICGSTAB])
@pytest.mark.parametrize("precond", [oracle.PREC_NONE, oracle.PREC_DIAG])
def test_solve_mtx_dat_x_is_one(method, precond):
    case = kat()["mtx"]
    m = mtx_matrix(case)
    b = oracle.matvec(m, np.ones(10))
    x, st, I, R = oracle.solve(m, b, method=method, precond=precond, tol=case["tol"])
    assert st == 0 and I[4] >= 1 and R[2] <= 1e-10
    assert np.allclose(x, case["x_expected"], rtol=0, atol=1e-9)


@pytest.mark.parametrize("ndof", [1, 2, 3])
@pytest.mark.parametrize("precond", [oracle.PREC_NONE, oracle.PREC_DIAG])
@pytest.mark.parametrize("method", [oracle.ITER_CG, oracle.ITER_BICGSTAB])
def test_solve_chain_x_is_one(ndof, precond, method):
    m = chain_matrix(10, ndof, seed=ndof)
    b = oracle.matvec(m, np.ones(10 * ndof))
    x, st, I, R = oracle.solve(m, b, method=method, precond=precond, tol=1e-10)
    assert st == 0
    assert np.allclose(x, 1.0, rtol=0, atol=1e-8)


@pytest.mark.parametrize("threads", [1, 2, 4])
def test_solve_c_test_case(threads):
    """The fixture's draw of the reference C test converges in the recorded
    number of CG iterations under every summation order tried (its residual
    passes 1e-10 at 7.6e-11, after 3.1e-10 one iteration earlier)."""
    case = kat()["solve_c"]
    m = from_mesh_case(case)
    b = oracle.matvec(m, np.ones(20))
    oracle.set_threads(threads)
    try:
        x, st, I, R = oracle.solve(m, b, tol=case["tol"])
    finally:
        oracle.set_threads(1)
    assert st == 0 and np.allclose(x, case["x_expected"], rtol=0, atol=1e-8)
    assert int(I[4]) == case["oracle_iter"] == 14


def test_solve_c_complex_test_case():
    """The complex twin (wrapper_test/solver/monolis_solver_c_test.c:83-158) drawn
    next from the same glibc stream: COCG reaches x = 1 + i in the recorded count."""
    case = kat()["solve_c_complex"]
    m = from_mesh_case_complex(case)
    b = oracle.matvec_C(m, np.full(20, 1 + 1j))
    x, st, I, R = oracle.solve_C(m, b, tol=case["tol"])
    assert st == 0 and int(I[4]) == case["oracle_iter"] == 35
    np.testing.assert_allclose(x, 1 + 1j, rtol=0, atol=1e-8)


def test_c_caller_draws_the_fixture_values():
    """monolis_amd/c/test_c_abi.c draws the chain values from srand(1) before its
    first library call; "draws" prints them without calling the library."""
    import os
    import subprocess
    from monolis_amd import build as mb
    path = os.path.join(mb.BINDIR, "test_c_abi")
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    out = subprocess.run([path, "draws"], capture_output=True, text=True, timeout=60, check=True).stdout
    vals = [float(v) for v in out.splitlines()[-1].split()]
    assert len(vals) == 112
    for off, case, col in ((0, kat()["solve_c"]["set"], 4), (56, kat()["solve_c_complex"]["set_complex"], 4)):
        diag = [s[col] - 5000.0 for s in case[:20]]
        offd = [s[col] for s in case[20::2]]
        assert vals[off:off + 20] == diag and vals[off + 20:off + 56] == offd


def _jacobi(m):
    d = 1.0 / m.to_scipy().diagonal()
    return spla.LinearOperator((m.n, m.n), matvec=lambda v: d * v)


@pytest.mark.parametrize("precond", [oracle.PREC_NONE, oracle.PREC_DIAG])
def test_cg_iterations_match_scipy(precond):
    """Independent CG (scipy): same iteration count below the 200-iteration
    residual replacement (x0 = 0 makes ||r0|| = ||b||, so the tests agree)."""
    m = W.poisson7(12)
    b = W.uniform_vector(m.nndof, seed=3)
    x, st, I, R = oracle.solve(m, b, precond=precond, tol=1e-8)
    assert st == 0 and I[4] < 200
    count = [0]
    M = _jacobi(m) if precond == oracle.PREC_DIAG else None
    xs, info = spla.cg(m.to_scipy(), b, rtol=1e-8, atol=0.0, M=M, callback=lambda _: count.__setitem__(0, count[0] + 1))
    assert info == 0
    assert abs(count[0] - int(I[4])) <= 1
    assert np.allclose(x, xs, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("precond", [oracle.PREC_NONE, oracle.PREC_DIAG])
def test_bicgstab_iterations_match_scipy(precond):
    m = W.convdiff7(10, 0.5)
    b = W.uniform_vector(m.nndof, seed=4)
    x, st, I, R = oracle.solve(m, b, method=oracle.ITER_BICGSTAB, precond=precond, tol=1e-8)
    assert st == 0
    count = [0]
    M = _jacobi(m) if precond == oracle.PREC_DIAG else None
    xs, info = spla.bicgstab(m.to_scipy(), b, rtol=1e-8, atol=0.0, M=M,
                             callback=lambda _: count.__setitem__(0, count[0] + 1))
    assert info == 0
    assert abs(count[0] - int(I[4])) <= 1
    r = b - m.to_scipy() @ x
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-8


def test_cg_block_jacobi_residual_history_vs_textbook():
    """Block-Jacobi PCG (BCSR3): residual history equals an independent dense
    PCG built on numpy.linalg.solve for the diagonal blocks."""
    m = W.stencil27_bcsr3(5, seed=1)
    A = m.dense()
    xt = W.uniform_vector(m.nndof, 2)
    b = A @ xt
    x, st, I, R = oracle.solve(m, b, tol=1e-10)
    assert st == 0
    blocks = [A[3 * i:3 * i + 3, 3 * i:3 * i + 3] for i in range(m.n)]

    def prec(r):
        return np.concatenate([np.linalg.solve(blocks[i], r[3 * i:3 * i + 3]) for i in range(m.n)])
    xx = np.zeros_like(b)
    r = b.copy()
    rr0 = r @ r
    p = None
    rho1 = None
    it = 0
    while True:
        it += 1
        z = prec(r)
        rho = r @ z
        p = z.copy() if p is None else z + (rho / rho1) * p
        q = A @ p
        a = rho / (p @ q)
        xx += a * p
        r -= a * q
        if np.sqrt((r @ r) / rr0) < 1e-10:
            break
        rho1 = rho
    assert it == int(I[4])
    assert np.allclose(x, xx, rtol=1e-9, atol=1e-12)
    assert np.allclose(x, xt, rtol=1e-7, atol=1e-8)


def test_residual_replacement_iteration_200():
    """Beyond 200 iterations the true residual replaces the recurrence: the
    oracle still converges and the final cur_resid is the true relative residual."""
    m = W.poisson7(64)
    b = W.uniform_vector(m.nndof, seed=0)
    x, st, I, R = oracle.solve(m, b, precond=oracle.PREC_NONE, tol=1e-10, max_iter=2000)
    assert st == 0 and I[4] > 200
    true = np.linalg.norm(b - m.to_scipy() @ x) / np.linalg.norm(b)
    assert true < 2e-10


def test_edge_cases():
    m = chain_matrix(10, 2, seed=9)
    # b = 0 -> x = 0, converged, cur_iter untouched (mat_converge.f90:40-48)
    I, R = oracle.default_arrays()
    I[4] = 77
    x, st, I, R = oracle.solve(m, np.zeros(20), x0=np.full(20, 3.0), Iarray=I, Rarray=R)
    assert st == 0 and np.all(x == 0.0) and I[4] == 77
    # warm start at the exact solution: ||r0|| = 0 also resets X to 0 (same rule)
    b = oracle.matvec(m, np.ones(20))
    I, R = oracle.default_arrays()
    I[10] = 0
    x, st, I, R = oracle.solve(m, b, x0=np.ones(20), Iarray=I, Rarray=R)
    assert st == 0 and np.all(x == 0.0)
    # max_iter = 1 without abort -> not converged after one iteration
    I, R = oracle.default_arrays(max_iter=1)
    I[16] = 0
    x, st, I, R = oracle.solve(m, b, Iarray=I, Rarray=R)
    assert st == 1 and I[4] == 1 and R[2] > 1e-8
