"""GPU parity of the HIP path against the CPU oracle and monolis's golden vectors.

Every test calls libmonolis_hip.so through its C-ABI (ctypes) and compares with
the oracle on the same seeded inputs.  Bars (stated per test):
  SpMV: max |y_gpu - y_orc| <= 1e-12 * max |y_orc| (and exact for the integer KATs);
  dot:  relative difference <= 1e-12;
  solves: same convergence status, iteration counts within +-2 % (+-1 below 50),
          relative residual below the tolerance, ||x_gpu - x_orc|| / ||x_orc|| <= 1e-6.
"""
import numpy as np
import pytest

import oracle
from monolis_amd import api, workloads as W

from _helpers import chain_matrix, from_mesh_case, from_mesh_case_complex, kat, mtx_matrix, random_bcsr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    oracle.set_threads(1)      # sequential sums: the reference's single-thread order
    yield


def _oracle_iter_spread(m, b, **kw):
    """Iteration counts of the oracle under several summation orders (OpenMP
    thread counts).  Near finite termination the count itself depends on the
    reduction order; the GPU (another order) must fall inside this spread."""
    its = []
    for th in (1, 2, 3, 4, 5, 8):
        oracle.set_threads(th)
        its.append(int(oracle.solve(m, b, **kw)[2][4]))
    oracle.set_threads(1)
    return min(its), max(its)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _iters_close(a, b):
    return abs(a - b) <= max(1, int(0.02 * b))


# --- SpMV -------------------------------------------------------------------------
@pytest.mark.parametrize("case", kat()["matvec"], ids=lambda c: c["name"])
def test_matvec_reference_kat(case):
    m = from_mesh_case(case)
    y = api.monolis_matvec_product_R(api.Monolis(m), np.array(case["x"]))
    assert np.array_equal(y[:m.nndof], np.array(case["y"]))


@pytest.mark.parametrize("ndof", [1, 2, 3, 4, 5, 6, 7, 8])
def test_matvec_random_pattern(ndof):
    m = random_bcsr(203, ndof, 0.05, seed=100 + ndof, diag_missing=(5, 77))
    x = np.random.default_rng(ndof).uniform(-1, 1, m.npndof)
    y = api.monolis_matvec_product_R(api.Monolis(m), x)[:m.nndof]
    assert _rel(y, oracle.matvec(m, x)) <= 1e-12


def test_matvec_empty_rows_and_ragged_slices():
    # rows of length 0, 1 and long rows in one slice; n not a multiple of 64
    m = random_bcsr(130, 3, 0.0, seed=3, diag_missing=tuple(range(0, 130, 7)))
    rng = np.random.default_rng(0)
    rows = [sorted(set(rng.integers(0, 130, size=int(rng.integers(0, 40))).tolist())) for _ in range(130)]
    m.index = np.zeros(131, np.int32)
    m.index[1:] = np.cumsum([len(r) for r in rows])
    m.item = np.array([c + 1 for r in rows for c in r], np.int32)
    m.A = rng.uniform(-1, 1, len(m.item) * 9)
    m.n_dof_list = None
    m.ensure_dof_index()
    x = rng.uniform(-1, 1, 390)
    y = api.monolis_matvec_product_R(api.Monolis(m), x)[:390]
    assert _rel(y, oracle.matvec(m, x)) <= 1e-12


def test_matvec_stencil27_device_resident():
    m = W.stencil27_bcsr3(24, seed=5)
    x = np.random.default_rng(1).uniform(-1, 1, m.nndof)
    dm = api.DeviceMatrix(m)
    dx = api.DeviceBuffer.from_numpy(x)
    dy = api.DeviceBuffer(m.nndof)
    dm.matvec(dx, dy)
    assert _rel(dy.numpy(), oracle.matvec(m, x)) <= 1e-12
    # FP32 layout: values and vectors rounded to float, same product within FP32 error
    dx32 = api.DeviceBuffer.from_numpy(x.astype(np.float32))
    dy32 = api.DeviceBuffer(m.nndof, np.float32)
    dm.matvec(dx32, dy32)
    ref = oracle.matvec(m, x)
    assert _rel(dy32.numpy().astype(np.float64), ref) <= 1e-5


def test_matvec_matrix_staged_by_threaded_upload():
    """Host arrays of 64 MB and more are copied by several threads, each taking
    a 64 KB-aligned share (the last one the ragged tail).  The 40^3-node BCSR3
    value array is 118 MB: the device values read back must equal the host's
    bit for bit, and the product must match the oracle.  The same holds for a
    device vector filled from a 72 MB host array."""
    m = W.stencil27_bcsr3(40, seed=9)
    assert m.A.nbytes >= 64 << 20 and m.A.nbytes % (1 << 16) != 0
    dm = api.DeviceMatrix(m)
    assert np.array_equal(dm.values(), m.A)
    x = np.random.default_rng(4).uniform(-1, 1, m.nndof)
    dx = api.DeviceBuffer.from_numpy(x)
    dy = api.DeviceBuffer(m.nndof)
    dm.matvec(dx, dy)
    assert _rel(dy.numpy(), oracle.matvec(m, x)) <= 1e-12
    # a device vector filled from a 72 MB host array (monolis_hip_memcpy)
    v = np.random.default_rng(5).uniform(-1, 1, 9_000_001)
    assert np.array_equal(api.DeviceBuffer.from_numpy(v).numpy(), v)


@pytest.mark.parametrize("gen", ["stencil27", "poisson7"])
def test_matvec_every_kernel_variant_is_bit_identical(gen):
    """The autotuner may pick any variant: all of them must give the same bits
    (same per-row accumulation order, only load scheduling differs)."""
    from monolis_amd import _lib
    m = W.stencil27_bcsr3(13, seed=4) if gen == "stencil27" else W.poisson7(23)
    x = np.random.default_rng(5).uniform(-1, 1, m.nndof)
    dm = api.DeviceMatrix(m)
    dx = api.DeviceBuffer.from_numpy(x)
    dy = api.DeviceBuffer(m.nndof)
    ref = oracle.matvec(m, x)
    outs = []
    for v in range(_lib.load().monolis_hip_num_variants()):
        dm.set_variant(v)
        dm.matvec(dx, dy)
        outs.append(dy.numpy())
    assert _rel(outs[0], ref) <= 1e-12
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    tune = dm.autotune(fp32=True, reps=2)
    assert 0 <= tune["best64"] < len(outs) and 0 <= tune["best32"] < len(outs)
    dm.matvec(dx, dy)
    assert np.array_equal(dy.numpy(), outs[0])


def test_autotune_solver_keeps_solution():
    """monolis_hip_matrix_autotune_solver (no counterpart in monolis): per-variant
    per-iteration times of a fixed-length block-Jacobi CG, one valid variant kept per
    precision, and the solve after tuning agrees with the one before.  The variants'
    SpMV products are bit-identical, but the (p, q) partial sums come one per
    workgroup and the variants' grids differ, so the dots (and x) may differ in the
    last bits: the iteration count within the usual tolerance, x to 1e-8, and the
    refined solve to tol."""
    from monolis_amd import _lib
    m = W.stencil27_bcsr3(12, seed=6)
    b = oracle.matvec(m, W.uniform_vector(m.nndof, seed=7))
    dm = api.DeviceMatrix(m)
    db = api.DeviceBuffer.from_numpy(b)
    x0 = api.DeviceBuffer(m.npndof)
    p0 = api.Prm().quiet()
    dm.solve(db, x0, p0)
    nv = _lib.load().monolis_hip_num_variants()
    t64 = dm.autotune_solver(fp32=False, iters=4, reps=1)
    t32 = dm.autotune_solver(fp32=True, iters=4, reps=1)
    for t in (t64, t32):
        assert 0 <= t["best"] < nv and len(t["ms_per_iter"]) == nv
        assert all(v > 0 for v in t["ms_per_iter"]), t
        assert t["ms_per_iter"][t["best"]] == min(t["ms_per_iter"])
    x1 = api.DeviceBuffer(m.npndof)
    p1 = api.Prm().quiet()
    dm.solve(db, x1, p1)
    assert p1.Iarray[4] > 0 and _iters_close(int(p1.Iarray[4]), int(p0.Iarray[4]))
    assert np.linalg.norm(x1.numpy() - x0.numpy()) <= 1e-8 * np.linalg.norm(x0.numpy())
    xr = api.DeviceBuffer(m.npndof)
    pr = api.Prm().quiet()
    dm.solve_refined(db, xr, pr, inner_tol=1e-4)
    r = b - oracle.matvec(m, xr.numpy()[:m.nndof])
    assert np.linalg.norm(r) / np.linalg.norm(b) < 1e-8
    with pytest.raises(RuntimeError):
        dm.autotune_solver(iters=0)


def _with_far_zero_block(m, far):
    """Copy of m with one extra stored block (row 0, column `far`) holding zeros:
    the same operator, but the column span exceeds the 16-bit delta range."""
    from monolis_amd import matrix as mm
    b2 = m.ndof * m.ndof
    n = m.n
    rows = [list(m.item[m.index[i]:m.index[i + 1]]) for i in range(n)]
    vals = [m.A[m.index[i] * b2:m.index[i + 1] * b2] for i in range(n)]
    cols0 = rows[0] + [far + 1]
    order = np.argsort(cols0, kind="stable")
    v0 = np.concatenate([vals[0].reshape(-1, b2), np.zeros((1, b2))])[order].reshape(-1)
    rows[0] = [cols0[k] for k in order]
    vals[0] = v0
    index = np.zeros(n + 1, np.int32)
    index[1:] = np.cumsum([len(r) for r in rows])
    item = np.array([c for r in rows for c in r], np.int32)
    return mm.Matrix(n, n, m.ndof, index, item, np.concatenate(vals)).ensure_dof_index()


@pytest.mark.parametrize("gen", ["poisson7", "stencil27"])
def test_matvec_16bit_column_stream_matches_32bit(gen):
    """Matrices whose |col - row| fits in 16 bits read a 16-bit delta copy of
    the column ids; a far zero block forces the 32-bit stream on the same
    operator.  Both must give the same bits in every kernel variant, and the
    same CG iterations."""
    from monolis_amd import _lib
    m = W.poisson7(42) if gen == "poisson7" else W.stencil27_bcsr3(33, seed=2)  # n = 74,088 / 35,937 block rows
    wide = _with_far_zero_block(m, m.n - 1)
    assert m.n - 1 > 32767
    x = np.random.default_rng(7).uniform(-1, 1, m.nndof)
    dx, dy = api.DeviceBuffer.from_numpy(x), api.DeviceBuffer(m.nndof)
    outs = {}
    for name, mat in (("narrow", m), ("wide", wide)):
        dm = api.DeviceMatrix(mat)
        res = []
        for v in range(_lib.load().monolis_hip_num_variants()):
            dm.set_variant(v)
            dm.matvec(dx, dy)
            res.append(dy.numpy())
        outs[name] = res
    ref = oracle.matvec(m, x)
    assert _rel(outs["narrow"][0], ref) <= 1e-12
    for a, b in zip(outs["narrow"], outs["wide"]):
        assert np.array_equal(a, b)
    b = oracle.matvec(m, np.ones(m.nndof))
    its = []
    for mat in (m, wide):
        mo = api.Monolis(mat)
        mo.prm.quiet()
        xs = np.zeros(mat.npndof)
        api.monolis_solve_R(mo, b, xs)
        assert mo.prm.Iarray[5] == 0
        its.append(int(mo.prm.Iarray[4]))
    assert its[0] == its[1]


def test_matvec_values_update():
    m = W.stencil27_bcsr3(10, seed=1)
    dm = api.DeviceMatrix(m)
    x = np.random.default_rng(2).uniform(-1, 1, m.nndof)
    dx = api.DeviceBuffer.from_numpy(x)
    dy = api.DeviceBuffer(m.nndof)
    m2 = W.stencil27_bcsr3(10, seed=9)
    dm.set_values(m2.A)
    dm.matvec(dx, dy)
    assert _rel(dy.numpy(), oracle.matvec(m2, x)) <= 1e-12


# --- BLAS-1 ---------------------------------------------------------------------------
def test_inner_product_kat_and_random():
    k = kat()["blas"]["inner_product"]
    mon = api.Monolis(chain_matrix(2, 2, 0))
    assert api.monolis_inner_product_R(mon, 2, np.array(k["x"]), np.array(k["y"])) == k["sum"]
    rng = np.random.default_rng(3)
    x, y = rng.uniform(-1, 1, 1_000_003), rng.uniform(-1, 1, 1_000_003)
    s = api.inner_product_device(api.DeviceBuffer.from_numpy(x), api.DeviceBuffer.from_numpy(y))
    ref = oracle.inner_product(x, y)
    assert abs(s - ref) <= 1e-12 * abs(ref) + 1e-12 * np.sqrt(len(x))
    # deterministic: same bits on a second call
    s2 = api.inner_product_device(api.DeviceBuffer.from_numpy(x), api.DeviceBuffer.from_numpy(y))
    assert s == s2


# --- solves through the flat ABI -------------------------------------------------------
def _solve_flat(m, b, method=1, precond=1, tol=1e-8, max_iter=1000, x0=None, init_x=True, abort=False):
    mon = api.Monolis(m, api.Prm().quiet())
    api.monolis_set_method(mon, method)
    api.monolis_set_precond(mon, precond)
    api.monolis_set_tolerance(mon, tol)
    api.monolis_set_maxiter(mon, max_iter)
    api.monolis_set_init_x(mon, init_x)
    api.monolis_set_error_abort(mon, abort)
    x = np.zeros(m.npndof) if x0 is None else np.array(x0, np.float64)
    api.monolis_solve_R(mon, b, x)
    return x, mon


@pytest.mark.parametrize("method", [1, 5])
@pytest.mark.parametrize("precond", [0, 1])
def test_solve_mtx_dat(method, precond):
    case = kat()["mtx"]
    m = mtx_matrix(case)
    b = oracle.matvec(m, np.ones(10))
    x, mon = _solve_flat(m, b, method, precond, tol=case["tol"])
    assert api.monolis_get_error_tag(mon) == 0
    assert api.monolis_get_converge_iter(mon) >= 1
    assert api.monolis_get_converge_residual(mon) <= 1e-10
    assert np.allclose(x, 1.0, rtol=0, atol=1e-9)


@pytest.mark.parametrize("ndof", [1, 2, 3])
@pytest.mark.parametrize("precond", [0, 1])
@pytest.mark.parametrize("method", [1, 5])
def test_solve_chain_matches_oracle(ndof, precond, method):
    m = chain_matrix(10, ndof, seed=ndof)
    b = oracle.matvec(m, np.ones(10 * ndof))
    x, mon = _solve_flat(m, b, method, precond, tol=1e-10)
    xo, st, I, R = oracle.solve(m, b, method=method, precond=precond, tol=1e-10)
    assert api.monolis_get_error_tag(mon) == 0 and st == 0
    # exact: the oracle's count, or (where the oracle's own count moves with its
    # summation order, e.g. ndof 3 BiCGSTAB: 25-26) a count inside that spread
    lo, hi = _oracle_iter_spread(m, b, method=method, precond=precond, tol=1e-10)
    it = api.monolis_get_converge_iter(mon)
    assert it == int(I[4]) or lo <= it <= hi, (it, int(I[4]), lo, hi)
    assert np.allclose(x, 1.0, rtol=0, atol=1e-8)


def test_solve_c_reference_case():
    """The fixture's draw of wrapper_test/solver/monolis_solver_c_test.c:8-81 (the
    matrix test_c_abi.c solves): status 0 and the oracle's 14 CG iterations."""
    case = kat()["solve_c"]
    m = from_mesh_case(case)
    b = oracle.matvec(m, np.ones(20))
    x, mon = _solve_flat(m, b, tol=case["tol"])
    xo, st, I, R = oracle.solve(m, b, tol=case["tol"])
    assert st == 0 and api.monolis_get_error_tag(mon) == 0
    assert api.monolis_get_converge_iter(mon) == int(I[4]) == case["oracle_iter"]
    assert api.monolis_get_converge_residual(mon) < case["tol"]
    assert np.allclose(x, case["x_expected"], rtol=0, atol=1e-8)


def test_solve_c_complex_reference_case():
    """Complex twin (monolis_solver_c_test.c:83-158), COCG: status 0 and the oracle's
    35 iterations +-1.  The oracle's residual at 34 iterations is 1.0022e-10
    against the 1e-10 tolerance, so another summation order may stop one
    iteration earlier."""
    case = kat()["solve_c_complex"]
    m = from_mesh_case_complex(case)
    b = oracle.matvec_C(m, np.full(20, 1 + 1j))
    mon = api.Monolis(m, api.Prm().quiet())
    api.monolis_set_method(mon, api.MONOLIS_ITER_COCG)
    api.monolis_set_tolerance(mon, case["tol"])
    x = np.zeros(20, np.complex128)
    api.monolis_solve_C(mon, b, x)
    xo, st, I, R = oracle.solve_C(m, b, tol=case["tol"])
    assert st == 0 and api.monolis_get_error_tag(mon) == 0
    assert abs(api.monolis_get_converge_iter(mon) - int(I[4])) <= 1 and int(I[4]) == case["oracle_iter"]
    np.testing.assert_allclose(x, 1 + 1j, rtol=0, atol=1e-8)


@pytest.mark.parametrize("ndof", [1, 3])
def test_cg_row_without_diagonal_or_padding(ndof):
    """A row without a diagonal block that fills its slice's width (no padding
    slot, which would hold the row's own column): row 5 has three off-diagonal
    blocks in a slice of width 3.  Five CG iterations (max_iter 5, no abort)
    match the oracle's x; the block-Jacobi factor of that row follows monolis's
    rule for a missing diagonal block."""
    from monolis_amd import matrix as mm
    n = 64
    rows = []
    for i in range(n):
        if i == 5:
            rows.append([3, 4, 6])
        else:
            rows.append(sorted({max(i - 1, 0), i, min(i + 1, n - 1)}))
    index = np.zeros(n + 1, np.int32)
    index[1:] = np.cumsum([len(r) for r in rows])
    item = np.array([c + 1 for r in rows for c in r], np.int32)
    rng = np.random.default_rng(31)
    A = rng.uniform(-1, 1, size=len(item) * ndof * ndof)
    for i, r in enumerate(rows):                       # dominant diagonal blocks where present
        for k, c in enumerate(r):
            if c == i:
                blk = A[(index[i] + k) * ndof * ndof:(index[i] + k + 1) * ndof * ndof].reshape(ndof, ndof)
                blk += 8.0 * np.eye(ndof)
    m = mm.Matrix(n, n, ndof, index, item, A).ensure_dof_index()
    assert max(len(r) for r in rows) == len(rows[5]) == 3 and 5 not in rows[5]
    b = rng.uniform(-1, 1, m.nndof)
    x, mon = _solve_flat(m, b, method=1, precond=1, tol=1e-30, max_iter=5)
    Io, Ro = oracle.default_arrays(method=oracle.ITER_CG, precond=oracle.PREC_DIAG, tol=1e-30, max_iter=5)
    Io[16] = 0                                           # report non-convergence instead of stopping
    xo, st, I, R = oracle.solve(m, b, Iarray=Io, Rarray=Ro)
    assert api.monolis_get_converge_iter(mon) == int(I[4]) == 5
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)


@pytest.mark.parametrize("method", [1, 3, 5])
def test_solver_time_split(method):
    """Rarray time_spmv (5), time_dotp (6), time_prec (7) and the two
    communication slots (8, 9) (src/define/def_solver_prm.f90:151-163).
    With is_measurement set the three computation slots are positive and sum
    to no more than time_sol (3), the one-GPU communication slots are 0, and
    the solve itself is unchanged (same iterations, bit-identical x); without
    the timing flags every slot of 5..9 is 0."""
    m = W.stencil27_bcsr3(16, seed=3)
    b = oracle.matvec(m, W.uniform_vector(m.nndof, seed=4))
    x0, mon0 = _solve_flat(m, b, method=method)
    mon = api.Monolis(m, api.Prm().quiet())
    api.monolis_set_method(mon, method)
    api.monolis_set_error_abort(mon, False)
    api.monolis_set_performance_measurement(mon, True)
    x = np.zeros(m.npndof)
    api.monolis_solve_R(mon, b, x)
    R = mon.prm.Rarray
    assert api.monolis_get_error_tag(mon) == 0 and api.monolis_get_error_tag(mon0) == 0
    assert api.monolis_get_converge_iter(mon) == api.monolis_get_converge_iter(mon0)
    assert np.array_equal(x, x0)
    assert R[5] > 0 and R[6] > 0 and R[7] > 0, R[3:10]
    assert R[8] == 0 and R[9] == 0
    assert R[5] + R[6] + R[7] <= R[3], R[3:10]
    assert np.all(mon0.prm.Rarray[5:10] == 0), mon0.prm.Rarray[3:10]


def test_solve_poisson32_cg_diag_matches_oracle():
    """The CPU-runnable workload: 7-point Poisson 32^3, CG + DIAG, tol 1e-8."""
    m = W.poisson7(32)
    b = W.uniform_vector(m.nndof, seed=0)
    x, mon = _solve_flat(m, b)
    xo, st, I, R = oracle.solve(m, b)
    it_gpu, it_orc = api.monolis_get_converge_iter(mon), int(I[4])
    assert st == 0 and api.monolis_get_error_tag(mon) == 0
    assert _iters_close(it_gpu, it_orc), (it_gpu, it_orc)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-6
    true = np.linalg.norm(b - m.to_scipy() @ x) / np.linalg.norm(b)
    assert true < 1e-8 * 1.01


def test_solve_residual_replacement_beyond_200():
    m = W.poisson7(64)
    b = W.uniform_vector(m.nndof, seed=0)
    x, mon = _solve_flat(m, b, precond=0, tol=1e-10, max_iter=2000)
    xo, st, I, R = oracle.solve(m, b, precond=0, tol=1e-10, max_iter=2000)
    it_gpu = api.monolis_get_converge_iter(mon)
    assert it_gpu > 200 and _iters_close(it_gpu, int(I[4])), (it_gpu, int(I[4]))
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-6


@pytest.mark.parametrize("precond", [0, 1])
def test_bicgstab_convdiff_matches_oracle(precond):
    m = W.convdiff7(16, 0.5)
    xt = W.uniform_vector(m.nndof, seed=1)
    b = oracle.matvec(m, xt)
    x, mon = _solve_flat(m, b, method=5, precond=precond)
    xo, st, I, R = oracle.solve(m, b, method=5, precond=precond)
    assert st == 0 and api.monolis_get_error_tag(mon) == 0
    assert _iters_close(api.monolis_get_converge_iter(mon), int(I[4]))
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-6


def test_block_jacobi_bcsr3_matches_oracle():
    m = W.stencil27_bcsr3(16, seed=3)
    xt = W.uniform_vector(m.nndof, seed=4)
    b = oracle.matvec(m, xt)
    x, mon = _solve_flat(m, b)
    xo, st, I, R = oracle.solve(m, b)
    assert _iters_close(api.monolis_get_converge_iter(mon), int(I[4]))
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-6
    assert np.linalg.norm(x - xt) / np.linalg.norm(xt) <= 1e-6


def test_edge_cases_flat_abi():
    m = chain_matrix(10, 2, seed=9)
    # b = 0 -> x = 0, cur_iter untouched
    mon = api.Monolis(m, api.Prm().quiet())
    mon.prm.Iarray[4] = 77
    x = np.full(20, 3.0)
    api.monolis_solve_R(mon, np.zeros(20), x)
    assert np.all(x == 0.0) and mon.prm.Iarray[4] == 77 and api.monolis_get_error_tag(mon) == 0
    # warm start at the exact solution: ||r0|| = 0 also resets x (monolis rule)
    b = oracle.matvec(m, np.ones(20))
    x, mon = _solve_flat(m, b, x0=np.ones(20), init_x=False)
    assert np.all(x == 0.0)
    # max_iter = 1 without abort -> status 1, one iteration
    x, mon = _solve_flat(m, b, max_iter=1)
    assert api.monolis_get_error_tag(mon) == 1 and api.monolis_get_converge_iter(mon) == 1
    # unsupported preconditioner (SOR) without abort -> error status, x untouched
    x0 = np.full(20, 5.0)
    x, mon = _solve_flat(m, b, precond=2, x0=x0)
    assert api.monolis_get_error_tag(mon) < 0 and np.all(x == 5.0)
    # zero diagonal block entry: factorisation replaces the pivot by 1, stays finite
    mz = chain_matrix(10, 3, seed=2)
    mz.A[0] = 0.0
    bz = oracle.matvec(mz, np.ones(30))
    x, mon = _solve_flat(mz, bz, max_iter=50)
    xo, st, I, R = oracle.solve(mz, bz, max_iter=50, tol=1e-8)
    assert np.all(np.isfinite(x))
    assert _iters_close(api.monolis_get_converge_iter(mon), int(I[4]))


# --- device-resident API ---------------------------------------------------------------
def test_device_solve_matches_flat_and_history():
    m = W.stencil27_bcsr3(12, seed=8)
    b = oracle.matvec(m, W.uniform_vector(m.nndof, 3))
    dm = api.DeviceMatrix(m)
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(m.npndof)
    prm = api.Prm().quiet()
    hist = dm.solve(db, dx, prm, hist_len=200)
    it = int(prm.Iarray[4])
    xo, st, I, R = oracle.solve(m, b)
    assert it == int(I[4]) or _iters_close(it, int(I[4]))
    assert abs(hist[it] - prm.Rarray[2]) == 0.0 and hist[it] < 1e-8 <= hist[it - 1]
    assert np.linalg.norm(dx.numpy() - xo) / np.linalg.norm(xo) <= 1e-6
    # repeated solve on the same handle: identical bits (deterministic reductions)
    dx2 = api.DeviceBuffer(m.npndof)
    dm.solve(db, dx2, api.Prm().quiet())
    assert np.array_equal(dx.numpy(), dx2.numpy())


@pytest.mark.parametrize("scaled", [False, True])
def test_refined_solve_reaches_fp64_accuracy(scaled):
    m = W.stencil27_bcsr3(16, seed=11)
    if scaled:
        m = W.scale_symmetric(m, 1e3, seed=2)
    xt = W.uniform_vector(m.nndof, seed=12)
    b = oracle.matvec(m, xt)
    dm = api.DeviceMatrix(m)
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(m.npndof)
    prm = api.Prm().quiet()
    api_m = api.Monolis(m, prm)
    api.monolis_set_tolerance(api_m, 1e-10)
    st = dm.solve_refined(db, dx, prm, inner_tol=1e-4, max_outer=30)
    assert st["converged"] and st["resid"] < 1e-10
    x = dx.numpy()
    true = np.linalg.norm(b - oracle.matvec(m, x)) / np.linalg.norm(b)
    assert true < 1e-10
    xo, s0, I, R = oracle.solve(m, b, tol=1e-10, max_iter=5000)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-6


@pytest.mark.parametrize("n", [16, 32])
def test_refined_bicgstab_convdiff_matches_fp64_oracle(n):
    """FP32-inner BiCGSTAB (monolis method 5, BiCGSTAB.f90:18-192) with point-Jacobi
    scaling inside FP64 refinement, on the nonsymmetric 7-point convection-diffusion
    operator (cell Peclet 0.5) -- the refined BiCGSTAB workload at test size.
    Bar: status 0, true relative residual < 1e-10 in FP64, and the solution within
    1e-6 (relative 2-norm) of the oracle's FP64 BiCGSTAB solution at the same tolerance."""
    m = W.convdiff7(n, peclet=0.5)
    b = W.uniform_vector(m.nndof, seed=40 + n)
    dm = api.DeviceMatrix(m)
    db = api.DeviceBuffer.from_numpy(b)
    dx = api.DeviceBuffer(m.npndof)
    prm = api.Prm().quiet()
    prm.Iarray[1] = 5           # BiCGSTAB
    prm.Iarray[2] = 1           # DIAG
    prm.Rarray[1] = 1e-10
    st = dm.solve_refined(db, dx, prm, inner_tol=1e-4, max_outer=30)
    assert st["converged"] and st["resid"] < 1e-10 and int(prm.Iarray[5]) == 0
    assert st["outer"] >= 2          # the FP32 inner solve alone cannot reach 1e-10
    x = dx.numpy()[:m.nndof]
    true = np.linalg.norm(b - oracle.matvec(m, x)) / np.linalg.norm(b)
    assert true < 1e-10
    xo, s0, I, R = oracle.solve(m, b, method=5, precond=1, tol=1e-10, max_iter=5000)
    assert s0 == 0
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-6


def test_invalid_csr_is_rejected_without_launching():
    m = chain_matrix(10, 2, seed=1)
    bad = m.item.copy()
    bad[3] = 11                      # column outside 1..NP
    from monolis_amd import matrix as mm
    mb = mm.Matrix(m.n, m.np_, 2, m.index.copy(), bad, m.A.copy()).ensure_dof_index()
    with pytest.raises(api._lib.MonolisHipError, match="item outside"):
        api.DeviceMatrix(mb)
    idx = m.index.copy()
    idx[4] = idx[5] + 1              # decreasing offsets
    mb = mm.Matrix(m.n, m.np_, 2, idx, m.item.copy(), m.A.copy()).ensure_dof_index()
    with pytest.raises(api._lib.MonolisHipError, match="non-decreasing"):
        api.DeviceMatrix(mb)


# --- C and Fortran callers of the flat ABI ----------------------------------------------
@pytest.mark.parametrize("exe,token", [("test_c_abi", "C_ABI_OK"), ("test_mtx_dat", "FORTRAN_OK"),
                                       ("test_cocg_chain", "FORTRAN_COCG_OK"),
                                       ("test_fortran_api", "FORTRAN_API_OK")])
def test_compiled_callers(exe, token):
    """Programs compiled against include/monolis_hip.h (C), mod_monolis_hip
    (amdflang, ISO_C_BINDING) and mod_monolis (monolis's Fortran API over
    libmonolis_front.so) run the reference's own test cases on the GPU."""
    import os
    import subprocess
    from monolis_amd import build as mb
    path = os.path.join(mb.BINDIR, exe)
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    args = [path]
    if exe == "test_c_abi":   # expected CG / COCG iteration counts of the fixture's draws
        args += [str(kat()["solve_c"]["oracle_iter"]), str(kat()["solve_c_complex"]["oracle_iter"])]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert token in r.stdout


def test_monolis_c_api_program_on_gpu():
    """monolis_amd/c/test_front.c: a C program written against monolis's C API
    (include/monolis.h), linked against libmonolis_front.so only -- SpMV against
    a host loop, every real method x NONE / DIAG on the chain of
    wrapper_test/solver/monolis_solver_c_test.c (x = 1), COCG (x = 1 + i),
    variable blocks."""
    import os
    import subprocess
    from monolis_amd import build as mb
    path = os.path.join(mb.BINDIR, "test_front")
    if not os.path.exists(path):
        pytest.fail(f"{path} not built (run __graft_entry__.build())")
    r = subprocess.run([path, "gpu"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok (0 failed)" in r.stdout


# --- BiCGSTAB exact-zero dots: sorted-sum fallback and resume ---------------------------
def test_bicgstab_breakdown_stage_d_diagonal_system():
    """A = 2I with DIAG: s = 0 exactly, so (t, t) = 0 -> sorted sum -> omega = 0
    (BiCGSTAB.f90:132-145) and x = b/2 after one iteration."""
    from monolis_amd import matrix as mm
    n = 100
    m = mm.Matrix(n, n, 1, np.arange(n + 1, dtype=np.int32), np.arange(1, n + 1, dtype=np.int32),
                  np.full(n, 2.0)).ensure_dof_index()
    b = np.random.default_rng(0).uniform(-1, 1, n)
    x, mon = _solve_flat(m, b, method=5, precond=1)
    xo, st, I, R = oracle.solve(m, b, method=5, precond=1)
    assert st == 0 and api.monolis_get_error_tag(mon) == 0
    assert api.monolis_get_converge_iter(mon) == int(I[4]) == 1
    assert np.array_equal(x, b / 2) and np.array_equal(xo, b / 2)


def test_bicgstab_breakdown_stage_b_orthogonal_shadow():
    """(rt, v) = 0 exactly at the first iteration: both paths recompute it with the
    sorted sum, get 0 again and follow monolis into a non-converged solve."""
    from monolis_amd import matrix as mm
    m = mm.Matrix(2, 2, 1, np.array([0, 1, 2], np.int32), np.array([2, 1], np.int32),
                  np.array([1.0, 1.0])).ensure_dof_index()
    b = np.array([1.0, 0.0])
    x, mon = _solve_flat(m, b, method=5, precond=0, max_iter=5)
    xo, st, I, R = oracle.solve(m, b, method=5, precond=0, max_iter=5)
    assert st == 1 and api.monolis_get_error_tag(mon) == 1
    assert api.monolis_get_converge_iter(mon) == int(I[4])
    assert np.array_equal(np.isfinite(x), np.isfinite(xo))
