"""Device assembly on the GPU against the oracle (monolis_oracle_asm.c): the
pattern of a simple mesh, batched element matrices (bit-for-bit equal to one
add-matrix call per element in order), Dirichlet conditions (bit-for-bit for a
sorted list), value read-back, and an assemble -> condition -> solve pipeline."""
import ctypes

import numpy as np
import pytest

import oracle
from monolis_amd import _lib, api
from monolis_amd import matrix as mm

from _helpers import element_matrices, hex_mesh

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _init(hip):
    oracle.set_threads(1)
    yield


def _oracle_mat(n_node, b, elems):
    index, item = oracle.pattern_by_simple_mesh(n_node, elems)
    return mm.Matrix(n_node, n_node, b, index, item, np.zeros(len(item) * b * b)).ensure_dof_index()


def test_pattern_reference_unit_test():
    """src_test/matrix/spmat_nzpattern_test.f90:38-73 built on the GPU."""
    dm = api.device_matrix_by_simple_mesh(4, 2, [[0, 1], [1, 2], [2, 3]])
    assert dm.mat.index.tolist() == [0, 2, 5, 8, 10]
    assert dm.mat.item.tolist() == [1, 2, 1, 2, 3, 2, 3, 4, 3, 4]
    assert not dm.values().any()


@pytest.mark.parametrize("case", ["hex10", "tets", "isolated_nodes"])
def test_pattern_matches_oracle(case):
    rng = np.random.default_rng(4)
    if case == "hex10":
        n_node, elems = hex_mesh(10, 9, 8)
    elif case == "tets":
        n_node = 3000
        elems = np.array([rng.choice(n_node, 4, replace=False) for _ in range(5000)], np.int32)
        elems[7, 3] = elems[7, 0]                       # degenerate element
    else:
        n_node = 500                                     # nodes 400..499 in no element
        elems = np.array([rng.choice(400, 3, replace=False) for _ in range(300)], np.int32)
    dm = api.device_matrix_by_simple_mesh(n_node, 3, elems)
    index, item = oracle.pattern_by_simple_mesh(n_node, elems)
    assert np.array_equal(dm.mat.index, index) and np.array_equal(dm.mat.item, item)


def test_pattern_rejects_bad_node_ids():
    h = ctypes.c_void_p()
    e = np.array([[0, 5]], np.int32)
    rc = _lib.load().monolis_hip_matrix_create_by_simple_mesh(ctypes.byref(h), 4, 2, 1, 1, api._p(e))
    assert rc == -2 and b"outside" in _lib.load().monolis_hip_last_error()


@pytest.mark.parametrize("b", [1, 2, 3])
def test_element_assembly_bit_identical(b):
    n_node, elems = hex_mesh(6, 5, 4)
    Ke = element_matrices(len(elems), 8 * b, seed=b, spd=False)
    dm = api.device_matrix_by_simple_mesh(n_node, b, elems)
    # two calls (elements split in two) equal one call per element in order
    half = len(elems) // 2
    dm.add_elements(elems[:half], Ke[:half])
    dm.add_elements(elems[half:], Ke[half:])
    mo = _oracle_mat(n_node, b, elems)
    oracle.add_elements(mo, elems, Ke)
    assert np.array_equal(dm.values(), mo.A)
    # the assembled device matrix multiplies as the oracle's
    x = np.random.default_rng(1).uniform(-1, 1, n_node * b)
    dx, dy = api.DeviceBuffer.from_numpy(x), api.DeviceBuffer(n_node * b)
    dm.matvec(dx, dy)
    ref = oracle.matvec(mo, x)
    assert np.max(np.abs(dy.numpy() - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_element_assembly_large_crosses_chunks_and_device_inputs(monkeypatch):
    """35,937 hexahedra (2.3 M node pairs, 20.7 M values) in chunks of 2^20
    node pairs; connectivity and element matrices passed as device buffers."""
    monkeypatch.setenv("MONOLIS_HIP_ASSEMBLY_CHUNK_PAIRS", str(1 << 20))
    b = 3
    n_node, elems = hex_mesh(33)
    Ke = element_matrices(len(elems), 8 * b, seed=11, spd=True)
    dm = api.device_matrix_by_simple_mesh(n_node, b, elems)
    de, dk = api.DeviceBuffer.from_numpy(elems), api.DeviceBuffer.from_numpy(Ke.reshape(-1))
    rc = _lib.load().monolis_hip_matrix_add_elements_R(ctypes.c_void_p(dm.h), len(elems), 8, ctypes.c_void_p(de.ptr),
                                                       ctypes.c_void_p(dk.ptr))
    assert rc == 0, _lib.load().monolis_hip_last_error()
    mo = _oracle_mat(n_node, b, elems)
    oracle.add_elements(mo, elems, Ke)
    assert np.array_equal(dm.values(), mo.A)


def test_element_assembly_rejects_block_outside_pattern():
    dm = api.device_matrix_by_simple_mesh(4, 1, [[0, 1], [1, 2], [2, 3]])
    with pytest.raises(_lib.MonolisHipError, match="not in the nonzero pattern"):
        dm.add_elements(np.array([[0, 3]], np.int32), np.ones((1, 2, 2)))
    with pytest.raises(_lib.MonolisHipError, match="outside"):
        dm.add_elements(np.array([[0, 9]], np.int32), np.ones((1, 2, 2)))


def test_dirichlet_reference_unit_test():
    """src_test/matrix/spmat_handler_test.f90:119-132 on the device matrix."""
    dm = api.device_matrix_by_simple_mesh(4, 2, [[0, 1], [1, 2], [2, 3]])
    dm.set_values(np.full(10 * 4, 2.0))
    B = np.zeros(8)
    dm.set_dirichlet(B, [1], [0], [1.0])
    assert B.tolist() == [-2.0, -2.0, 1.0, -2.0, -2.0, -2.0, 0.0, 0.0]
    mo = _oracle_mat(4, 2, [[0, 1], [1, 2], [2, 3]])
    mo.A[:] = 2.0
    oracle.dirichlet(mo, np.zeros(8), [1], [0], [1.0])
    assert np.array_equal(dm.values(), mo.A)


@pytest.mark.parametrize("b", [1, 3])
def test_dirichlet_list_bit_identical_to_sequential(b):
    n_node, elems = hex_mesh(8)
    Ke = element_matrices(len(elems), 8 * b, seed=5)
    dm = api.device_matrix_by_simple_mesh(n_node, b, elems)
    dm.add_elements(elems, Ke)
    mo = _oracle_mat(n_node, b, elems)
    oracle.add_elements(mo, elems, Ke)
    rng = np.random.default_rng(9)
    f = rng.uniform(-1, 1, n_node * b)
    nodes = np.sort(rng.choice(n_node, 60, replace=False)).astype(np.int32)
    nodes = np.repeat(nodes, b)                      # every DOF of those nodes ...
    dofs = np.tile(np.arange(b, dtype=np.int32), 60)
    keep = rng.random(len(nodes)) < 0.7              # ... or most of them, sorted by (node, dof)
    nodes, dofs = nodes[keep], dofs[keep]
    val = rng.uniform(-1, 1, len(nodes))
    Bg = f.copy()
    dm.set_dirichlet(Bg, nodes, dofs, val)
    Bo = f.copy()
    oracle.dirichlet(mo, Bo, nodes, dofs, val)
    assert np.array_equal(dm.values(), mo.A)
    assert np.array_equal(Bg, Bo)


def test_dirichlet_errors():
    dm = api.device_matrix_by_simple_mesh(4, 2, [[0, 1], [1, 2], [2, 3]])
    B = np.zeros(8)
    with pytest.raises(_lib.MonolisHipError, match="twice"):
        dm.set_dirichlet(B, [1, 1], [0, 0], [1.0, 2.0])
    with pytest.raises(_lib.MonolisHipError, match="outside"):
        dm.set_dirichlet(B, [4], [0], [1.0])
    with pytest.raises(_lib.MonolisHipError, match="outside"):
        dm.set_dirichlet(B, [0], [2], [1.0])
    # a row without its diagonal block (user BCSR): monolis stops, here an error
    m = mm.Matrix(2, 2, 1, np.array([0, 1, 2], np.int32), np.array([2, 1], np.int32),
                  np.array([1.0, 1.0])).ensure_dof_index()
    dm2 = api.DeviceMatrix(m)
    with pytest.raises(_lib.MonolisHipError, match="diagonal block"):
        dm2.set_dirichlet(np.zeros(2), [0], [0], [1.0])


def test_assemble_condition_solve_pipeline():
    """Hexahedral mesh, 3 DOF per node: pattern, element matrices and Dirichlet
    conditions on the GPU, then the device CG; against the same pipeline in
    the oracle."""
    b = 3
    n_node, elems = hex_mesh(12)
    Ke = element_matrices(len(elems), 8 * b, seed=21)
    dm = api.device_matrix_by_simple_mesh(n_node, b, elems)
    dm.add_elements(elems, Ke)
    bottom = np.arange(13 * 13, dtype=np.int32)          # z = 0 face, all DOFs fixed to 0
    nodes, dofs = np.repeat(bottom, b), np.tile(np.arange(b, dtype=np.int32), len(bottom))
    f = np.random.default_rng(2).uniform(-1, 1, n_node * b)
    Bg = f.copy()
    dm.set_dirichlet(Bg, nodes, dofs, np.zeros(len(nodes)))
    db, dx = api.DeviceBuffer.from_numpy(Bg), api.DeviceBuffer(n_node * b)
    prm = api.Prm().quiet()
    prm.Rarray[1] = 1e-10
    dm.solve(db, dx, prm)
    mo = _oracle_mat(n_node, b, elems)
    oracle.add_elements(mo, elems, Ke)
    Bo = f.copy()
    oracle.dirichlet(mo, Bo, nodes, dofs, np.zeros(len(nodes)))
    xo, st, I, R = oracle.solve(mo, Bo, tol=1e-10)
    assert prm.Iarray[5] == 0 and st == 0
    assert abs(int(prm.Iarray[4]) - int(I[4])) <= max(1, int(0.02 * I[4]))
    x = dx.numpy()
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert not x[bottom[0] * b:(bottom[-1] + 1) * b].any()


# --- complex matrices (monolis_*_C) ------------------------------------------------------
def _complex_elements(n_elem, nzm, seed):
    """Complex symmetric element matrices: SPD real part + i * 0.3 * SPD."""
    return element_matrices(n_elem, nzm, seed=seed) + 0.3j * element_matrices(n_elem, nzm, seed=seed + 100)


@pytest.mark.parametrize("b", [1, 2, 3])
def test_complex_element_assembly_bit_identical(b):
    """monolis_add_matrix_to_sparse_matrix_main_C (spmat_handler_util.f90:501-600) adds
    complex values in the real routine's order, so the real and imaginary parts are
    each the real assembly of Ke.real and Ke.imag (the oracle's), bit for bit."""
    n_node, elems = hex_mesh(5, 4, 4)
    Ke = element_matrices(len(elems), 8 * b, seed=b, spd=False) + 1j * element_matrices(len(elems), 8 * b,
                                                                                          seed=b + 50, spd=False)
    dm = api.device_matrix_by_simple_mesh_C(n_node, b, elems)
    index, item = oracle.pattern_by_simple_mesh(n_node, elems)
    assert np.array_equal(dm.mat.index, index) and np.array_equal(dm.mat.item, item)
    half = len(elems) // 2
    dm.add_elements(elems[:half], Ke[:half])
    dm.add_elements(elems[half:], Ke[half:])
    mr, mi = _oracle_mat(n_node, b, elems), _oracle_mat(n_node, b, elems)
    oracle.add_elements(mr, elems, np.ascontiguousarray(Ke.real))
    oracle.add_elements(mi, elems, np.ascontiguousarray(Ke.imag))
    A = dm.values()
    assert np.array_equal(A.real, mr.A) and np.array_equal(A.imag, mi.A)
    # the assembled device matrix multiplies as the oracle's complex SpMV
    x = np.random.default_rng(1).uniform(-1, 1, n_node * b) + 1j * np.random.default_rng(2).uniform(-1, 1, n_node * b)
    mo = mm.Matrix(n_node, n_node, b, index, item, A.copy()).ensure_dof_index()
    dx, dy = api.DeviceBuffer.from_numpy(x), api.DeviceBuffer(n_node * b, np.complex128)
    dm.matvec(dx, dy)
    ref = oracle.matvec_C(mo, x)
    assert np.max(np.abs(dy.numpy() - ref)) <= 1e-12 * np.max(np.abs(ref))
    with pytest.raises(_lib.MonolisHipError, match="not in the nonzero pattern"):
        dm.add_elements(np.array([[0, n_node - 1]], np.int32), np.ones((1, 2 * b, 2 * b), np.complex128))


def test_complex_dirichlet_reference_unit_test():
    """src_test/matrix/spmat_handler_test.f90:236-248 on the device: A = (2, 2), B = 0,
    condition on node 2 (1-based), DOF 1, value (1, 0)."""
    dm = api.device_matrix_by_simple_mesh_C(4, 2, [[0, 1], [1, 2], [2, 3]])
    dm.set_values(np.full(10 * 4, 2.0 + 2.0j))
    B = np.zeros(8, np.complex128)
    dm.set_dirichlet(B, [1], [0], [1.0 + 0.0j])
    assert B.tolist() == [-2 - 2j, -2 - 2j, 1 + 0j, -2 - 2j, -2 - 2j, -2 - 2j, 0j, 0j]
    A = np.full(10 * 4, 2.0 + 2.0j)
    oracle.dirichlet_C(dm.mat.index, dm.mat.item, A, 4, 2, np.zeros(8, np.complex128), [1], [0], [1.0 + 0.0j])
    assert np.array_equal(dm.values(), A)


@pytest.mark.parametrize("b", [1, 3])
def test_complex_dirichlet_list_bit_identical_to_sequential(b):
    n_node, elems = hex_mesh(7)
    Ke = _complex_elements(len(elems), 8 * b, seed=5)
    dm = api.device_matrix_by_simple_mesh_C(n_node, b, elems)
    dm.add_elements(elems, Ke)
    A = dm.values().copy()
    rng = np.random.default_rng(9)
    f = rng.uniform(-1, 1, n_node * b) + 1j * rng.uniform(-1, 1, n_node * b)
    nodes = np.sort(rng.choice(n_node, 40, replace=False)).astype(np.int32)
    nodes = np.repeat(nodes, b)
    dofs = np.tile(np.arange(b, dtype=np.int32), 40)
    keep = rng.random(len(nodes)) < 0.7
    nodes, dofs = nodes[keep], dofs[keep]
    val = rng.uniform(-1, 1, len(nodes)) + 1j * rng.uniform(-1, 1, len(nodes))
    Bg = f.copy()
    dm.set_dirichlet(Bg, nodes, dofs, val)
    Bo = f.copy()
    oracle.dirichlet_C(dm.mat.index, dm.mat.item, A, n_node, b, Bo, nodes, dofs, val)
    assert np.array_equal(dm.values(), A)
    assert np.array_equal(Bg, Bo)
    with pytest.raises(_lib.MonolisHipError, match="twice"):
        dm.set_dirichlet(np.zeros(n_node * b, np.complex128), [0, 0], [0, 0], [1.0, 2.0])


def test_complex_assemble_condition_cocg_pipeline():
    """Complex symmetric element matrices, Dirichlet conditions and COCG on the device,
    against the same pipeline through the oracle (assembly restated as two real
    assemblies, complex conditions, the oracle's COCG)."""
    b = 3
    n_node, elems = hex_mesh(8)
    Ke = _complex_elements(len(elems), 8 * b, seed=21)
    dm = api.device_matrix_by_simple_mesh_C(n_node, b, elems)
    dm.add_elements(elems, Ke)
    bottom = np.arange(9 * 9, dtype=np.int32)
    nodes, dofs = np.repeat(bottom, b), np.tile(np.arange(b, dtype=np.int32), len(bottom))
    rng = np.random.default_rng(2)
    f = rng.uniform(-1, 1, n_node * b) + 1j * rng.uniform(-1, 1, n_node * b)
    A0 = dm.values().copy()
    Bg = f.copy()
    dm.set_dirichlet(Bg, nodes, dofs, np.zeros(len(nodes), np.complex128))
    db, dx = api.DeviceBuffer.from_numpy(Bg), api.DeviceBuffer(n_node * b, np.complex128)
    prm = api.Prm().quiet()
    prm.Iarray[1], prm.Iarray[2], prm.Iarray[3] = api.MONOLIS_ITER_COCG, 1, 2000   # COCG + block-Jacobi
    prm.Rarray[1] = 1e-10
    dm.solve(db, dx, prm)
    Bo = f.copy()
    oracle.dirichlet_C(dm.mat.index, dm.mat.item, A0, n_node, b, Bo, nodes, dofs, np.zeros(len(nodes), np.complex128))
    mo = mm.Matrix(n_node, n_node, b, dm.mat.index, dm.mat.item, A0).ensure_dof_index()
    xo, st, I, R = oracle.solve_C(mo, Bo, tol=1e-10, precond=1, max_iter=2000)
    assert prm.Iarray[5] == 0 and st == 0
    assert abs(int(prm.Iarray[4]) - int(I[4])) <= max(1, int(0.02 * I[4]))
    x = dx.numpy()
    assert np.linalg.norm(x - xo) <= 1e-6 * np.linalg.norm(xo)
    assert not x[bottom[0] * b:(bottom[-1] + 1) * b].any()


# --- variable block sizes (monolis n_dof = -1) -------------------------------------------
def _v_elements(elems, lst, seed, cplx=False, spd=False):
    """Element matrices of a variable-block mesh: element e is (sum of its nodes'
    block sizes)^2, row-major; returned as a list and as the concatenation."""
    rng = np.random.default_rng(seed)
    out = []
    for e in elems:
        S = int(lst[e].sum())
        R = rng.uniform(-1, 1, (S, S))
        K = R @ R.T / S + 0.05 * np.eye(S) if spd else R
        if cplx:
            K = K + 0.3j * (R.T @ R / S if spd else rng.uniform(-1, 1, (S, S)))
        out.append(K)
    return out, np.concatenate([k.reshape(-1) for k in out])


def _v_dense_sum(n_dofs, dofi, lst, elems, Ks, dtype):
    D = np.zeros((n_dofs, n_dofs), dtype)
    for e, K in zip(elems, Ks):  # in element order, entry by entry (np.add.at is unbuffered)
        d = np.concatenate([np.arange(dofi[v], dofi[v] + lst[v]) for v in e])
        np.add.at(D, (np.repeat(d, len(d)), np.tile(d, len(d))), K.reshape(-1))
    return D


@pytest.mark.parametrize("cplx", [False, True])
def test_vblock_element_assembly_matches_sequential_dense_sum(cplx):
    """monolis_add_matrix_to_sparse_matrix_R/_C on a variable-block matrix
    (spmat_handler_util.f90:398-600 through n_dof_index / n_dof_index2): the device
    values equal the dense element-by-element sums, bit for bit."""
    n_node, elems = hex_mesh(4, 3, 3)
    lst = np.random.default_rng(4).integers(1, 4, n_node).astype(np.int32)
    dm = api.device_matrix_by_simple_mesh_V(n_node, lst, elems, complex_values=cplx)
    index = np.zeros(n_node + 1, np.int32)
    item = np.zeros(len(dm.mat.item), np.int32)
    get = _lib.load().monolis_hip_matrix_get_pattern_C if cplx else _lib.load().monolis_hip_matrix_get_pattern
    assert get(ctypes.c_void_p(dm.h), index.ctypes.data_as(ctypes.c_void_p), item.ctypes.data_as(ctypes.c_void_p)) == 0
    assert np.array_equal(index, dm.mat.index) and np.array_equal(item, dm.mat.item)
    Ks, flat = _v_elements(elems, lst, seed=3, cplx=cplx)
    half = len(elems) // 2
    cut = sum(k.size for k in Ks[:half])
    dm.add_elements(elems[:half], flat[:cut])
    dm.add_elements(elems[half:], flat[cut:])
    dofi = dm.mat.n_dof_index
    D = _v_dense_sum(int(dofi[-1]), dofi, lst, elems, Ks, np.complex128 if cplx else np.float64)
    dm.mat.A = dm.values()
    assert np.array_equal(dm.mat.dense(), D)
    # the expanded device matrix multiplies as the dense sum
    rng = np.random.default_rng(6)
    x = rng.uniform(-1, 1, D.shape[0]) + (1j * rng.uniform(-1, 1, D.shape[0]) if cplx else 0)
    dt = np.complex128 if cplx else np.float64
    dx, dy = api.DeviceBuffer.from_numpy(x.astype(dt)), api.DeviceBuffer(D.shape[0], dt)
    dm.matvec(dx, dy)
    ref = D @ x
    assert np.max(np.abs(dy.numpy() - ref)) <= 1e-12 * np.max(np.abs(ref))


def test_vblock_uniform_sizes_equal_fixed_block_assembly_and_conditions():
    """With every block size b the variable-block layout is the fixed-block one:
    element sums and Dirichlet conditions bit-identical to the fixed-block device path."""
    b = 2
    n_node, elems = hex_mesh(5, 4, 3)
    Ke = element_matrices(len(elems), 8 * b, seed=8, spd=False)
    dv = api.device_matrix_by_simple_mesh_V(n_node, np.full(n_node, b, np.int32), elems)
    dfx = api.device_matrix_by_simple_mesh(n_node, b, elems)
    dv.add_elements(elems, Ke.reshape(-1))
    dfx.add_elements(elems, Ke)
    assert np.array_equal(dv.values(), dfx.values())
    rng = np.random.default_rng(1)
    f = rng.uniform(-1, 1, n_node * b)
    nodes = np.sort(rng.choice(n_node, 20, replace=False)).astype(np.int32)
    dofs = (nodes % b).astype(np.int32)
    val = rng.uniform(-1, 1, len(nodes))
    Bv, Bf = f.copy(), f.copy()
    dv.set_dirichlet(Bv, nodes, dofs, val)
    dfx.set_dirichlet(Bf, nodes, dofs, val)
    assert np.array_equal(dv.values(), dfx.values()) and np.array_equal(Bv, Bf)


def test_vblock_dirichlet_and_cg_solve_are_the_reduced_solve():
    """Variable blocks: conditions on a list of (node, DOF), then the device CG (with
    diag_V) on the conditioned system; its solution is that of the reduced system
    (dense brute force), the condition DOFs hold their values."""
    n_node, elems = hex_mesh(6)
    lst = np.random.default_rng(12).integers(1, 4, n_node).astype(np.int32)
    dm = api.device_matrix_by_simple_mesh_V(n_node, lst, elems)
    Ks, flat = _v_elements(elems, lst, seed=13, spd=True)
    dm.add_elements(elems, flat)
    dofi = dm.mat.n_dof_index
    m = int(dofi[-1])
    A0 = _v_dense_sum(m, dofi, lst, elems, Ks, np.float64)
    nodes = np.arange(0, n_node, 3, dtype=np.int32)
    dofs = np.array([(v // 3) % lst[v] for v in nodes], np.int32)
    rng = np.random.default_rng(14)
    val = rng.uniform(-1, 1, len(nodes))
    f = rng.uniform(-1, 1, m)
    Bg = f.copy()
    dm.set_dirichlet(Bg, nodes, dofs, val)
    db, dx = api.DeviceBuffer.from_numpy(Bg), api.DeviceBuffer(m)
    prm = api.Prm().quiet()
    prm.Rarray[1] = 1e-10
    dm.solve(db, dx, prm)
    assert prm.Iarray[5] == 0
    x = dx.numpy()
    fixed = dofi[nodes] + dofs
    free = np.setdiff1d(np.arange(m), fixed)
    xr = np.linalg.solve(A0[np.ix_(free, free)], f[free] - A0[np.ix_(free, fixed)] @ val)
    np.testing.assert_allclose(x[fixed], val, rtol=0, atol=1e-9)
    np.testing.assert_allclose(x[free], xr, rtol=1e-7, atol=1e-7)
    with pytest.raises(_lib.MonolisHipError, match="block size"):
        big = int(np.argmax(lst == 1))
        dm.set_dirichlet(np.zeros(m), [big], [1], [0.0])
