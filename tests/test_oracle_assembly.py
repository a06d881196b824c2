"""Pins the oracle's assembly restatement (monolis_oracle_asm.c): the
reference's own unit-test values and dense brute force."""
import numpy as np
import pytest

import oracle
from _helpers import element_matrices, hex_mesh
from monolis_amd import matrix as mm


def _mat(n_node, b, elems):
    index, item = oracle.pattern_by_simple_mesh(n_node, elems)
    return mm.Matrix(n_node, n_node, b, index, item, np.zeros(len(item) * b * b)).ensure_dof_index()


def test_pattern_reference_unit_test():
    """src_test/matrix/spmat_nzpattern_test.f90:38-73 (4-node chain, 2-node elements)."""
    index, item = oracle.pattern_by_simple_mesh(4, [[0, 1], [1, 2], [2, 3]])
    assert index.tolist() == [0, 2, 5, 8, 10]
    assert item.tolist() == [1, 2, 1, 2, 3, 2, 3, 4, 3, 4]


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_pattern_matches_pair_sets(seed):
    rng = np.random.default_rng(seed)
    n_node, n_elem, n_base = 60, 45, 4
    elems = np.array([rng.choice(n_node, n_base, replace=False) for _ in range(n_elem)], np.int32)
    elems[3, 2] = elems[3, 1]                     # a degenerate element repeats a node
    index, item = oracle.pattern_by_simple_mesh(n_node, elems)
    pairs = {(i, i) for i in range(n_node)}
    for e in elems:
        pairs |= {(int(a), int(c)) for a in e for c in e}
    rows = [sorted(c for (r, c) in pairs if r == i) for i in range(n_node)]
    assert index.tolist() == [0] + np.cumsum([len(r) for r in rows]).tolist()
    assert item.tolist() == [c + 1 for r in rows for c in r]
    # agrees with the host builder used elsewhere
    m = mm.nonzero_pattern_by_simple_mesh(n_node, 1, elems.tolist())
    assert np.array_equal(m.index, index) and np.array_equal(m.item, item)


def test_element_add_reference_unit_test():
    """src_test/matrix/spmat_handler_test.f90:46-104: set_block, add_scalar, then
    monolis_add_matrix_to_sparse_matrix_R with a one-node element on node 2."""
    m = _mat(4, 2, [[0, 1], [1, 2], [2, 3]])
    mm.set_scalar(m, 1, 1, 0, 0, 1.0)
    mm.set_scalar(m, 1, 1, 0, 1, 2.0)
    mm.set_scalar(m, 1, 1, 1, 0, 3.0)
    mm.set_scalar(m, 1, 1, 1, 1, 4.0)
    mm.add_scalar(m, 1, 1, 0, 0, 1.0)
    oracle.add_elements(m, [[1]], np.array([[[1.0, 2.0], [3.0, 4.0]]]))
    assert m.A[12:16].tolist() == [3.0, 4.0, 6.0, 8.0]      # A(13..16) of the reference test


def test_dirichlet_reference_unit_test():
    """src_test/matrix/spmat_handler_test.f90:119-132: A = 2 everywhere, B = 0,
    condition on node 2 (1-based), DOF 1, value 1."""
    m = _mat(4, 2, [[0, 1], [1, 2], [2, 3]])
    m.A[:] = 2.0
    B = np.zeros(8)
    oracle.dirichlet(m, B, [1], [0], [1.0])
    assert B.tolist() == [-2.0, -2.0, 1.0, -2.0, -2.0, -2.0, 0.0, 0.0]


@pytest.mark.parametrize("b", [1, 2, 3])
def test_element_add_matches_dense_sequential_sum(b):
    n_node, elems = hex_mesh(3, 2, 2)
    m = _mat(n_node, b, elems)
    Ke = element_matrices(len(elems), 8 * b, seed=b, spd=False)
    oracle.add_elements(m, elems, Ke)
    # dense, element by element in order (np.add.at is unbuffered and in order)
    D = np.zeros((n_node * b, n_node * b))
    dofs = (elems[:, :, None] * b + np.arange(b)).reshape(len(elems), -1)
    rr = np.repeat(dofs[:, :, None], dofs.shape[1], axis=2).reshape(-1)
    cc = np.repeat(dofs[:, None, :], dofs.shape[1], axis=1).reshape(-1)
    np.add.at(D, (rr, cc), Ke.reshape(-1))
    assert np.array_equal(m.dense(), D)


def test_element_add_rejects_block_outside_pattern():
    m = _mat(4, 1, [[0, 1], [1, 2], [2, 3]])
    with pytest.raises(KeyError):
        oracle.add_elements(m, [[0, 3]], np.ones((1, 2, 2)))


@pytest.mark.parametrize("b", [1, 3])
def test_dirichlet_solution_is_the_reduced_solve(b):
    n_node, elems = hex_mesh(3)
    m = _mat(n_node, b, elems)
    oracle.add_elements(m, elems, element_matrices(len(elems), 8 * b, seed=7))
    A0 = m.dense().copy()
    rng = np.random.default_rng(3)
    f = rng.uniform(-1, 1, n_node * b)
    nodes = np.arange(0, n_node, 5, dtype=np.int32)
    dof = (nodes % b).astype(np.int32)
    val = rng.uniform(-1, 1, len(nodes))
    B = f.copy()
    oracle.dirichlet(m, B, nodes, dof, val)
    x = np.linalg.solve(m.dense(), B)
    fixed = nodes * b + dof
    free = np.setdiff1d(np.arange(n_node * b), fixed)
    xr = np.linalg.solve(A0[np.ix_(free, free)], f[free] - A0[np.ix_(free, fixed)] @ val)
    np.testing.assert_allclose(x[fixed], val, rtol=0, atol=1e-12)
    np.testing.assert_allclose(x[free], xr, rtol=1e-9, atol=1e-9)
    # condition rows and columns are unit rows / zero columns
    Ad = m.dense()
    assert np.array_equal(Ad[fixed][:, fixed], np.eye(len(fixed)))
    assert not Ad[np.ix_(free, fixed)].any() and not Ad[np.ix_(fixed, free)].any()


# --- complex conditions (monolis_set_Dirichlet_bc_C) -------------------------------------
def test_dirichlet_C_reference_unit_test():
    """src_test/matrix/spmat_handler_test.f90:236-248: A = (2, 2) everywhere, B = 0,
    condition on node 2 (1-based), DOF 1, value (1, 0)."""
    m = _mat(4, 2, [[0, 1], [1, 2], [2, 3]])
    A = np.full(len(m.A), 2.0 + 2.0j)
    B = np.zeros(8, np.complex128)
    oracle.dirichlet_C(m.index, m.item, A, 4, 2, B, [1], [0], [1.0 + 0.0j])
    assert B.tolist() == [-2 - 2j, -2 - 2j, 1 + 0j, -2 - 2j, -2 - 2j, -2 - 2j, 0j, 0j]


@pytest.mark.parametrize("b", [1, 2, 3])
def test_dirichlet_C_reduces_to_the_real_conditions(b):
    """Closed forms of the complex product: with real A, B and values the complex
    conditions equal the real ones (imaginary parts stay 0); with values i*v the
    real parts of B stay as they are and the imaginary parts receive the real update
    of a zero right-hand side with values v (B.im - v*A), bit for bit."""
    n_node, elems = hex_mesh(3)
    m = _mat(n_node, b, elems)
    oracle.add_elements(m, elems, element_matrices(len(elems), 8 * b, seed=11))
    rng = np.random.default_rng(5)
    f = rng.uniform(-1, 1, n_node * b)
    nodes = np.arange(1, n_node, 4, dtype=np.int32)
    dof = ((nodes // 4) % b).astype(np.int32)
    v = rng.uniform(-1, 1, len(nodes))
    # real values
    mr = _mat(n_node, b, elems)
    mr.A[:] = m.A
    Br = f.copy()
    oracle.dirichlet(mr, Br, nodes, dof, v)
    A = m.A.astype(np.complex128)
    B = f.astype(np.complex128)
    oracle.dirichlet_C(m.index, m.item, A, n_node, b, B, nodes, dof, v.astype(np.complex128))
    assert np.array_equal(B.real, Br) and not B.imag.any()
    assert np.array_equal(A.real, mr.A) and not A.imag.any()
    # imaginary values
    mi = _mat(n_node, b, elems)
    mi.A[:] = m.A
    Bi = np.zeros(n_node * b)
    oracle.dirichlet(mi, Bi, nodes, dof, v)
    A = m.A.astype(np.complex128)
    B = f.astype(np.complex128)
    oracle.dirichlet_C(m.index, m.item, A, n_node, b, B, nodes, dof, 1j * v)
    fixed = nodes * b + dof
    free = np.setdiff1d(np.arange(n_node * b), fixed)
    assert np.array_equal(B.real[free], f[free]) and np.array_equal(B.imag[free], Bi[free])
    assert np.array_equal(B[fixed], 1j * v)
    assert np.array_equal(A.real, mi.A) and not A.imag.any()
