"""The oracle's variable-block path (monolis n_dof = -1): monolis_matvec_V_R
(src/linalg/matvec.f90:231-289) and diag_V (src/prec/diag/diag_V.f90), pinned
on the CPU against:
  * the reference's own matvec_V test inputs (src_test/linalg/matvec_test.f90:42-108:
    4 nodes with 1, 2, 3, 1 DOFs on two triangles, A(k) = mod(k, 7) + 1,
    X(k) = mod(k, 5) + 1), with the dense product as the exact answer;
  * dense products and dense solves of random variable-block matrices;
  * per-block numpy solves for diag_V (apply(D_i v) = D_i^-1 v);
  * planted mistakes (a transposed block) that the checks must catch.
"""
import numpy as np
import pytest

import oracle
from monolis_amd import matrix as mm

from _helpers import random_vblock


def reference_matvec_v_case():
    m = mm.nonzero_pattern_by_simple_mesh_V(4, [1, 2, 3, 1], [(0, 1, 2), (1, 2, 3)])
    k = np.arange(1, len(m.A) + 1)
    m.A = (k % 7 + 1).astype(np.float64)
    x = (np.arange(1, m.npndof + 1) % 5 + 1).astype(np.float64)
    return m, x


def test_matvec_v_reference_inputs_exact():
    m, x = reference_matvec_v_case()
    assert m.npndof == 7 and list(m.n_dof_index) == [0, 1, 3, 6, 7]
    y = oracle.matvec(m, x)
    assert np.array_equal(y[:7], m.dense() @ x)   # small integers: exact


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_matvec_v_random_vs_dense(seed):
    m = random_vblock(40, 0.1, seed, maxb=5, spd=False)
    x = np.random.default_rng(seed).uniform(-1, 1, m.npndof)
    y = oracle.matvec(m, x)
    assert np.allclose(y, m.dense() @ x, rtol=1e-13, atol=1e-13)
    # planted mistake: transpose one off-diagonal block of size > 1 x 1
    bad = m.A.copy()
    for i in range(m.n):
        for e in range(m.index[i], m.index[i + 1]):
            j = m.item[e] - 1
            n1, n2 = m.n_dof_list[i], m.n_dof_list[j]
            if i != j and n1 == n2 and n1 > 1:
                a0 = m.n_dof_index2[e]
                bad[a0:a0 + n1 * n2] = bad[a0:a0 + n1 * n2].reshape(n1, n2).T.reshape(-1)
                mb = mm.Matrix(m.n, m.np_, -1, m.index, m.item, bad, n_dof_list=m.n_dof_list,
                               n_dof_index=m.n_dof_index, n_dof_index2=m.n_dof_index2)
                assert not np.allclose(oracle.matvec(mb, x), y)
                return
    pytest.skip("no square off-diagonal block larger than 1x1 in this pattern")


def test_diag_v_apply_solves_each_block():
    m = random_vblock(30, 0.1, 7, maxb=6, spd=True)
    D = oracle.diag_setup(m)
    v = np.random.default_rng(8).uniform(-1, 1, m.nndof)
    z = oracle.diag_apply(m, D, v)
    Dense = m.dense()
    for i in range(m.n):
        s = slice(m.n_dof_index[i], m.n_dof_index[i + 1])
        assert np.allclose(z[s], np.linalg.solve(Dense[s, s], v[s]), rtol=1e-12, atol=1e-12)


def test_diag_v_zero_pivot_rule():
    m = random_vblock(6, 0.3, 9, sizes=[2, 3, 1, 2, 3, 1], spd=True)
    # zero the first diagonal entry of block row 1: the pivot becomes 1 (diag_V.f90:60)
    e = [e for e in range(m.index[1], m.index[2]) if m.item[e] == 2][0]
    m.A[m.n_dof_index2[e]] = 0.0
    doff, D = oracle.diag_setup(m)
    assert D[doff[1]] == 1.0


@pytest.mark.parametrize("method", [oracle.ITER_CG, oracle.ITER_BICGSTAB])
def test_solve_v_matches_dense(method):
    m = random_vblock(50, 0.08, 11, maxb=4, spd=True)
    b = np.random.default_rng(12).uniform(-1, 1, m.nndof)
    x, st, I, R = oracle.solve(m, b, method=method, precond=oracle.PREC_DIAG, tol=1e-12)
    assert st == 0
    xd = np.linalg.solve(m.dense(), b)
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
