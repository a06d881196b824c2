"""Writes tests/golden/monolis_kat.json: known answers taken from monolis's own tests.

Every entry is data (inputs and expected outputs) copied from the reference
test named in its "source" field; nothing here runs reference code.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))

CHAIN5 = [[0, 1], [1, 2], [2, 3], [3, 4]]
# (i, j, value) triples added to sub-block (0,0); sub-blocks (1,1) and (2,2)
# repeat them scaled by 2 and 4 (wrapper_test/linalg/monolis_matvec_c_test.c:38-50, :103-128, :200-237)
R11_ENTRIES = [[0, 0, 2.0], [0, 1, 1.0], [1, 0, 1.0], [1, 1, 2.0], [1, 2, 3.0], [2, 1, 1.0], [2, 2, 2.0],
               [2, 3, 4.0], [3, 2, 1.0], [3, 3, 2.0], [3, 4, 5.0], [4, 3, 1.0], [4, 4, 2.0]]


def matvec_kats():
    out = []
    for ndof, expect, src in [
        (1, [3.0, 6.0, 7.0, 8.0, 3.0], "wrapper_test/linalg/monolis_matvec_c_test.c:63-67"),
        (2, [3.0, 6.0, 6.0, 12.0, 7.0, 14.0, 8.0, 16.0, 3.0, 6.0],
         "wrapper_test/linalg/monolis_matvec_c_test.c:141-150"),
        (3, [3.0, 6.0, 12.0, 6.0, 12.0, 24.0, 7.0, 14.0, 28.0, 8.0, 16.0, 32.0, 3.0, 6.0, 12.0],
         "wrapper_test/linalg/monolis_matvec_c_test.c:238-252"),
    ]:
        adds = []
        for d in range(ndof):
            for i, j, v in R11_ENTRIES:
                adds.append([i, j, d, d, v * (2 ** d)])
        out.append({"name": f"R{ndof}{ndof}", "source": src, "n_node": 5, "ndof": ndof, "elems": CHAIN5,
                    "add": adds, "x": [1.0] * (5 * ndof), "y": expect})
    return out


def glibc_rand_sequence(count: int):
    """glibc rand() from the default seed, as the reference C test draws it."""
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)
    return [int(libc.rand()) for _ in range(count)]


def solver_c_cases():
    """wrapper_test/solver/monolis_solver_c_test.c:8-158: 10-node chain, 2 DOF.
    Real case: diagonal rand()%1001+5000 and off-diagonal blocks rand()%1001,
    each written untransposed at (i, i+1) and (i+1, i) (a nonsymmetric matrix),
    b = A 1, tol 1e-10, x = 1 expected.  Complex case, drawn next from the same
    stream: the same values v as v + 0.1 v i, solved by COCG to x = 1 + i.
    Both draws start from srand(1); monolis_amd/c/test_c_abi.c draws the same
    values before its first library call.  ("oracle_iter" is the iteration
    count of this repository's CPU oracle on the drawn matrix, one thread; the
    reference test itself checks only x.)"""
    n_node, ndof = 10, 2
    elems = [[i, i + 1] for i in range(9)]
    r = iter(glibc_rand_sequence(1000))
    out = []
    for name, src, scale_im, x_re, x_im, iters in (
            ("solve_c_R", "wrapper_test/solver/monolis_solver_c_test.c:8-81", 0.0, 1.0, 0.0, 14),
            ("solve_c_C", "wrapper_test/solver/monolis_solver_c_test.c:83-158", 0.1, 1.0, 1.0, 35)):
        sets = []
        for i in range(n_node):
            for j in range(ndof):
                v = float(next(r) % 1001 + 5000)
                sets.append([i, i, j, j, v, scale_im * v])
        for e in elems:
            for j in range(ndof):
                for k in range(ndof):
                    v = float(next(r) % 1001)
                    sets.append([e[0], e[1], j, k, v, scale_im * v])
                    sets.append([e[1], e[0], j, k, v, scale_im * v])
        case = {"name": name, "source": src, "n_node": n_node, "ndof": ndof, "elems": elems, "tol": 1e-10,
                "oracle_iter": iters}
        if scale_im == 0.0:
            case["set"] = [s[:5] for s in sets]
            case["x_expected"] = [x_re] * (n_node * ndof)
        else:
            case["set_complex"] = sets
            case["x_expected_re_im"] = [x_re, x_im]
        out.append(case)
    return out


def mtx_case():
    """src_test/input.f/mtx.dat (10x10, diagonal 4, off-diagonal 1) solved by every
    (method, precond) pair with x = 1 expected (src_test/input.f/main.f90:199-231)."""
    entries = [[i, i, 4.0] for i in range(10)] + [[i, i + 1, 1.0] for i in range(9)]
    return {"name": "mtx_dat", "source": "src_test/input.f/mtx.dat, src_test/input.f/main.f90:176-231",
            "n_node": 10, "ndof": 1, "entries": entries, "tol": 1e-10, "x_expected": [1.0] * 10}


def blas_kats():
    return {
        "inner_product": {"source": "src_test/linalg/inner_product_test.f90:63-73",
                          "x": [1.0, 1.0, 1.0, 1.0], "y": [1.0, 2.0, 3.0, 4.0], "sum": 10.0},
        "axpby": {"source": "src_test/linalg/vec_util_test.f90:100-118",
                  "alpha": 2.0, "x": [1.0] * 4, "beta": 2.0, "y": [2.0] * 4, "z": [6.0] * 4},
        "converge": {"source": "src_test/linalg/mat_converge_test.f90:36-55",
                     "N": 2, "ndof": 2, "B2": 4.0, "R": [2.0] * 4, "iter": 10,
                     "cur_iter": 10, "cur_resid": 2.0, "is_converge": False},
    }


def main():
    solve_r, solve_c = solver_c_cases()
    data = {"matvec": matvec_kats(), "solve_c": solve_r, "solve_c_complex": solve_c, "mtx": mtx_case(),
            "blas": blas_kats()}
    path = os.path.join(HERE, "monolis_kat.json")
    with open(path, "w") as f:
        json.dump(data, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
