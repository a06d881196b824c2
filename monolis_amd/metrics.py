"""Algorithmic byte counts (the roofline numerators) for the Krylov path.

Fixed definitions, int32 indices, every matrix value and index read once,
the input vector read once and the output written once:

    SpMV_B    = nnzb * (b^2 * s + 4) + (N_b + 1) * 4 + 2 * N_b * b * s
    CG_iter_B = SpMV_B + N_b * b^2 * s + 11 * N_b * b * s

with s = 8 (FP64) or 4 (FP32).  The 11 vector passes of a CG iteration are:
preconditioner r -> z (2), p update (3), x / r update and ||r||^2 (6).
"""
from __future__ import annotations

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published peak


def spmv_bytes(nb: int, nnzb: int, b: int, s: int) -> int:
    return nnzb * (b * b * s + 4) + (nb + 1) * 4 + 2 * nb * b * s


def cg_iter_bytes(nb: int, nnzb: int, b: int, s: int) -> int:
    return spmv_bytes(nb, nnzb, b, s) + nb * b * b * s + 11 * nb * b * s
