// Complex FP64 path: SpMV, unconjugated dot, complex block-Jacobi and COCG
// (monolis src/iterative/COCG.f90) on a complex Matrix (Matrix::is_complex).
#pragma once

#include <vector>

#include "mh_internal.hpp"

namespace mh {

// y = A x over the N block rows (complex values, monolis matvec_*_C).
int zspmv(const Matrix& M, const zd* x, zd* y, hipStream_t s);
// Complex block-Jacobi factors (diag_nn_setup_C).  Returns -6 when a diagonal
// block has a zero pivot (the reference stops there).
int zdiag_setup(Matrix& M, hipStream_t s);
// out[0] + i out[1] = sum_i x_i y_i (no conjugation); out is a device pointer.
int zdot(long long m, const zd* x, const zd* y, double* out, double* partials, hipStream_t s);

struct ZSolveResult {
  int iter = 0;
  double resid_abs = 0.0;  // |sqrt(R2 / B2)|, the quantity tested against tol
  double resid_re = 0.0;   // its real part, what monolis stores as cur_resid
  int converged = 0;
  int zero_rhs = 0;
  std::vector<double> hist;  // |resid| per iteration (index = iteration)
};

// COCG with precond 0 (NONE) or 1 (DIAG); P.method must be 13.
int cocg_solve(Matrix& M, const SolveParams& P, const zd* b, zd* x, Workspace& W, ZSolveResult* out,
               hipStream_t s);

}  // namespace mh
