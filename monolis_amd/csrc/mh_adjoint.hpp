// Adjoint support: device transpose, sparsity-restricted outer product and
// scalar access to stored entries.
#pragma once

#include "mh_internal.hpp"

namespace mh {

// dA[e*b*b + r*b + c] = -lambda[i*b + r] * x[j*b + c] for every stored block
// e = (i, j); asynchronous on s.
int outer_grad(const Matrix& M, const double* lambda, const double* x, double* dA, hipStream_t s);

// Stored entry (ci, cj), block component (si, sj), all 0-based.  mode 0 reads,
// 1 sets to v, 2 adds v; *value receives the entry after the operation and
// *found whether the block is in the pattern (nothing changes if it is not).
int matrix_scalar(Matrix* M, int ci, int cj, int si, int sj, double v, int mode, double* value, int* found,
                  hipStream_t s);

}  // namespace mh
