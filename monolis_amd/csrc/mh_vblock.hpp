// diag_V preconditioner apply as a row op (one thread per block row), used
// standalone and inside the CG / BiCGSTAB drivers for variable block sizes.
#pragma once

#include "mh_internal.hpp"
#include "mh_kernels.hpp"

namespace mh {

// dst = M_V^-1 src (src/prec/diag/diag_V.f90:177-204): per block row, forward
// substitution with the unit-lower factor, then backward substitution with
// the reciprocal pivots.  With DOT, also the partial (src, dst) and, in
// finalize, CG's rho = (r, z) and beta = rho / rho_prev (0 when `first`),
// CG.f90:71-84, or with `first` = 2 GropCG's gamma = (r, u).
template <typename T, bool DOT>
struct OpVPrec {
  static constexpr int KCLASS = 2;  // kClsPrec: applies the preconditioner
  using VT = T;
  static constexpr int BS = 1;
  static constexpr int NV = DOT ? 1 : 0;
  const T* __restrict__ src;
  T* __restrict__ dst;
  const int* __restrict__ list;
  const int* __restrict__ dof_index;
  const long long* __restrict__ doff;
  const T* __restrict__ D;
  KState* st;  // null for a standalone apply
  int first;
  __device__ bool skip() const { return st != nullptr && st->done != 0; }
  __device__ void begin() {}
  __device__ void row(int i, bool valid, double (&part)[1], const WaveRows<T, 1>&) const {
    if (!valid) return;
    const int nd = list[i];
    const int o = dof_index[i];
    const T* L = D + doff[i];
    T t[kMaxVBlock];
    for (int j = 0; j < nd; ++j) t[j] = src[o + j];
    for (int j = 1; j < nd; ++j)
      for (int k = 0; k < j; ++k) t[j] = t[j] - L[nd * j + k] * t[k];
    for (int j = nd - 1; j >= 0; --j) {
      for (int k = nd - 1; k > j; --k) t[j] = t[j] - L[nd * j + k] * t[k];
      t[j] = L[(nd + 1) * j] * t[j];
    }
    for (int j = 0; j < nd; ++j) {
      dst[o + j] = t[j];
      if constexpr (DOT) part[0] += (double)src[o + j] * (double)t[j];
    }
  }
  __device__ void finalize(const double (&tot)[1]) const {
    if constexpr (DOT) {
      if (first == 2) {
        st->gamma = tot[0];
        return;
      }
      const double rho = tot[0];
      st->beta = first ? 0.0 : rho / st->rho;
      st->rho = rho;
    }
  }
};

}  // namespace mh
