// Adjoint (reverse-mode) support for x = A^{-1} b:
//   * a device transpose of the BCSR matrix, cached on the handle, for adjoint
//     solves A^T lambda = dL/dx of nonsymmetric systems;
//   * the sparsity-restricted outer product dL/dA = -lambda x^T (one value per
//     stored block entry and block component);
//   * device-side scalar access to stored entries, the device counterpart of
//     monolis_{set,add,get}_scalar_{to,from}_sparse_matrix_main_R
//     (src/matrix/spmat_handler_util.f90:32-74, 207-252, 306-348).
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "mh_adjoint.hpp"

namespace mh {

namespace {

// keys[e] = (column block) * n + row block, vals[e] = e, for every stored block.
__global__ void k_transpose_keys(int n, const int* __restrict__ rowptr, const int* __restrict__ item,
                                 unsigned long long* __restrict__ keys, int* __restrict__ vals) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int e = rowptr[i]; e < rowptr[i + 1]; ++e) {
    keys[e] = (unsigned long long)(item[e] - 1) * (unsigned long long)n + (unsigned long long)i;
    vals[e] = e;
  }
}

// rowptrT[c] = first sorted position whose column block is >= c (c = 0..n);
// itemT[p] = row block + 1 of sorted entry p; iperm[perm[p]] = p.
__global__ void k_transpose_index(int n, int nz, const unsigned long long* __restrict__ keys,
                                  const int* __restrict__ perm, int* __restrict__ rowptrT, int* __restrict__ itemT,
                                  int* __restrict__ iperm) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= n) {
    const unsigned long long target = (unsigned long long)t * (unsigned long long)n;
    int lo = 0, hi = nz;
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (keys[mid] < target) lo = mid + 1;
      else hi = mid;
    }
    rowptrT[t] = lo;
  }
  if (t < nz) {
    itemT[t] = (int)(keys[t] % (unsigned long long)n) + 1;
    iperm[perm[t]] = t;
  }
}

// AT block iperm[e] = (A block e)^T, read from the SELL value planes.
__global__ void k_transpose_values(int n, int b, const int* __restrict__ slice_ptr, const int* __restrict__ rowptr,
                                   const double* __restrict__ val, const int* __restrict__ iperm,
                                   double* __restrict__ AT) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int bb = b * b;
  const int s = i / kWave, lane = i % kWave;
  const int k0 = slice_ptr[s];
  const int e0 = rowptr[i], len = rowptr[i + 1] - e0;
  for (int k = 0; k < len; ++k) {
    const size_t slot = (size_t)(k0 + k);
    double* dst = AT + (size_t)iperm[e0 + k] * bb;
    for (int r = 0; r < b; ++r)
      for (int c = 0; c < b; ++c) dst[c * b + r] = val[sell_vpos<double>(slot, r * b + c, bb, lane)];
  }
}

// dA[e*bb + r*b + c] = -lambda[i*b + r] * x[j*b + c] for block e = (i, j).
__global__ void k_outer_grad(int n, int nz, int b, const int* __restrict__ rowptr, const int* __restrict__ item,
                             const double* __restrict__ lambda, const double* __restrict__ x,
                             double* __restrict__ dA) {
  const int bb = b * b;
  const size_t total = (size_t)nz * bb;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
    const int e = (int)(t / bb), rc = (int)(t % bb);
    // row block of entry e: last i with rowptr[i] <= e
    int lo = 0, hi = n;
    while (hi - lo > 1) {
      const int mid = lo + (hi - lo) / 2;
      if (rowptr[mid] <= e) lo = mid;
      else hi = mid;
    }
    const int j = item[e] - 1;
    dA[t] = -lambda[(size_t)lo * b + rc / b] * x[(size_t)j * b + rc % b];
  }
}

// Scalar access to the stored entry (ci, cj, si, sj) (0-based).  mode 0 get,
// 1 set, 2 add.  out[0] = found, out[1] = value after the operation.
__global__ void k_scalar_access(int ci, int cj, int si, int sj, int b, double v, int mode,
                                const int* __restrict__ slice_ptr, const int* __restrict__ rowptr,
                                const int* __restrict__ item, double* __restrict__ val64, float* __restrict__ val32,
                                double* __restrict__ out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  // binary search of the (ascending) block columns of row ci (spmat_handler_util.f90:62-64)
  int lo = rowptr[ci], hi = rowptr[ci + 1] - 1, found = -1;
  while (lo <= hi) {
    const int mid = lo + (hi - lo) / 2;
    const int c = item[mid] - 1;
    if (c == cj) {
      found = mid;
      break;
    }
    if (c < cj) lo = mid + 1;
    else hi = mid - 1;
  }
  out[0] = found >= 0 ? 1.0 : 0.0;
  out[1] = 0.0;
  if (found < 0) return;
  const int bb = b * b;
  const size_t slot = (size_t)(slice_ptr[ci / kWave] + (found - rowptr[ci]));
  const size_t p = sell_vpos<double>(slot, si * b + sj, bb, ci % kWave);
  double a = val64[p];
  if (mode == 1) a = v;
  else if (mode == 2) a = a + v;
  if (mode != 0) {
    val64[p] = a;
    if (val32) val32[sell_vpos<float>(slot, si * b + sj, bb, ci % kWave)] = (float)a;
  }
  out[1] = a;
}

}  // namespace

int matrix_transpose(Matrix* M, hipStream_t s) {
  if (M->vb) {
    set_error("transposes of variable-block matrices are not supported");
    return -2;
  }
  if (M->np != M->n) {
    set_error("transpose needs a square local matrix (np == n), got n=%d np=%d", M->n, M->np);
    return -2;
  }
  const int n = M->n, nz = M->nz, b = M->b;
  if (!M->transpose) {
    DevArray<unsigned long long> k0, k1;
    DevArray<int> v0, rowptrT, itemT;
    MH_TRY(k0.alloc((size_t)nz + 1));
    MH_TRY(k1.alloc((size_t)nz + 1));
    MH_TRY(v0.alloc((size_t)nz + 1));
    MH_TRY(M->tperm.alloc((size_t)nz + 1));
    MH_TRY(M->tiperm.alloc((size_t)nz + 1));
    MH_TRY(rowptrT.alloc((size_t)n + 1));
    MH_TRY(itemT.alloc((size_t)nz + 1));
    if (n > 0) {
      k_transpose_keys<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(n, M->rowptr.p, M->csr_item.p, k0.p, v0.p);
      MH_HIP(hipGetLastError());
    }
    int end_bit = 1;
    while (end_bit < 64 && (1ull << end_bit) <= (unsigned long long)n * (unsigned long long)n) ++end_bit;
    if (nz > 0) {
      size_t tmp_bytes = 0;
      MH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0.p, k1.p, v0.p, M->tperm.p, nz, 0, end_bit,
                                                s));
      DevArray<unsigned char> tmp;
      MH_TRY(tmp.alloc(tmp_bytes));
      MH_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, k0.p, k1.p, v0.p, M->tperm.p, nz, 0, end_bit, s));
      MH_HIP(hipStreamSynchronize(s));
    }
    const int m = std::max(n + 1, nz);
    k_transpose_index<<<(m + kBlock - 1) / kBlock, kBlock, 0, s>>>(n, nz, k1.p, M->tperm.p, rowptrT.p, itemT.p,
                                                                  M->tiperm.p);
    MH_HIP(hipGetLastError());
    DevArray<double> AT;
    MH_TRY(AT.alloc((size_t)nz * b * b + 1));
    if (n > 0) {
      k_transpose_values<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(n, b, M->slice_ptr.p, M->rowptr.p,
                                                                     M->val64.p, M->tiperm.p, AT.p);
      MH_HIP(hipGetLastError());
    }
    Matrix* T = nullptr;
    MH_TRY(matrix_create(&T, n, n, nz, b, rowptrT.p, itemT.p, AT.p, s));
    MH_HIP(hipStreamSynchronize(s));  // temporaries are released on return
    M->transpose = T;
    M->transpose_stale = false;
    return 0;
  }
  if (M->transpose_stale) {
    DevArray<double> AT;
    MH_TRY(AT.alloc((size_t)nz * b * b + 1));
    if (n > 0) {
      k_transpose_values<<<(n + kBlock - 1) / kBlock, kBlock, 0, s>>>(n, b, M->slice_ptr.p, M->rowptr.p,
                                                                     M->val64.p, M->tiperm.p, AT.p);
      MH_HIP(hipGetLastError());
    }
    MH_TRY(matrix_set_values(M->transpose, AT.p, s));
    MH_HIP(hipStreamSynchronize(s));
    M->transpose_stale = false;
  }
  return 0;
}

int outer_grad(const Matrix& M, const double* lambda, const double* x, double* dA, hipStream_t s) {
  if (M.vb) {
    set_error("the pattern gradient of variable-block matrices is not supported");
    return -2;
  }
  const size_t total = (size_t)M.nz * M.b * M.b;
  if (total == 0) return 0;
  size_t g = (total + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  k_outer_grad<<<(unsigned)g, kBlock, 0, s>>>(M.n, M.nz, M.b, M.rowptr.p, M.csr_item.p, lambda, x, dA);
  MH_HIP(hipGetLastError());
  return 0;
}

int matrix_scalar(Matrix* M, int ci, int cj, int si, int sj, double v, int mode, double* value, int* found,
                  hipStream_t s) {
  if (M->vb) {
    set_error("scalar access to variable-block matrices is not supported (use set_values)");
    return -2;
  }
  if (ci < 0 || ci >= M->n || cj < 0 || cj >= M->np || si < 0 || si >= M->b || sj < 0 || sj >= M->b) {
    set_error("scalar access (%d, %d, %d, %d) outside the matrix (n=%d np=%d n_dof=%d)", ci, cj, si, sj, M->n,
              M->np, M->b);
    return -2;
  }
  Context* c = nullptr;
  MH_TRY(context(&c));
  DevArray<double> out;
  MH_TRY(out.alloc(2));
  k_scalar_access<<<1, 64, 0, s>>>(ci, cj, si, sj, M->b, v, mode, M->slice_ptr.p, M->rowptr.p, M->csr_item.p,
                                   M->val64.p, M->val32_ready ? M->val32.p : nullptr, out.p);
  MH_HIP(hipGetLastError());
  double h[2] = {0.0, 0.0};
  MH_HIP(hipMemcpyAsync(h, out.p, sizeof(h), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  *found = h[0] != 0.0;
  if (value) *value = h[1];
  if (mode != 0 && *found) {
    M->diag_ready = M->diag32_ready = false;
    M->transpose_stale = true;
  }
  return 0;
}

}  // namespace mh
