// Variable block sizes (monolis n_dof = -1).
//
// monolis stores block (i, j) of size n_i x n_j row-major at
// A(n_dof_index2(e) + 1 ...), with n_i = n_dof_list(i) and vector entries of
// block row i at n_dof_index(i) + 1 .. n_dof_index(i + 1).  Its SpMV
// (monolis_matvec_V_R, src/linalg/matvec.f90:231-289) computes, for scalar row
// (i, k), sum over the blocks j of row i (ascending) and over l = 1..n_j of
// A * x.  The GPU stores the scalar expansion of that matrix: scalar row r =
// n_dof_index(i) + k holds, in this same order, the entries of every block of
// row i, with scalar column n_dof_index(j) + l.  The expansion is a plain
// scalar matrix for the SELL-64 kernels, so the variable-block SpMV runs at
// scalar-CSR speed and sums in monolis's order.
//
// The block-Jacobi preconditioner keeps the blocks: diag_V
// (src/prec/diag/diag_V.f90:13-86 setup, :177-204 apply) factors each
// diagonal block with the no-pivot LU of the fixed-size paths (reciprocal
// pivots, zero pivot -> 1) and applies it per block row.  Complex values
// (monolis_matvec_V_C, matvec.f90:293-351; diag_V.f90:88-158, :208-250) take
// the same expansion with 16-byte entries; their factorisation stops on a zero
// pivot instead (diag_V.f90:123), recorded here and reported when a solve asks
// for the preconditioner.
#include <algorithm>
#include <memory>
#include <vector>

#include "mh_launch.hpp"
#include "mh_vblock.hpp"

namespace mh {

namespace {

// Scalar expansion of the values (and the column ids when fill_col).
template <typename V>
__global__ void k_v_fill(int nb, const int* __restrict__ list, const int* __restrict__ dof_index,
                         const int* __restrict__ index2, const int* __restrict__ index, const int* __restrict__ item,
                         const int* __restrict__ srowptr, const V* __restrict__ A, int* __restrict__ item_s,
                         V* __restrict__ A_s, int fill_col) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int n1 = list[i];
  for (int k = 0; k < n1; ++k) {
    int pos = srowptr[dof_index[i] + k];
    for (int e = index[i]; e < index[i + 1]; ++e) {
      const int j = item[e] - 1;
      const int n2 = list[j];
      const long long kn = (long long)index2[e] + (long long)n2 * k;
      const int c0 = dof_index[j];
      for (int l = 0; l < n2; ++l, ++pos) {
        if (fill_col) item_s[pos] = c0 + l + 1;
        A_s[pos] = A[kn + l];
      }
    }
  }
}

// diag_V setup: block row i's diagonal block, factored in place at doff[i].
// Rows without a diagonal block keep zero factors (their apply gives 0).
__global__ void k_vdiag_setup(int nb, const int* __restrict__ list, const int* __restrict__ index2,
                              const int* __restrict__ index, const int* __restrict__ item,
                              const long long* __restrict__ doff, const double* __restrict__ A,
                              double* __restrict__ D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int nd = list[i];
  double* LU = D + doff[i];
  for (int e = 0; e < nd * nd; ++e) LU[e] = 0.0;
  int found = -1;
  for (int e = index[i]; e < index[i + 1]; ++e)
    if (item[e] - 1 == i) found = e;
  if (found < 0) return;
  const double* src = A + index2[found];
  for (int e = 0; e < nd * nd; ++e) LU[e] = src[e];
  // diag_V.f90:58-72: no-pivot Doolittle, L unit lower, pivots as reciprocals
  for (int k = 0; k < nd; ++k) {
    if (LU[k * nd + k] == 0.0) LU[k * nd + k] = 1.0;
    LU[k * nd + k] = 1.0 / LU[k * nd + k];
    for (int l = k + 1; l < nd; ++l) {
      LU[l * nd + k] = LU[l * nd + k] * LU[k * nd + k];
      for (int j = k + 1; j < nd; ++j) LU[l * nd + j] = LU[l * nd + j] - LU[l * nd + k] * LU[k * nd + j];
    }
  }
}

// diag_V_C setup (diag_V.f90:88-158): as k_vdiag_setup in complex arithmetic;
// a zero pivot stops the factorisation of that row and records the smallest
// such row in *bad.
__global__ void k_zvdiag_setup(int nb, const int* __restrict__ list, const int* __restrict__ index2,
                               const int* __restrict__ index, const int* __restrict__ item,
                               const long long* __restrict__ doff, const zd* __restrict__ A, zd* __restrict__ D,
                               int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const int nd = list[i];
  zd* LU = D + doff[i];
  for (int e = 0; e < nd * nd; ++e) LU[e] = zd(0.0);
  int found = -1;
  for (int e = index[i]; e < index[i + 1]; ++e)
    if (item[e] - 1 == i) found = e;
  if (found < 0) return;
  const zd* src = A + index2[found];
  for (int e = 0; e < nd * nd; ++e) LU[e] = src[e];
  for (int k = 0; k < nd; ++k) {
    const zd piv = LU[k * nd + k];
    if (piv.re == 0.0 && piv.im == 0.0) {
      atomicMin(bad, i);
      return;
    }
    LU[k * nd + k] = zdiv(zd(1.0), piv);
    for (int l = k + 1; l < nd; ++l) {
      LU[l * nd + k] = LU[l * nd + k] * LU[k * nd + k];
      for (int j = k + 1; j < nd; ++j) LU[l * nd + j] = LU[l * nd + j] - LU[l * nd + k] * LU[k * nd + j];
    }
  }
}

__global__ void k_to_fp32_v(long long m, const double* __restrict__ a, float* __restrict__ b) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

unsigned grid_1d(long long m) {
  long long g = (m + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  return (unsigned)(g > 0 ? g : 1);
}

// Device or host source -> host vector.
template <typename U>
int to_host(std::vector<U>& dst, const U* src, size_t count, hipStream_t s) {
  dst.resize(count);
  if (count == 0) return 0;
  if (is_device_pointer(src)) {
    MH_HIP(hipMemcpyAsync(dst.data(), src, sizeof(U) * count, hipMemcpyDeviceToHost, s));
    MH_HIP(hipStreamSynchronize(s));
  } else {
    std::copy(src, src + count, dst.begin());
  }
  return 0;
}

template <typename V>
DevArray<V>& vkeep(VBlock& Vb);
template <>
DevArray<double>& vkeep<double>(VBlock& Vb) {
  return Vb.A64;
}
template <>
DevArray<zd>& vkeep<zd>(VBlock& Vb) {
  return Vb.Az;
}

// Values of the expansion (device) from monolis's A (host or device), and the
// diag_V factors from the same A.
template <typename V>
int fill_vblock_values(Matrix* M, const V* A, hipStream_t s, DevArray<int>* item_s, DevArray<V>* A_s) {
  constexpr bool cplx = sizeof(V) == sizeof(zd);
  VBlock& Vb = *M->vb;
  const V* dA = A;
  DevArray<V> tmp;
  if (!is_device_pointer(A)) {
    MH_TRY(tmp.alloc((size_t)Vb.nzs + 1));
    MH_TRY(upload(tmp.p, A, sizeof(V) * (size_t)Vb.nzs, s));
    dA = tmp.p;
  }
  {  // the monolis-layout copy that device assembly updates
    DevArray<V>& keep = vkeep<V>(Vb);
    if (keep.n < (size_t)Vb.nzs + 1) MH_TRY(keep.alloc((size_t)Vb.nzs + 1));
    if (dA != keep.p && Vb.nzs > 0) MH_HIP(hipMemcpyAsync(keep.p, dA, sizeof(V) * (size_t)Vb.nzs, hipMemcpyDeviceToDevice, s));
  }
  DevArray<int> bad;
  if constexpr (cplx) {
    if (Vb.Dz.n < Vb.D64.n) MH_TRY(Vb.Dz.alloc(Vb.D64.n));
    MH_TRY(bad.alloc(1));
    MH_HIP(hipMemsetAsync(bad.p, 0x7f, sizeof(int), s));
  }
  if (Vb.nb > 0) {
    const unsigned grid = (unsigned)((Vb.nb + kBlock - 1) / kBlock);
    k_v_fill<V><<<grid, kBlock, 0, s>>>(Vb.nb, Vb.list.p, Vb.dof_index.p, Vb.index2.p, Vb.index.p, Vb.item.p,
                                        Vb.srowptr.p, dA, item_s ? item_s->p : nullptr, A_s->p, item_s != nullptr);
    MH_HIP(hipGetLastError());
    if constexpr (cplx) {
      k_zvdiag_setup<<<grid, kBlock, 0, s>>>(Vb.nb, Vb.list.p, Vb.index2.p, Vb.index.p, Vb.item.p, Vb.doff.p, dA,
                                             Vb.Dz.p, bad.p);
    } else {
      k_vdiag_setup<<<grid, kBlock, 0, s>>>(Vb.nb, Vb.list.p, Vb.index2.p, Vb.index.p, Vb.item.p, Vb.doff.p, dA,
                                            Vb.D64.p);
    }
    MH_HIP(hipGetLastError());
  }
  Vb.d32_ready = false;
  if constexpr (cplx) {
    int hbad = 0x7f7f7f7f;
    MH_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    MH_HIP(hipStreamSynchronize(s));
    Vb.zbad = hbad == 0x7f7f7f7f ? -1 : hbad;
  }
  MH_HIP(hipStreamSynchronize(s));  // tmp is released on return
  return 0;
}

// Shared by the real and complex builds: validates monolis's block arrays,
// uploads them, expands the values and builds the scalar SELL-64 matrix.
template <typename VT>
int vblock_build(Matrix** out, int n, int np, int nz, const int* n_dof_list, const int* n_dof_index,
                 const int* n_dof_index2, const int* index, const int* item, const VT* A, hipStream_t s);

template <typename VT>
int vblock_build(Matrix** out, int n, int np, int nz, const int* n_dof_list, const int* n_dof_index,
                 const int* n_dof_index2, const int* index, const int* item, const VT* A, hipStream_t s) {
  constexpr bool cplx = sizeof(VT) == sizeof(zd);
  *out = nullptr;
  if (n < 0 || np < n || nz < 0) {
    set_error("invalid matrix sizes n=%d np=%d nz=%d", n, np, nz);
    return -2;
  }
  // the block arrays are small (O(np + nz) ints): validate them on the host
  std::vector<int> list, dofi, idx, itm, idx2;
  MH_TRY(to_host(list, n_dof_list, (size_t)np, s));
  MH_TRY(to_host(dofi, n_dof_index, (size_t)np + 1, s));
  MH_TRY(to_host(idx, index, (size_t)np + 1, s));
  MH_TRY(to_host(itm, item, (size_t)nz, s));
  MH_TRY(to_host(idx2, n_dof_index2, (size_t)nz + 1, s));
  if (idx[0] != 0 || idx[np] != nz) {
    set_error("invalid BCSR index (index[0] = %d, index[np] = %d, nz = %d)", idx[0], idx[np], nz);
    return -2;
  }
  int maxb = 0;
  if (dofi[0] != 0) {
    set_error("n_dof_index[0] must be 0");
    return -2;
  }
  for (int i = 0; i < np; ++i) {
    if (list[i] < 1 || dofi[i + 1] - dofi[i] != list[i]) {
      set_error("n_dof_list / n_dof_index disagree at block row %d", i);
      return -2;
    }
    maxb = std::max(maxb, list[i]);
    if (idx[i + 1] < idx[i]) {
      set_error("index not non-decreasing at row %d", i);
      return -2;
    }
  }
  if (maxb > kMaxVBlock) {
    set_error("variable block of size %d is larger than the supported %d", maxb, kMaxVBlock);
    return -2;
  }
  if (idx2[0] != 0) {
    set_error("n_dof_index2[0] must be 0");
    return -2;
  }
  std::vector<int> slen((size_t)dofi[np] + 1, 0);
  for (int i = 0; i < np; ++i) {
    int rowlen = 0;
    for (int e = idx[i]; e < idx[i + 1]; ++e) {
      const int j = itm[e] - 1;
      if (j < 0 || j >= np) {
        set_error("item outside 1..NP at entry %d", e);
        return -2;
      }
      if ((long long)idx2[e + 1] - idx2[e] != (long long)list[i] * list[j]) {
        set_error("n_dof_index2 disagrees with the block sizes at entry %d", e);
        return -2;
      }
      rowlen += list[j];
    }
    if (i < n)
      for (int k = 0; k < list[i]; ++k) slen[(size_t)dofi[i] + k] = rowlen;
  }
  const int ns = dofi[n], nps = dofi[np];
  std::vector<int> srow((size_t)nps + 1, 0);
  long long acc = 0;
  for (int r = 0; r < nps; ++r) {
    srow[r] = (int)acc;
    acc += slen[r];
    if (acc >= (1LL << 31)) {
      set_error("scalar expansion of the variable-block matrix has more than 2^31 entries");
      return -2;
    }
  }
  srow[nps] = (int)acc;
  if (acc != idx2[idx[n]]) {  // entries of the internal rows
    set_error("n_dof_index2 total %d does not match the block sizes (%lld)", idx2[idx[n]], acc);
    return -2;
  }
  auto* V = new VBlock();
  std::unique_ptr<VBlock> vguard(V);
  V->nb = n;
  V->npb = np;
  V->nzb = nz;
  V->maxb = maxb;
  V->nzs = idx2[nz];
  std::vector<long long> doff((size_t)n + 1, 0);
  for (int i = 0; i < n; ++i) doff[i + 1] = doff[i] + (long long)list[i] * list[i];
  MH_TRY(V->list.alloc((size_t)np + 1));
  MH_TRY(V->dof_index.alloc((size_t)np + 1));
  MH_TRY(V->index.alloc((size_t)np + 1));
  MH_TRY(V->item.alloc((size_t)nz + 1));
  MH_TRY(V->index2.alloc((size_t)nz + 1));
  MH_TRY(V->doff.alloc((size_t)n + 1));
  MH_TRY(V->srowptr.alloc((size_t)nps + 1));
  MH_TRY(V->D64.alloc((size_t)doff[n] + 1));
  MH_TRY(upload(V->list.p, list.data(), sizeof(int) * np, s));
  MH_TRY(upload(V->dof_index.p, dofi.data(), sizeof(int) * ((size_t)np + 1), s));
  MH_TRY(upload(V->index.p, idx.data(), sizeof(int) * ((size_t)np + 1), s));
  MH_TRY(upload(V->item.p, itm.data(), sizeof(int) * nz, s));
  MH_TRY(upload(V->index2.p, idx2.data(), sizeof(int) * ((size_t)nz + 1), s));
  MH_TRY(upload(V->doff.p, doff.data(), sizeof(long long) * ((size_t)n + 1), s));
  MH_TRY(upload(V->srowptr.p, srow.data(), sizeof(int) * ((size_t)nps + 1), s));
  // expansion: scalar column ids and values, then the SELL build
  Matrix tmpM;
  tmpM.vb = V;
  vguard.release();
  DevArray<int> item_s;
  DevArray<VT> A_s;
  int rc = 0;
  if ((rc = item_s.alloc((size_t)acc + 1)) || (rc = A_s.alloc((size_t)acc + 1)) ||
      (rc = fill_vblock_values<VT>(&tmpM, A, s, &item_s, &A_s))) {
    return rc;  // tmpM owns V and frees it
  }
  Matrix* M = nullptr;
  if constexpr (cplx)
    rc = matrix_create_complex(&M, ns, nps, (int)acc, 1, V->srowptr.p, item_s.p, (const double*)A_s.p, s);
  else
    rc = matrix_create(&M, ns, nps, (int)acc, 1, V->srowptr.p, item_s.p, A_s.p, s);
  if (rc) return rc;
  MH_HIP(hipStreamSynchronize(s));
  M->vb = V;
  tmpM.vb = nullptr;
  *out = M;
  return 0;
}

}  // namespace

int matrix_create_vblock(Matrix** out, int n, int np, int nz, const int* n_dof_list, const int* n_dof_index,
                         const int* n_dof_index2, const int* index, const int* item, const double* A,
                         hipStream_t s) {
  return vblock_build<double>(out, n, np, nz, n_dof_list, n_dof_index, n_dof_index2, index, item, A, s);
}

int matrix_create_vblock_complex(Matrix** out, int n, int np, int nz, const int* n_dof_list, const int* n_dof_index,
                                 const int* n_dof_index2, const int* index, const int* item, const zd* A,
                                 hipStream_t s) {
  return vblock_build<zd>(out, n, np, nz, n_dof_list, n_dof_index, n_dof_index2, index, item, A, s);
}

int matrix_set_values_vblock(Matrix* M, const double* A, hipStream_t s) {
  DevArray<double> A_s;
  MH_TRY(A_s.alloc((size_t)M->vb->nzs + 1));
  MH_TRY(fill_vblock_values<double>(M, A, s, nullptr, &A_s));
  MH_TRY(matrix_set_values(M, A_s.p, s));
  MH_HIP(hipStreamSynchronize(s));
  return 0;
}

int matrix_set_values_vblock_complex(Matrix* M, const zd* A, hipStream_t s) {
  DevArray<zd> A_s;
  MH_TRY(A_s.alloc((size_t)M->vb->nzs + 1));
  MH_TRY(fill_vblock_values<zd>(M, A, s, nullptr, &A_s));
  MH_TRY(matrix_set_values(M, (const double*)A_s.p, s));
  MH_HIP(hipStreamSynchronize(s));
  return 0;
}

int vdiag_make_fp32(Matrix& M, hipStream_t s) {
  VBlock& V = *M.vb;
  if (V.d32_ready) return 0;
  const long long m = (long long)V.D64.n;
  if ((long long)V.D32.n < m) MH_TRY(V.D32.alloc((size_t)m));
  k_to_fp32_v<<<grid_1d(m), kBlock, 0, s>>>(m, V.D64.p, V.D32.p);
  MH_HIP(hipGetLastError());
  V.d32_ready = true;
  return 0;
}

template <typename T>
int vdiag_apply(const Matrix& M, const T* x, T* y, hipStream_t s) {
  const VBlock& V = *M.vb;
  if (V.nb == 0) return 0;
  const T* D = sizeof(T) == 8 ? (const T*)V.D64.p : (const T*)V.D32.p;
  OpVPrec<T, false> op{x, y, V.list.p, V.dof_index.p, V.doff.p, D, nullptr, 0};
  return launch_rows(V.nb, op, nullptr, s);
}
template int vdiag_apply<double>(const Matrix&, const double*, double*, hipStream_t);
template int vdiag_apply<float>(const Matrix&, const float*, float*, hipStream_t);

}  // namespace mh
