// Device-side matrix assembly, the GPU counterpart of monolis's host assembly:
//   * the nonzero pattern of a simple mesh (monolis_get_nonzero_pattern_by_simple_mesh_R,
//     src/matrix/spmat_nzpattern.f90:15-49, with the nodal graph of
//     spmat_nzpattern_util.f90:13-58): (row, column) keys of every node pair of
//     every element plus the diagonal, radix-sorted and made unique;
//   * batched element matrices (monolis_add_matrix_to_sparse_matrix_main_R,
//     src/matrix/spmat_handler_util.f90:398-497): the node pairs of all
//     elements keyed by their block position, stably sorted (element order
//     kept), then every block component summed left to right onto its current
//     value, so the result equals one add-matrix call per element in element
//     order, bit for bit;
//   * Dirichlet conditions (monolis_set_Dirichlet_bc_main_R,
//     src/matrix/spmat_handler_util.f90:604-677) for a whole list in one pass;
//   * value and pattern read-back in monolis's layout.
// The value kernels are shared by real and complex matrices (complex element
// matrices: monolis_add_matrix_to_sparse_matrix_main_C, spmat_handler_util.f90:
// 501-600, the same loop on complex(kdouble); complex conditions:
// monolis_set_Dirichlet_bc_main_C, :681-749, whose B update multiplies two
// complex numbers as (ar*br - ai*bi, ar*bi + ai*br) without contraction).
#include <algorithm>
#include <cstdlib>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "mh_assembly.hpp"
#include "mh_launch.hpp"

namespace mh {

namespace {

int grid_for(long long n) {
  const long long g = (n + kBlock - 1) / kBlock;
  return (int)(g > 0 ? g : 1);
}

int bits_for(unsigned long long v) {
  int b = 1;
  while (b < 64 && (v >> b) != 0) ++b;
  return b;
}

// Device copy of a host array (device pointers are used in place).
template <typename T>
int stage(const T* src, size_t count, DevArray<T>& tmp, const T** out, hipStream_t s) {
  if (count == 0 || is_device_pointer(src)) {
    *out = src;
    return 0;
  }
  MH_TRY(tmp.alloc(count));
  MH_TRY(upload(tmp.p, src, count * sizeof(T), s));
  *out = tmp.p;
  return 0;
}

int read_flag(const int* d, int* h, hipStream_t s) {
  MH_HIP(hipMemcpyAsync(h, d, sizeof(int), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  return 0;
}

__global__ void k_check_range(long long m, const int* __restrict__ v, int hi, int* __restrict__ bad) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x) {
    const int x = v[i];
    if (x < 0 || x >= hi) atomicOr(bad, 1);
  }
}

// keys of every ordered node pair (a != c) of every element, then the diagonal
__global__ void k_mesh_keys(int n_node, int n_base, int n_elem, const int* __restrict__ elem,
                            unsigned long long* __restrict__ keys) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = (long long)n_base * (n_base - 1);
  const long long npair = (long long)n_elem * per;
  if (t < npair) {
    const long long e = t / per;
    const int q = (int)(t % per);
    const int a = q / (n_base - 1);
    int c = q % (n_base - 1);
    if (c >= a) ++c;
    keys[t] = ((unsigned long long)elem[e * n_base + a] << 32) | (unsigned)elem[e * n_base + c];
  } else if (t < npair + n_node) {
    const unsigned i = (unsigned)(t - npair);
    keys[t] = ((unsigned long long)i << 32) | i;
  }
}

// index[r] = first key of row r, item = column + 1 (keys sorted and unique)
__global__ void k_pattern_csr(int n_node, int nz, const unsigned long long* __restrict__ keys,
                              int* __restrict__ index, int* __restrict__ item) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t <= n_node) {
    const unsigned long long target = (unsigned long long)t << 32;
    int lo = 0, hi = nz;
    while (lo < hi) {
      const int mid = lo + (hi - lo) / 2;
      if (keys[mid] < target) lo = mid + 1;
      else hi = mid;
    }
    index[t] = lo;
  }
  if (t < nz) item[t] = (int)(keys[t] & 0xffffffffull) + 1;
}

// Block position of (row, col) in the CSR, -1 when absent (item sorted per row).
__device__ __forceinline__ int find_block(const int* __restrict__ rowptr, const int* __restrict__ item, int row,
                                          int col) {
  int lo = rowptr[row];
  const int end = rowptr[row + 1];
  int hi = end;
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (item[mid] < col + 1) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && item[lo] == col + 1) ? lo : -1;
}

// One thread per (element, a, c) node pair: key = block position of
// (elem[a], elem[c]) or `none` (sorted last) when the block is absent; value =
// the pair id e*n_base^2 + a*n_base + c.  A stable sort by key keeps the pairs
// of one block in element order.
__global__ void k_elem_targets(int ne, int n_base, const int* __restrict__ elem, const int* __restrict__ rowptr,
                               const int* __restrict__ item, unsigned long long* __restrict__ keys,
                               int* __restrict__ ids, int* __restrict__ bad, unsigned long long none) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = (long long)n_base * n_base;
  if (t >= (long long)ne * per) return;
  const int e = (int)(t / per);
  const int q = (int)(t % per);
  const int a = q / n_base, c = q % n_base;
  const int blk = find_block(rowptr, item, elem[(size_t)e * n_base + a], elem[(size_t)e * n_base + c]);
  if (blk < 0) atomicOr(bad, 1);
  keys[t] = blk >= 0 ? (unsigned long long)blk : none;
  ids[t] = (int)t;
}

// SELL position of block `blk`, component e (row found by bisection).
template <typename V>
__device__ __forceinline__ size_t sell_pos(int blk, int e, int bb, int n, const int* __restrict__ rowptr,
                                           const int* __restrict__ slice_ptr) {
  int lo = 0, hi = n - 1;  // first row r with rowptr[r + 1] > blk
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (rowptr[mid + 1] > blk) hi = mid;
    else lo = mid + 1;
  }
  const int r = lo;
  const size_t slot = (size_t)slice_ptr[r / kWave] + (size_t)(blk - rowptr[r]);
  return sell_vpos<V>(slot, e, bb, r % kWave);
}

// Thread (i, component): at the head i of each run of equal block keys, adds
// the component of every pair in the run, in order, onto the stored value.
template <typename V>
__global__ void k_elem_accumulate(long long m, const unsigned long long* __restrict__ keys,
                                  const int* __restrict__ ids, const V* __restrict__ Ke, int n_base, int b,
                                  int n, const int* __restrict__ rowptr, const int* __restrict__ slice_ptr,
                                  V* __restrict__ val, unsigned long long none) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int bb = b * b;
  if (t >= m * bb) return;
  const long long i = t / bb;
  const int ee = (int)(t % bb);
  const unsigned long long key = keys[i];
  if (key == none || (i > 0 && keys[i - 1] == key)) return;
  const int k1 = ee / b, k2 = ee % b;
  const long long per = (long long)n_base * n_base;
  const long long nzm = (long long)n_base * b;
  const size_t pos = sell_pos<V>((int)key, ee, bb, n, rowptr, slice_ptr);
  V acc = val[pos];
  for (long long j = i; j < m && keys[j] == key; ++j) {
    const long long pid = ids[j];
    const long long e = pid / per;
    const int q = (int)(pid % per);
    const int a = q / n_base, c = q % n_base;
    acc = acc + Ke[e * nzm * nzm + (long long)(a * b + k1) * nzm + (c * b + k2)];
  }
  val[pos] = acc;
}

template <typename V>
__global__ void k_bc_mark(int nbc, const int* __restrict__ node, const int* __restrict__ dof,
                          const V* __restrict__ v, int b, int* __restrict__ isbc, V* __restrict__ bcval,
                          int* __restrict__ dup) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbc) return;
  const int d = node[i] * b + dof[i];
  if (atomicAdd(&isbc[d], 1) != 0) atomicOr(dup, 1);
  bcval[d] = v[i];
}

__global__ void k_bc_check_diag(int nbc, const int* __restrict__ node, const int* __restrict__ rowptr,
                                const int* __restrict__ item, int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbc) return;
  if (find_block(rowptr, item, node[i], node[i]) < 0) atomicOr(bad, 1);
}

// val * A of a condition column, as monolis multiplies (no fused multiply-add).
__device__ __forceinline__ double bc_prod(double v, double a) {
#pragma clang fp contract(off)
  return v * a;
}
__device__ __forceinline__ zd bc_prod(zd v, zd a) {
#pragma clang fp contract(off)
  return zd(v.re * a.re - v.im * a.im, v.re * a.im + v.im * a.re);
}

// One thread per block row: rows of condition DOFs become unit rows with
// rhs = val; other rows move val * A(r, c) of condition columns to the
// right-hand side (blocks in column order, then component order) and zero them.
template <typename V>
__global__ void k_bc_apply(int n, int b, const int* __restrict__ slice_ptr, const int* __restrict__ rowptr,
                           const int* __restrict__ col, V* __restrict__ val, const int* __restrict__ isbc,
                           const V* __restrict__ bcval, V* __restrict__ rhs) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = r / kWave, lane = r % kWave;
  const int k0 = slice_ptr[s];
  const int len = rowptr[r + 1] - rowptr[r];
  const int bb = b * b;
  for (int k1 = 0; k1 < b; ++k1) {
    const bool rowbc = isbc[r * b + k1] != 0;
    V acc = rhs[(size_t)r * b + k1];
    for (int k = 0; k < len; ++k) {
      const size_t slot = (size_t)k0 + k;
      const int c = col[slot * kWave + lane];
      for (int k2 = 0; k2 < b; ++k2) {
        const bool colbc = isbc[c * b + k2] != 0;
        const size_t pos = sell_vpos<V>(slot, k1 * b + k2, bb, lane);
        if (colbc && !rowbc) {
          const V t = bc_prod(bcval[c * b + k2], val[pos]);
          acc = acc - t;  // B(n) = B(n) - val*A(m)  (spmat_handler_util.f90:636, :723)
        }
        if (rowbc || colbc) val[pos] = V((rowbc && c == r && k1 == k2) ? 1.0 : 0.0);
      }
    }
    rhs[(size_t)r * b + k1] = rowbc ? bcval[r * b + k1] : acc;
  }
}

template <typename V>
__global__ void k_sell_gather(int n, int b, const int* __restrict__ slice_ptr, const int* __restrict__ rowptr,
                              const V* __restrict__ val, V* __restrict__ A) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int s = r / kWave, lane = r % kWave;
  const int k0 = slice_ptr[s];
  const int j0 = rowptr[r], len = rowptr[r + 1] - j0;
  const int bb = b * b;
  for (int k = 0; k < len; ++k)
    for (int e = 0; e < bb; ++e)
      A[((size_t)j0 + k) * bb + e] = val[sell_vpos<V>((size_t)k0 + k, e, bb, lane)];
}

// ---- variable block sizes (monolis n_dof = -1): values in monolis's layout,
// block (i, j) at index2[e] with n_i x n_j row-major components ----------------
// Element e's matrix is (sum_a n(elem[e][a]))^2 row-major; sz[e] is its size.
__global__ void k_v_elem_sizes(int ne, int n_base, const int* __restrict__ elem, const int* __restrict__ list,
                               long long* __restrict__ sz) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  long long d = 0;
  for (int a = 0; a < n_base; ++a) d += list[elem[(size_t)e * n_base + a]];
  sz[e] = d * d;
}

// Block row of block position blk (bisection over the block CSR offsets).
__device__ __forceinline__ int block_row(int blk, int n, const int* __restrict__ index) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (index[mid + 1] > blk) hi = mid;
    else lo = mid + 1;
  }
  return lo;
}

// k_elem_accumulate for variable blocks: thread (i, component k1*maxb + k2) at
// the head i of a run of equal block keys adds that component of every pair in
// the run, in order, onto monolis's A (entry (k1, k2) of block (r, c) at
// index2[blk] + n_c*k1 + k2; element entry (off_a + k1, off_c + k2), off_a the
// sum of the block sizes of the element's nodes before a).
template <typename V>
__global__ void k_v_elem_accumulate(long long m, const unsigned long long* __restrict__ keys,
                                    const int* __restrict__ ids, const V* __restrict__ Ke,
                                    const long long* __restrict__ koff, const int* __restrict__ elem, int n_base,
                                    int maxb, int n, const int* __restrict__ list, const int* __restrict__ index,
                                    const int* __restrict__ item, const int* __restrict__ index2, V* __restrict__ A,
                                    unsigned long long none) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int mm = maxb * maxb;
  if (t >= m * mm) return;
  const long long i = t / mm;
  const int comp = (int)(t % mm);
  const unsigned long long key = keys[i];
  if (key == none || (i > 0 && keys[i - 1] == key)) return;
  const int blk = (int)key;
  const int r = block_row(blk, n, index), c = item[blk] - 1;
  const int n1 = list[r], n2 = list[c];
  const int k1 = comp / maxb, k2 = comp % maxb;
  if (k1 >= n1 || k2 >= n2) return;
  const long long pos = (long long)index2[blk] + (long long)n2 * k1 + k2;
  const long long per = (long long)n_base * n_base;
  V acc = A[pos];
  for (long long j = i; j < m && keys[j] == key; ++j) {
    const long long pid = ids[j];
    const long long e = pid / per;
    const int q = (int)(pid % per);
    const int a = q / n_base, cc = q % n_base;
    long long S = 0, ro = 0, co = 0;
    for (int u = 0; u < n_base; ++u) {
      const int d = list[elem[e * n_base + u]];
      if (u < a) ro += d;
      if (u < cc) co += d;
      S += d;
    }
    acc = acc + Ke[koff[e] + (ro + k1) * S + co + k2];
  }
  A[pos] = acc;
}

template <typename V>
__global__ void k_v_bc_mark(int nbc, const int* __restrict__ node, const int* __restrict__ dof,
                            const V* __restrict__ v, const int* __restrict__ dof_index, const int* __restrict__ list,
                            int* __restrict__ isbc, V* __restrict__ bcval, int* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbc) return;
  if (dof[i] >= list[node[i]]) {  // monolis_stop_by_submatrix_access
    atomicOr(flags + 2, 1);
    return;
  }
  const int d = dof_index[node[i]] + dof[i];
  if (atomicAdd(&isbc[d], 1) != 0) atomicOr(flags + 1, 1);
  bcval[d] = v[i];
}

// k_bc_apply on monolis's variable-block layout (one thread per block row).
template <typename V>
__global__ void k_v_bc_apply(int n, const int* __restrict__ list, const int* __restrict__ dof_index,
                             const int* __restrict__ index, const int* __restrict__ item,
                             const int* __restrict__ index2, V* __restrict__ A, const int* __restrict__ isbc,
                             const V* __restrict__ bcval, V* __restrict__ rhs) {
#pragma clang fp contract(off)
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int n1 = list[r];
  for (int k1 = 0; k1 < n1; ++k1) {
    const int d = dof_index[r] + k1;
    const bool rowbc = isbc[d] != 0;
    V acc = rhs[d];
    for (int e = index[r]; e < index[r + 1]; ++e) {
      const int c = item[e] - 1;
      const int n2 = list[c];
      for (int k2 = 0; k2 < n2; ++k2) {
        const int dc = dof_index[c] + k2;
        const bool colbc = isbc[dc] != 0;
        const long long pos = (long long)index2[e] + (long long)n2 * k1 + k2;
        if (colbc && !rowbc) {
          const V t = bc_prod(bcval[dc], A[pos]);
          acc = acc - t;
        }
        if (rowbc || colbc) A[pos] = V((rowbc && c == r && k1 == k2) ? 1.0 : 0.0);
      }
    }
    rhs[d] = rowbc ? bcval[d] : acc;
  }
}

int require_assemblable(const Matrix& M, const char* what, bool cplx) {
  if (M.vb || M.is_complex != cplx) {
    set_error("%s: %s fixed-block matrices only", what, cplx ? "complex" : "real");
    return -2;
  }
  if (M.np != M.n) {
    set_error("%s: square local matrices only (n == np)", what);
    return -2;
  }
  return 0;
}

void mark_values_changed(Matrix* M) {
  M->diag_ready = M->diag32_ready = M->diagz_ready = false;
  M->val32_ready = false;  // the FP32 copy is rebuilt on its next use
  M->transpose_stale = true;
}

template <typename V>
DevArray<V>& vb_values(VBlock& Vb);
template <>
DevArray<double>& vb_values<double>(VBlock& Vb) {
  return Vb.A64;
}
template <>
DevArray<zd>& vb_values<zd>(VBlock& Vb) {
  return Vb.Az;
}

// Re-expands a variable-block matrix from its kept monolis-layout values.
int vb_reexpand(Matrix* M, const double* A, hipStream_t s) { return matrix_set_values_vblock(M, A, s); }
int vb_reexpand(Matrix* M, const zd* A, hipStream_t s) { return matrix_set_values_vblock_complex(M, A, s); }

int require_vblock(const Matrix& M, const char* what, bool cplx) {
  if (!M.vb || M.is_complex != cplx) {
    set_error("%s: %s matrices only", what, cplx ? "complex" : "real");
    return -2;
  }
  if (M.vb->npb != M.vb->nb) {
    set_error("%s: square local matrices only (n == np)", what);
    return -2;
  }
  return 0;
}

template <typename V>
V* values_of(Matrix* M);
template <>
double* values_of<double>(Matrix* M) {
  return M->val64.p;
}
template <>
zd* values_of<zd>(Matrix* M) {
  return M->valz.p;
}

}  // namespace

int pattern_by_simple_mesh(int n_node, int n_base, int n_elem, const int* elem, DevArray<int>& index,
                           DevArray<int>& item, int* nz_out, hipStream_t s) {
  if (n_node < 1 || n_base < 1 || n_elem < 0) {
    set_error("invalid mesh sizes n_node=%d n_base=%d n_elem=%d", n_node, n_base, n_elem);
    return -2;
  }
  const long long nelem_ids = (long long)n_elem * n_base;
  const long long npair = (long long)n_elem * n_base * (n_base - 1);
  const long long K = npair + n_node;
  if (K >= (1LL << 31)) {
    set_error("mesh too large for one pattern build (%lld node pairs)", K);
    return -2;
  }
  DevArray<int> etmp;
  const int* de = nullptr;
  MH_TRY(stage(elem, (size_t)nelem_ids, etmp, &de, s));
  DevArray<int> flag;
  MH_TRY(flag.alloc(1));
  MH_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
  if (nelem_ids > 0) {
    k_check_range<<<std::min(grid_for(nelem_ids), 65536), kBlock, 0, s>>>(nelem_ids, de, n_node, flag.p);
    MH_HIP(hipGetLastError());
  }
  int bad = 0;
  MH_TRY(read_flag(flag.p, &bad, s));
  if (bad) {
    set_error("element connectivity holds a node id outside 0..%d", n_node - 1);
    return -2;
  }
  DevArray<unsigned long long> k0, k1;
  MH_TRY(k0.alloc((size_t)K));
  MH_TRY(k1.alloc((size_t)K));
  k_mesh_keys<<<grid_for(K), kBlock, 0, s>>>(n_node, n_base, n_elem, de, k0.p);
  MH_HIP(hipGetLastError());
  const int end_bit = 32 + bits_for((unsigned long long)n_node);
  size_t tmp_bytes = 0, tmp2 = 0;
  MH_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, k0.p, k1.p, (int)K, 0, end_bit, s));
  MH_HIP(hipcub::DeviceSelect::Unique(nullptr, tmp2, k1.p, k0.p, flag.p, (int)K, s));
  DevArray<unsigned char> tmp;
  MH_TRY(tmp.alloc(std::max(tmp_bytes, tmp2)));
  MH_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tmp_bytes, k0.p, k1.p, (int)K, 0, end_bit, s));
  MH_HIP(hipcub::DeviceSelect::Unique(tmp.p, tmp2, k1.p, k0.p, flag.p, (int)K, s));
  int nz = 0;
  MH_TRY(read_flag(flag.p, &nz, s));
  MH_TRY(index.alloc((size_t)n_node + 1));
  MH_TRY(item.alloc((size_t)nz));
  k_pattern_csr<<<grid_for(std::max<long long>(nz, (long long)n_node + 1)), kBlock, 0, s>>>(n_node, nz, k0.p,
                                                                                            index.p, item.p);
  MH_HIP(hipGetLastError());
  MH_HIP(hipStreamSynchronize(s));
  *nz_out = nz;
  return 0;
}

template <typename V>
static int add_elements_T(Matrix* M, int n_elem, int n_base, const int* elem, const V* Ke, hipStream_t s) {
  MH_TRY(require_assemblable(*M, "element assembly", sizeof(V) == sizeof(zd)));
  if (n_elem < 0 || n_base < 1) {
    set_error("invalid element sizes n_elem=%d n_base=%d", n_elem, n_base);
    return -2;
  }
  if (n_elem == 0 || M->n == 0) return 0;
  const int b = M->b, bb = b * b;
  const long long nzm = (long long)n_base * b;
  const long long per_elem = nzm * nzm;       // values per element matrix
  const long long pairs_elem = (long long)n_base * n_base;
  if (per_elem > (1LL << 30)) {
    set_error("element matrix too large (%lld entries)", per_elem);
    return -2;
  }
  // node pairs per chunk: bounded temporaries, element order kept across chunks
  // (MONOLIS_HIP_ASSEMBLY_CHUNK_PAIRS overrides the 2^24 default)
  long long max_pairs = 1LL << 24;
  if (const char* env = getenv("MONOLIS_HIP_ASSEMBLY_CHUNK_PAIRS")) {
    const long long v = atoll(env);
    if (v > 0) max_pairs = v;
  }
  const int chunk = (int)std::max<long long>(1, std::min<long long>(n_elem, max_pairs / pairs_elem));
  DevArray<int> etmp;
  const int* de = nullptr;
  MH_TRY(stage(elem, (size_t)n_elem * n_base, etmp, &de, s));
  DevArray<int> flag;
  MH_TRY(flag.alloc(1));
  MH_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
  k_check_range<<<std::min(grid_for((long long)n_elem * n_base), 65536), kBlock, 0, s>>>((long long)n_elem * n_base,
                                                                                        de, M->n, flag.p);
  MH_HIP(hipGetLastError());
  int bad = 0;
  MH_TRY(read_flag(flag.p, &bad, s));
  if (bad) {
    set_error("element connectivity holds a node id outside 0..%d", M->n - 1);
    return -2;
  }
  const bool ke_dev = is_device_pointer(Ke);
  const long long cap = (long long)chunk * pairs_elem;
  DevArray<unsigned long long> k0, k1;
  DevArray<int> i0, i1;
  DevArray<V> kt;
  MH_TRY(k0.alloc((size_t)cap));
  MH_TRY(k1.alloc((size_t)cap));
  MH_TRY(i0.alloc((size_t)cap));
  MH_TRY(i1.alloc((size_t)cap));
  if (!ke_dev) MH_TRY(kt.alloc((size_t)chunk * per_elem));
  // keys are block positions < nz; `none` (an absent block) sorts last
  const int end_bit = bits_for((unsigned long long)M->nz) + 1;
  const unsigned long long none = 1ull << (end_bit - 1);
  size_t tmp_bytes = 0;
  MH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0.p, k1.p, i0.p, i1.p, (int)cap, 0, end_bit, s));
  DevArray<unsigned char> tmp;
  MH_TRY(tmp.alloc(tmp_bytes));
  for (int e0 = 0; e0 < n_elem; e0 += chunk) {
    const int ne = std::min(chunk, n_elem - e0);
    const long long m = (long long)ne * pairs_elem;
    const V* kc = nullptr;
    if (ke_dev) {
      kc = Ke + (size_t)e0 * per_elem;
    } else {
      MH_TRY(upload(kt.p, Ke + (size_t)e0 * per_elem, (size_t)ne * per_elem * sizeof(V), s));
      kc = kt.p;
    }
    k_elem_targets<<<grid_for(m), kBlock, 0, s>>>(ne, n_base, de + (size_t)e0 * n_base, M->rowptr.p, M->csr_item.p,
                                                   k0.p, i0.p, flag.p, none);
    MH_HIP(hipGetLastError());
    MH_TRY(read_flag(flag.p, &bad, s));
    if (bad) {
      // monolis stops here (monolis_stop_by_matrix_assemble); earlier chunks stay added
      set_error("element %d..%d couples nodes whose block is not in the nonzero pattern", e0, e0 + ne - 1);
      mark_values_changed(M);
      return -2;
    }
    size_t tb = tmp_bytes;
    MH_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, k0.p, k1.p, i0.p, i1.p, (int)m, 0, end_bit, s));
    k_elem_accumulate<V><<<grid_for(m * bb), kBlock, 0, s>>>(m, k1.p, i1.p, kc, n_base, b, M->n, M->rowptr.p,
                                                              M->slice_ptr.p, values_of<V>(M), none);
    MH_HIP(hipGetLastError());
  }
  MH_HIP(hipStreamSynchronize(s));
  mark_values_changed(M);
  return 0;
}

// Element assembly on a variable-block matrix: monolis's layout values kept in
// the VBlock are updated (sorted node pairs, sums in element order, as for fixed
// blocks), then the scalar expansion and the diag_V factors are rebuilt.
template <typename V>
static int add_elements_V_T(Matrix* M, int n_elem, int n_base, const int* elem, const V* Ke, hipStream_t s) {
  MH_TRY(require_vblock(*M, "element assembly", sizeof(V) == sizeof(zd)));
  VBlock& Vb = *M->vb;
  if (n_elem < 0 || n_base < 1) {
    set_error("invalid element sizes n_elem=%d n_base=%d", n_elem, n_base);
    return -2;
  }
  if (n_elem == 0 || Vb.nb == 0) return 0;
  const int n = Vb.nb;
  const long long pairs_elem = (long long)n_base * n_base;
  long long max_pairs = 1LL << 24;
  if (const char* env = getenv("MONOLIS_HIP_ASSEMBLY_CHUNK_PAIRS")) {
    const long long v = atoll(env);
    if (v > 0) max_pairs = v;
  }
  const int chunk = (int)std::max<long long>(1, std::min<long long>(n_elem, max_pairs / pairs_elem));
  DevArray<int> etmp;
  const int* de = nullptr;
  MH_TRY(stage(elem, (size_t)n_elem * n_base, etmp, &de, s));
  DevArray<int> flag;
  MH_TRY(flag.alloc(1));
  MH_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), s));
  k_check_range<<<std::min(grid_for((long long)n_elem * n_base), 65536), kBlock, 0, s>>>((long long)n_elem * n_base,
                                                                                        de, n, flag.p);
  MH_HIP(hipGetLastError());
  int bad = 0;
  MH_TRY(read_flag(flag.p, &bad, s));
  if (bad) {
    set_error("element connectivity holds a node id outside 0..%d", n - 1);
    return -2;
  }
  // element offsets into Ke: exclusive sum of (sum of the nodes' block sizes)^2
  DevArray<long long> sz, koff;
  MH_TRY(sz.alloc((size_t)n_elem + 1));
  MH_TRY(koff.alloc((size_t)n_elem + 1));
  MH_HIP(hipMemsetAsync(sz.p + n_elem, 0, sizeof(long long), s));
  k_v_elem_sizes<<<grid_for(n_elem), kBlock, 0, s>>>(n_elem, n_base, de, Vb.list.p, sz.p);
  MH_HIP(hipGetLastError());
  size_t scan_bytes = 0;
  MH_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, sz.p, koff.p, n_elem + 1, s));
  DevArray<unsigned char> scan_tmp;
  MH_TRY(scan_tmp.alloc(scan_bytes ? scan_bytes : 1));
  MH_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, scan_bytes, sz.p, koff.p, n_elem + 1, s));
  long long total = 0;
  MH_HIP(hipMemcpyAsync(&total, koff.p + n_elem, sizeof(long long), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  DevArray<V> kt;
  const V* dk = Ke;
  if (!is_device_pointer(Ke)) {
    MH_TRY(kt.alloc((size_t)total + 1));
    MH_TRY(upload(kt.p, Ke, (size_t)total * sizeof(V), s));
    dk = kt.p;
  }
  const long long cap = (long long)chunk * pairs_elem;
  DevArray<unsigned long long> k0, k1;
  DevArray<int> i0, i1;
  MH_TRY(k0.alloc((size_t)cap));
  MH_TRY(k1.alloc((size_t)cap));
  MH_TRY(i0.alloc((size_t)cap));
  MH_TRY(i1.alloc((size_t)cap));
  const int end_bit = bits_for((unsigned long long)Vb.nzb) + 1;
  const unsigned long long none = 1ull << (end_bit - 1);
  size_t tmp_bytes = 0;
  MH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0.p, k1.p, i0.p, i1.p, (int)cap, 0, end_bit, s));
  DevArray<unsigned char> tmp;
  MH_TRY(tmp.alloc(tmp_bytes ? tmp_bytes : 1));
  V* A = vb_values<V>(Vb).p;
  const int mm = Vb.maxb * Vb.maxb;
  for (int e0 = 0; e0 < n_elem; e0 += chunk) {
    const int ne = std::min(chunk, n_elem - e0);
    const long long m = (long long)ne * pairs_elem;
    k_elem_targets<<<grid_for(m), kBlock, 0, s>>>(ne, n_base, de + (size_t)e0 * n_base, Vb.index.p, Vb.item.p, k0.p,
                                                   i0.p, flag.p, none);
    MH_HIP(hipGetLastError());
    MH_TRY(read_flag(flag.p, &bad, s));
    if (bad) {
      set_error("element %d..%d couples nodes whose block is not in the nonzero pattern", e0, e0 + ne - 1);
      MH_TRY(vb_reexpand(M, A, s));  // earlier chunks stay added, as monolis stops there
      mark_values_changed(M);
      return -2;
    }
    size_t tb = tmp_bytes;
    MH_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, k0.p, k1.p, i0.p, i1.p, (int)m, 0, end_bit, s));
    k_v_elem_accumulate<V><<<grid_for(m * mm), kBlock, 0, s>>>(m, k1.p, i1.p, dk, koff.p + e0,
                                                                de + (size_t)e0 * n_base, n_base, Vb.maxb, n,
                                                                Vb.list.p, Vb.index.p, Vb.item.p, Vb.index2.p, A,
                                                                none);
    MH_HIP(hipGetLastError());
  }
  MH_TRY(vb_reexpand(M, A, s));
  mark_values_changed(M);
  return 0;
}

int matrix_add_elements(Matrix* M, int n_elem, int n_base, const int* elem, const double* Ke, hipStream_t s) {
  if (M->vb) return add_elements_V_T<double>(M, n_elem, n_base, elem, Ke, s);
  return add_elements_T<double>(M, n_elem, n_base, elem, Ke, s);
}
int matrix_add_elements_C(Matrix* M, int n_elem, int n_base, const int* elem, const zd* Ke, hipStream_t s) {
  if (M->vb) return add_elements_V_T<zd>(M, n_elem, n_base, elem, Ke, s);
  return add_elements_T<zd>(M, n_elem, n_base, elem, Ke, s);
}

template <typename V>
static int dirichlet_T(Matrix* M, V* rhs, int n_bc, const int* node, const int* dof, const V* val, hipStream_t s) {
  MH_TRY(require_assemblable(*M, "Dirichlet conditions", sizeof(V) == sizeof(zd)));
  if (n_bc < 0) {
    set_error("negative number of conditions");
    return -2;
  }
  if (n_bc == 0 || M->n == 0) return 0;
  const int b = M->b;
  DevArray<int> nt, dt;
  DevArray<V> vt;
  const int *dn = nullptr, *dd = nullptr;
  const V* dv = nullptr;
  MH_TRY(stage(node, (size_t)n_bc, nt, &dn, s));
  MH_TRY(stage(dof, (size_t)n_bc, dt, &dd, s));
  MH_TRY(stage(val, (size_t)n_bc, vt, &dv, s));
  DevArray<int> flag;
  MH_TRY(flag.alloc(2));
  MH_HIP(hipMemsetAsync(flag.p, 0, 2 * sizeof(int), s));
  k_check_range<<<std::min(grid_for(n_bc), 65536), kBlock, 0, s>>>(n_bc, dn, M->n, flag.p);
  k_check_range<<<std::min(grid_for(n_bc), 65536), kBlock, 0, s>>>(n_bc, dd, b, flag.p + 1);
  MH_HIP(hipGetLastError());
  int bad[2] = {0, 0};
  MH_HIP(hipMemcpyAsync(bad, flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  if (bad[0] || bad[1]) {
    set_error("Dirichlet condition outside the matrix (%s)", bad[0] ? "node id" : "DOF index");
    return -2;
  }
  MH_HIP(hipMemsetAsync(flag.p, 0, 2 * sizeof(int), s));
  k_bc_check_diag<<<grid_for(n_bc), kBlock, 0, s>>>(n_bc, dn, M->rowptr.p, M->csr_item.p, flag.p);
  MH_HIP(hipGetLastError());
  DevArray<int> isbc;
  DevArray<V> bcval;
  const size_t m = (size_t)M->n * b;
  MH_TRY(isbc.alloc(m));
  MH_TRY(bcval.alloc(m));
  MH_HIP(hipMemsetAsync(isbc.p, 0, m * sizeof(int), s));
  k_bc_mark<V><<<grid_for(n_bc), kBlock, 0, s>>>(n_bc, dn, dd, dv, b, isbc.p, bcval.p, flag.p + 1);
  MH_HIP(hipGetLastError());
  MH_HIP(hipMemcpyAsync(bad, flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  if (bad[0]) {
    // monolis stops (monolis_stop_by_set_DBC)
    set_error("Dirichlet condition on a node without a diagonal block");
    return -2;
  }
  if (bad[1]) {
    set_error("the same DOF is listed twice among the Dirichlet conditions");
    return -2;
  }
  DevArray<V> rtmp;
  V* dr = rhs;
  const bool rhs_dev = is_device_pointer(rhs);
  if (!rhs_dev) {
    MH_TRY(rtmp.alloc(m));
    MH_TRY(upload(rtmp.p, rhs, m * sizeof(V), s));
    dr = rtmp.p;
  }
  k_bc_apply<V><<<grid_for(M->n), kBlock, 0, s>>>(M->n, b, M->slice_ptr.p, M->rowptr.p, M->col.p, values_of<V>(M),
                                                  isbc.p, bcval.p, dr);
  MH_HIP(hipGetLastError());
  if (!rhs_dev) MH_HIP(hipMemcpyAsync(rhs, dr, m * sizeof(V), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  mark_values_changed(M);
  return 0;
}

// Dirichlet conditions on a variable-block matrix (monolis's layout values kept
// in the VBlock, then the expansion and the diag_V factors are rebuilt).
template <typename V>
static int dirichlet_V_T(Matrix* M, V* rhs, int n_bc, const int* node, const int* dof, const V* val, hipStream_t s) {
  MH_TRY(require_vblock(*M, "Dirichlet conditions", sizeof(V) == sizeof(zd)));
  VBlock& Vb = *M->vb;
  if (n_bc < 0) {
    set_error("negative number of conditions");
    return -2;
  }
  if (n_bc == 0 || Vb.nb == 0) return 0;
  DevArray<int> nt, dt;
  DevArray<V> vt;
  const int *dn = nullptr, *dd = nullptr;
  const V* dv = nullptr;
  MH_TRY(stage(node, (size_t)n_bc, nt, &dn, s));
  MH_TRY(stage(dof, (size_t)n_bc, dt, &dd, s));
  MH_TRY(stage(val, (size_t)n_bc, vt, &dv, s));
  DevArray<int> flag;
  MH_TRY(flag.alloc(3));
  MH_HIP(hipMemsetAsync(flag.p, 0, 3 * sizeof(int), s));
  k_check_range<<<std::min(grid_for(n_bc), 65536), kBlock, 0, s>>>(n_bc, dn, Vb.nb, flag.p);
  k_check_range<<<std::min(grid_for(n_bc), 65536), kBlock, 0, s>>>(n_bc, dd, Vb.maxb, flag.p + 1);
  MH_HIP(hipGetLastError());
  int bad[3] = {0, 0, 0};
  MH_HIP(hipMemcpyAsync(bad, flag.p, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  if (bad[0] || bad[1]) {
    set_error("Dirichlet condition outside the matrix (%s)", bad[0] ? "node id" : "DOF index");
    return -2;
  }
  MH_HIP(hipMemsetAsync(flag.p, 0, 3 * sizeof(int), s));
  k_bc_check_diag<<<grid_for(n_bc), kBlock, 0, s>>>(n_bc, dn, Vb.index.p, Vb.item.p, flag.p);
  MH_HIP(hipGetLastError());
  const size_t m = (size_t)M->n;  // scalar rows = DOFs
  DevArray<int> isbc;
  DevArray<V> bcval;
  MH_TRY(isbc.alloc(m));
  MH_TRY(bcval.alloc(m));
  MH_HIP(hipMemsetAsync(isbc.p, 0, m * sizeof(int), s));
  k_v_bc_mark<V><<<grid_for(n_bc), kBlock, 0, s>>>(n_bc, dn, dd, dv, Vb.dof_index.p, Vb.list.p, isbc.p, bcval.p,
                                                   flag.p);
  MH_HIP(hipGetLastError());
  MH_HIP(hipMemcpyAsync(bad, flag.p, 3 * sizeof(int), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  if (bad[0]) {
    set_error("Dirichlet condition on a node without a diagonal block");
    return -2;
  }
  if (bad[2]) {  // monolis_stop_by_submatrix_access
    set_error("Dirichlet condition on a DOF beyond its node's block size");
    return -2;
  }
  if (bad[1]) {
    set_error("the same DOF is listed twice among the Dirichlet conditions");
    return -2;
  }
  DevArray<V> rtmp;
  V* dr = rhs;
  const bool rhs_dev = is_device_pointer(rhs);
  if (!rhs_dev) {
    MH_TRY(rtmp.alloc(m));
    MH_TRY(upload(rtmp.p, rhs, m * sizeof(V), s));
    dr = rtmp.p;
  }
  V* A = vb_values<V>(Vb).p;
  k_v_bc_apply<V><<<grid_for(Vb.nb), kBlock, 0, s>>>(Vb.nb, Vb.list.p, Vb.dof_index.p, Vb.index.p, Vb.item.p,
                                                     Vb.index2.p, A, isbc.p, bcval.p, dr);
  MH_HIP(hipGetLastError());
  if (!rhs_dev) MH_HIP(hipMemcpyAsync(rhs, dr, m * sizeof(V), hipMemcpyDeviceToHost, s));
  MH_TRY(vb_reexpand(M, A, s));
  mark_values_changed(M);
  return 0;
}

int matrix_dirichlet(Matrix* M, double* rhs, int n_bc, const int* node, const int* dof, const double* val,
                     hipStream_t s) {
  if (M->vb) return dirichlet_V_T<double>(M, rhs, n_bc, node, dof, val, s);
  return dirichlet_T<double>(M, rhs, n_bc, node, dof, val, s);
}
int matrix_dirichlet_C(Matrix* M, zd* rhs, int n_bc, const int* node, const int* dof, const zd* val, hipStream_t s) {
  if (M->vb) return dirichlet_V_T<zd>(M, rhs, n_bc, node, dof, val, s);
  return dirichlet_T<zd>(M, rhs, n_bc, node, dof, val, s);
}

template <typename V>
static int get_values_T(const Matrix& M, V* A, hipStream_t s) {
  if (M.vb && M.is_complex == (sizeof(V) == sizeof(zd))) {  // monolis's layout, kept in the VBlock
    const size_t nv = (size_t)M.vb->nzs;
    if (nv) MH_HIP(hipMemcpyAsync(A, vb_values<V>(*M.vb).p, nv * sizeof(V), hipMemcpyDefault, s));
    MH_HIP(hipStreamSynchronize(s));
    return 0;
  }
  if (M.vb || M.is_complex != (sizeof(V) == sizeof(zd))) {
    set_error("value read-back: %s fixed-block matrices only", sizeof(V) == sizeof(zd) ? "complex" : "real");
    return -2;
  }
  const size_t nv = (size_t)M.nz * M.b * M.b;
  if (nv == 0) return 0;
  DevArray<V> tmp;
  V* dA = A;
  const bool dev = is_device_pointer(A);
  if (!dev) {
    MH_TRY(tmp.alloc(nv));
    dA = tmp.p;
    if (M.n < M.np) MH_HIP(hipMemsetAsync(dA, 0, nv * sizeof(V), s));
  }
  if (M.n > 0) {
    k_sell_gather<V><<<grid_for(M.n), kBlock, 0, s>>>(M.n, M.b, M.slice_ptr.p, M.rowptr.p,
                                                      values_of<V>(const_cast<Matrix*>(&M)), dA);
    MH_HIP(hipGetLastError());
  }
  if (!dev) MH_HIP(hipMemcpyAsync(A, dA, nv * sizeof(V), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  return 0;
}

int matrix_get_values(const Matrix& M, double* A, hipStream_t s) { return get_values_T<double>(M, A, s); }
int matrix_get_values_C(const Matrix& M, zd* A, hipStream_t s) { return get_values_T<zd>(M, A, s); }

int matrix_get_pattern(const Matrix& M, int* index, int* item, hipStream_t s) {
  if (M.vb) {  // the block pattern, not the scalar expansion's
    const VBlock& V = *M.vb;
    if (index) MH_HIP(hipMemcpyAsync(index, V.index.p, sizeof(int) * ((size_t)V.npb + 1), hipMemcpyDefault, s));
    if (item && V.nzb > 0) MH_HIP(hipMemcpyAsync(item, V.item.p, sizeof(int) * (size_t)V.nzb, hipMemcpyDefault, s));
    MH_HIP(hipStreamSynchronize(s));
    return 0;
  }
  if (index) MH_HIP(hipMemcpyAsync(index, M.rowptr.p, sizeof(int) * ((size_t)M.np + 1), hipMemcpyDefault, s));
  if (item && M.nz > 0) MH_HIP(hipMemcpyAsync(item, M.csr_item.p, sizeof(int) * (size_t)M.nz, hipMemcpyDefault, s));
  MH_HIP(hipStreamSynchronize(s));
  return 0;
}

}  // namespace mh
