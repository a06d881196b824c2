// monolis's user-facing C API (include/monolis.h) over the GPU path.
//
// Matrices are kept the way monolis keeps them: host arrays in MONOLIS_MAT
// (CSR index 0-based, item 1-based, values row-major per block, n_dof_index /
// n_dof_index2 prefix sums, a CSC view for Dirichlet conditions).  Pattern
// building and assembly are host bookkeeping, as in monolis; solves, products
// and inner products call the flat entry points of monolis_hip.h, which run on
// the GPU and fail loudly without one.
#include <algorithm>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/monolis.h"
#include "../../include/monolis_hip.h"

namespace {

using zc = std::complex<double>;
static_assert(sizeof(zc) == sizeof(double _Complex), "complex layout");

[[noreturn]] void stop(const char* what) {
  fprintf(stderr, "** monolis error: %s\n", what);
  fflush(stderr);
  exit(EXIT_FAILURE);
}

template <typename T>
T* alloc_zero(size_t n) {
  void* p = calloc(n > 0 ? n : 1, sizeof(T));
  if (!p) stop("host allocation failed");
  return static_cast<T*>(p);
}

template <typename T>
void release(T*& p) {
  free(p);
  p = nullptr;
}

zc* zptr(double _Complex* p) { return reinterpret_cast<zc*>(p); }

// ---------------------------------------------------------------------------
// Pattern construction (spmat_nzpattern_util.f90:13-90)
// ---------------------------------------------------------------------------
// Rows of the nodal graph with the diagonal: row i = {i} + graph row i,
// ascending (the graph rows are taken as they are, as monolis does).
void set_csr_from_graph(MONOLIS_MAT* M, int n_node, const std::vector<int>& gindex, const std::vector<int>& gitem) {
  release(M->CSR.index);
  release(M->CSR.item);
  M->CSR.index = alloc_zero<int>((size_t)n_node + 1);
  for (int i = 0; i < n_node; ++i) M->CSR.index[i + 1] = gindex[i + 1] + i + 1;
  const int nz = M->CSR.index[n_node];
  M->CSR.item = alloc_zero<int>((size_t)nz);
  for (int i = 0; i < n_node; ++i) {
    int* row = M->CSR.item + M->CSR.index[i];
    row[0] = i + 1;
    int k = 1;
    for (int e = gindex[i]; e < gindex[i + 1]; ++e) row[k++] = gitem[e] + 1;
    std::sort(row, row + k);
  }
}

// Nodal graph of elements given as node lists (gedatsu's conversion as
// monolis_get_nonzero_pattern_by_simple_mesh_R calls it): nodes sharing an
// element are adjacent, each neighbour once, the node itself excluded.
void graph_from_elements(int n_node, int n_elem, const int* eindex, const int* enodes, std::vector<int>& gindex,
                         std::vector<int>& gitem) {
  std::vector<long long> pairs;
  for (int e = 0; e < n_elem; ++e)
    for (int a = eindex[e]; a < eindex[e + 1]; ++a)
      for (int c = eindex[e]; c < eindex[e + 1]; ++c) {
        const int na = enodes[a], nc = enodes[c];
        if (na < 0 || na >= n_node || nc < 0 || nc >= n_node) stop("element node id outside 0..n_node-1");
        if (na != nc) pairs.push_back((long long)na * n_node + nc);
      }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  gindex.assign((size_t)n_node + 1, 0);
  gitem.resize(pairs.size());
  for (size_t k = 0; k < pairs.size(); ++k) {
    ++gindex[(size_t)(pairs[k] / n_node) + 1];
    gitem[k] = (int)(pairs[k] % n_node);
  }
  for (int i = 0; i < n_node; ++i) gindex[i + 1] += gindex[i];
}

// monolis_get_CSC_format (spmat_nzpattern_util.f90): rows of every column in
// ascending order, with the CSR position of each entry (both 1-based).
void set_csc(MONOLIS_MAT* M) {
  const int np = M->NP;
  const int nz = M->CSR.index[np];
  release(M->CSC.index);
  release(M->CSC.item);
  release(M->CSC.perm);
  M->CSC.index = alloc_zero<int>((size_t)np + 1);
  M->CSC.item = alloc_zero<int>((size_t)nz);
  M->CSC.perm = alloc_zero<int>((size_t)nz);
  for (int e = 0; e < nz; ++e) ++M->CSC.index[M->CSR.item[e]];
  for (int c = 0; c < np; ++c) M->CSC.index[c + 1] += M->CSC.index[c];
  std::vector<int> fill(M->CSC.index, M->CSC.index + np);
  for (int i = 0; i < np; ++i)
    for (int e = M->CSR.index[i]; e < M->CSR.index[i + 1]; ++e) {
      const int c = M->CSR.item[e] - 1;
      const int p = fill[c]++;
      M->CSC.item[p] = i + 1;
      M->CSC.perm[p] = e + 1;
    }
}

// monolis_set_n_dof_index (spmat_nzpattern_util.f90:62-90), then zeroed values
// and vectors of the pattern's size.
void set_dofs_and_values(MONOLIS_MAT* M, const int* n_dof_list, int n_dof, bool cplx) {
  const int np = M->NP;
  const int nz = M->CSR.index[np];
  release(M->n_dof_list);
  release(M->n_dof_index);
  release(M->n_dof_index2);
  M->n_dof_list = alloc_zero<int>((size_t)np);
  M->n_dof_index = alloc_zero<int>((size_t)np + 1);
  M->n_dof_index2 = alloc_zero<int>((size_t)nz + 1);
  for (int i = 0; i < np; ++i) {
    M->n_dof_list[i] = n_dof_list ? n_dof_list[i] : n_dof;
    if (M->n_dof_list[i] < 1) stop("block sizes must be >= 1");
    M->n_dof_index[i + 1] = M->n_dof_index[i] + M->n_dof_list[i];
  }
  for (int i = 0; i < np; ++i)
    for (int e = M->CSR.index[i]; e < M->CSR.index[i + 1]; ++e) {
      const long long add = (long long)M->n_dof_list[i] * M->n_dof_list[M->CSR.item[e] - 1];
      const long long v = M->n_dof_index2[e] + add;
      if (v >= (1LL << 31)) stop("matrix too large for 32-bit entry offsets");
      M->n_dof_index2[e + 1] = (int)v;
    }
  const size_t nza = (size_t)M->n_dof_index2[nz];
  const size_t vec = (size_t)M->n_dof_index[np];
  monolis_mat_finalize_val_R(&M->R);
  monolis_mat_finalize_val_C(&M->C);
  if (cplx) {
    M->C.A = reinterpret_cast<double _Complex*>(alloc_zero<zc>(nza));
    M->C.X = reinterpret_cast<double _Complex*>(alloc_zero<zc>(vec));
    M->C.B = reinterpret_cast<double _Complex*>(alloc_zero<zc>(vec));
  } else {
    M->R.A = alloc_zero<double>(nza);
    M->R.X = alloc_zero<double>(vec);
    M->R.B = alloc_zero<double>(vec);
  }
}

void pattern(MONOLIS* mat, int n_node, const std::vector<int>& gindex, const std::vector<int>& gitem,
             const int* n_dof_list, int n_dof, bool cplx) {
  if (n_node < 0) stop("n_node < 0");
  MONOLIS_MAT* M = &mat->mat;
  M->N = n_node;
  M->NP = n_node;
  M->NDOF = n_dof_list ? -1 : n_dof;
  set_csr_from_graph(M, n_node, gindex, gitem);
  set_csc(M);
  set_dofs_and_values(M, n_dof_list, n_dof, cplx);
}

void pattern_simple_mesh(MONOLIS* mat, int n_node, int n_base, const int* n_dof_list, int n_dof, int n_elem,
                         int** elem, bool cplx) {
  std::vector<int> eindex((size_t)n_elem + 1), enodes((size_t)n_elem * n_base);
  for (int e = 0; e < n_elem; ++e) {
    eindex[e + 1] = (e + 1) * n_base;
    for (int k = 0; k < n_base; ++k) enodes[(size_t)e * n_base + k] = elem[e][k];
  }
  std::vector<int> gi, gt;
  graph_from_elements(n_node, n_elem, eindex.data(), enodes.data(), gi, gt);
  pattern(mat, n_node, gi, gt, n_dof_list, n_dof, cplx);
}

void pattern_connectivity(MONOLIS* mat, int n_node, int n_dof, int n_elem, const int* conn_index,
                          const int* conn_item, bool cplx) {
  std::vector<int> gi, gt;
  graph_from_elements(n_node, n_elem, conn_index, conn_item, gi, gt);
  pattern(mat, n_node, gi, gt, nullptr, n_dof, cplx);
}

void pattern_graph(MONOLIS* mat, int n_node, const int* n_dof_list, int n_dof, const int* index, const int* item,
                   bool cplx) {
  std::vector<int> gi(index, index + n_node + 1), gt(item, item + index[n_node]);
  pattern(mat, n_node, gi, gt, n_dof_list, n_dof, cplx);
}

// ---------------------------------------------------------------------------
// Assembly (spmat_handler_util.f90)
// ---------------------------------------------------------------------------
int block_ndof(const MONOLIS_MAT* M, int i) { return M->n_dof_index[i + 1] - M->n_dof_index[i]; }

// Position (0-based) of block (ci, cj) in CSR order, or -1 (binary search of the
// ascending row, monolis_bsearch_I).
int find_block(const MONOLIS_MAT* M, int ci, int cj) {
  const int* b = M->CSR.item + M->CSR.index[ci];
  const int* e = M->CSR.item + M->CSR.index[ci + 1];
  const int* p = std::lower_bound(b, e, cj + 1);
  return (p != e && *p == cj + 1) ? (int)(p - M->CSR.item) : -1;
}

void check_node(const MONOLIS_MAT* M, int i) {
  if (i < 0 || i >= M->NP) {
    char msg[128];
    snprintf(msg, sizeof msg, "node %d outside the matrix (0..%d)", i, M->NP - 1);
    stop(msg);
  }
}

// Entry (ci, cj, sub_i, sub_j) of the value array, stopping as monolis does
// when the DOF is outside its block or the block is not in the pattern
// (spmat_handler_util.f90:24-38).
long long entry(const MONOLIS_MAT* M, int ci, int cj, int si, int sj) {
  check_node(M, ci);
  check_node(M, cj);
  const int ni = block_ndof(M, ci), nj = block_ndof(M, cj);
  if (si < 0 || si >= ni || sj < 0 || sj >= nj) stop("DOF index outside its block (monolis_stop_by_submatrix_access)");
  const int e = find_block(M, ci, cj);
  if (e < 0) stop("block is not in the nonzero pattern (monolis_stop_by_matrix_assemble)");
  return (long long)M->n_dof_index2[e] + (long long)nj * si + sj;
}

// Value arrays of the pattern: a pattern is built for real or for complex
// values, and the other kind is absent.
double* vals_R(const MONOLIS_MAT* M) {
  if (!M->R.A) stop("the matrix holds no real values (its pattern was built for complex values, or none was built)");
  return M->R.A;
}
zc* vals_C(const MONOLIS_MAT* M) {
  if (!M->C.A) stop("the matrix holds no complex values (its pattern was built for real values, or none was built)");
  return reinterpret_cast<zc*>(M->C.A);
}

// monolis_set_matrix_BCSR_mat_val_* copies n_dof^2 nz values into the value
// array (spmat_handler.f90:337-352); sizes other than the pattern's would
// write outside it, so they stop the program.
void check_mat_val(const MONOLIS_MAT* M, int n_dof, int nz) {
  if (M->NDOF != n_dof || M->CSR.index == nullptr || M->CSR.index[M->NP] != nz)
    stop("set_matrix_BCSR_mat_val: n_dof / nz differ from the matrix's pattern");
}

template <typename T>
void add_matrix_main(MONOLIS_MAT* M, T* A, int n1, int n2, const int* c1, const int* c2, T** val) {
  // monolis_add_matrix_to_sparse_matrix_main_R (spmat_handler_util.f90:398-497):
  // the row and column nodes are visited in ascending node order (stable for
  // repeated nodes), and each component is added once per element call.
  for (int a = 0; a < n1; ++a) check_node(M, c1[a]);
  for (int c = 0; c < n2; ++c) check_node(M, c2[c]);
  const int b = M->NDOF;
  if (b < 1) stop("element matrices need a fixed block size (n_dof >= 1)");
  std::vector<int> o1(n1), o2(n2);
  for (int a = 0; a < n1; ++a) o1[a] = a;
  for (int c = 0; c < n2; ++c) o2[c] = c;
  std::stable_sort(o1.begin(), o1.end(), [&](int x, int y) { return c1[x] < c1[y]; });
  std::stable_sort(o2.begin(), o2.end(), [&](int x, int y) { return c2[x] < c2[y]; });
  for (int ia : o1)
    for (int ic : o2) {
      const int e = find_block(M, c1[ia], c2[ic]);
      if (e < 0) stop("block is not in the nonzero pattern (monolis_stop_by_matrix_assemble)");
      T* blk = A + (size_t)e * b * b;
      for (int k1 = 0; k1 < b; ++k1)
        for (int k2 = 0; k2 < b; ++k2) blk[b * k1 + k2] = blk[b * k1 + k2] + val[ia * b + k1][ic * b + k2];
    }
}

// monolis_set_Dirichlet_bc_main_R (spmat_handler_util.f90:604-677): the
// condition's column (rows ascending, through the CSC view) moves val * A to
// the right-hand side and is zeroed; the row is zeroed with 1 on the diagonal
// and B(dof) = val.
template <typename T>
void dirichlet_main(MONOLIS_MAT* M, T* A, T* B, int node, int dof, T val) {
  check_node(M, node);
  const int nd1 = block_ndof(M, node);
  if (dof < 0 || dof >= nd1) stop("DOF index outside its block (monolis_stop_by_submatrix_access)");
  for (int p = M->CSC.index[node]; p < M->CSC.index[node + 1]; ++p) {
    const int jn = M->CSC.item[p] - 1;
    const int kn = M->CSC.perm[p] - 1;
    const int nd2 = block_ndof(M, jn);
    for (int k = 0; k < nd2; ++k) {
      const long long m = (long long)M->n_dof_index2[kn] + (long long)nd1 * k + dof;
      const long long n = (long long)M->n_dof_index[jn] + k;
      const T t = val * A[m];
      B[n] = B[n] - t;
      A[m] = T(0);
    }
  }
  bool diag = false;
  for (int e = M->CSR.index[node]; e < M->CSR.index[node + 1]; ++e) {
    const int jn = M->CSR.item[e] - 1;
    const int nd2 = block_ndof(M, jn);
    const long long m = (long long)M->n_dof_index2[e] + (long long)nd2 * dof;
    for (int k = 0; k < nd2; ++k) A[m + k] = T(0);
    if (jn == node) {
      A[(long long)M->n_dof_index2[e] + (long long)(nd2 + 1) * dof] = T(1);
      diag = true;
    }
  }
  if (!diag) stop("Dirichlet condition on a node without its diagonal block (monolis_stop_by_set_DBC)");
  B[M->n_dof_index[node] + dof] = val;
}

template <typename T>
void set_bcsr(MONOLIS* mat, int n, int np, int n_dof, int nz, const T* A, const int* index, const int* item,
              bool cplx) {
  if (n < 0 || np < n || nz < 0 || n_dof < 1) stop("invalid BCSR sizes");
  MONOLIS_MAT* M = &mat->mat;
  M->N = n;
  M->NP = np;
  M->NDOF = n_dof;
  release(M->CSR.index);
  release(M->CSR.item);
  M->CSR.index = alloc_zero<int>((size_t)np + 1);
  M->CSR.item = alloc_zero<int>((size_t)nz);
  for (int i = 0; i <= np; ++i) M->CSR.index[i] = index[i];
  if (M->CSR.index[0] != 0 || M->CSR.index[np] != nz) stop("invalid BCSR index (index[0] != 0 or index[np] != nz)");
  for (int e = 0; e < nz; ++e) {
    if (item[e] < 0 || item[e] >= np) stop("BCSR item outside 0..np-1");
    M->CSR.item[e] = item[e] + 1;  // monolis_spmat_handler_c.c stores item + 1
  }
  set_csc(M);
  set_dofs_and_values(M, nullptr, n_dof, cplx);
  const size_t nza = (size_t)nz * n_dof * n_dof;
  if (cplx)
    std::copy(A, A + nza, reinterpret_cast<T*>(M->C.A));
  else
    std::copy(A, A + nza, reinterpret_cast<T*>(M->R.A));
}

// Rows multiplied / solved on this rank (monolis_solver_c.c:14-15).
int rows_of(const MONOLIS* mat, const MONOLIS_COM* com) {
  return com->comm_size > 1 ? com->n_internal_vertex : mat->mat.N;
}

void sizes(const MONOLIS* mat, const MONOLIS_COM* com, int* n, int* nz, int* nndof, int* npndof, int* nzndof2) {
  const MONOLIS_MAT* M = &mat->mat;
  *n = rows_of(mat, com);
  *nz = M->CSR.index[M->NP];
  monolis_get_vec_size(*n, M->NP, M->NDOF, *nz, M->n_dof_index, M->n_dof_index2, nndof, npndof, nzndof2);
}

}  // namespace

extern "C" {

// ---- communication table ------------------------------------------------
void monolis_com_initialize_by_self(MONOLIS_COM* com) {
  memset(com, 0, sizeof(*com));
  com->comm_size = 1;
  com->recv_index = alloc_zero<int>(1);
  com->send_index = alloc_zero<int>(1);
  com->recv_neib_pe = alloc_zero<int>(1);
  com->send_neib_pe = alloc_zero<int>(1);
  com->recv_item = alloc_zero<int>(1);
  com->send_item = alloc_zero<int>(1);
}

void monolis_com_finalize(MONOLIS_COM* com) {
  release(com->recv_index);
  release(com->send_index);
  release(com->recv_neib_pe);
  release(com->send_neib_pe);
  release(com->recv_item);
  release(com->send_item);
}

// ---- initialisation -------------------------------------------------------
void monolis_global_initialize(void) {}
void monolis_global_finalize(void) {}

void monolis_prm_initialize(MONOLIS_PRM* prm) {
  // def_solver_prm.f90:183-222 / monolis_def_solver_prm_c.c
  memset(prm, 0, sizeof(*prm));
  strcpy(prm->com_top_dir_name, ".");
  strcpy(prm->com_part_dir_name, "parted.0");
  strcpy(prm->com_file_name, "node.dat");
  prm->Iarray[MONOLIS_PRM_I_METHOD] = MONOLIS_ITER_CG;
  prm->Iarray[MONOLIS_PRM_I_PRECOND] = MONOLIS_PREC_DIAG;
  prm->Iarray[MONOLIS_PRM_I_MAX_ITER] = 1000;
  prm->Iarray[MONOLIS_PRM_I_CUR_ITER] = 0;
  prm->Iarray[MONOLIS_PRM_I_IERR] = -1;
  prm->Iarray[MONOLIS_PRM_I_IS_INIT_X] = 1;
  prm->Iarray[MONOLIS_PRM_I_IS_ERROR_ABORT] = 1;
  prm->Iarray[MONOLIS_PRM_I_SHOW_ITERLOG] = 1;
  prm->Iarray[MONOLIS_PRM_I_SHOW_TIME] = 1;
  prm->Iarray[MONOLIS_PRM_I_SHOW_SUMMARY] = 1;
  prm->Iarray[MONOLIS_PRM_I_IDRS_DIM] = 4;
  prm->Iarray[MONOLIS_PRM_I_SPMV_FORMAT] = MONOLIS_SPMV_ELL;
  prm->Rarray[MONOLIS_PRM_R_TOL] = 1.0e-8;
}

void monolis_prm_finalize(MONOLIS_PRM* prm) {
  memset(prm->Iarray, 0, sizeof(prm->Iarray));
  memset(prm->Rarray, 0, sizeof(prm->Rarray));
}

void monolis_mat_initialize_val_R(MONOLIS_MAT_VAL_R* v) { memset(v, 0, sizeof(*v)); }
void monolis_mat_initialize_val_C(MONOLIS_MAT_VAL_C* v) { memset(v, 0, sizeof(*v)); }
void monolis_mat_initialize_SCSR(MONOLIS_MAT_SEPARATED_CSR* v) { memset(v, 0, sizeof(*v)); }
void monolis_mat_initialize_CSR(MONOLIS_MAT_CSR* v) { memset(v, 0, sizeof(*v)); }
void monolis_mat_initialize_CSC(MONOLIS_MAT_CSC* v) { memset(v, 0, sizeof(*v)); }
void monolis_mat_initialize(MONOLIS_MAT* M) { memset(M, 0, sizeof(*M)); }

void monolis_mat_finalize_val_R(MONOLIS_MAT_VAL_R* v) {
  release(v->A);
  release(v->U);
  release(v->D);
  release(v->L);
  release(v->X);
  release(v->B);
}
void monolis_mat_finalize_val_C(MONOLIS_MAT_VAL_C* v) {
  release(v->A);
  release(v->U);
  release(v->D);
  release(v->L);
  release(v->X);
  release(v->B);
}
void monolis_mat_finalize_SCSR(MONOLIS_MAT_SEPARATED_CSR* v) {
  release(v->indexU);
  release(v->indexL);
  release(v->itemU);
  release(v->itemL);
}
void monolis_mat_finalize_CSR(MONOLIS_MAT_CSR* v) {
  release(v->index);
  release(v->item);
}
void monolis_mat_finalize_CSC(MONOLIS_MAT_CSC* v) {
  release(v->index);
  release(v->item);
  release(v->perm);
}
void monolis_mat_finalize(MONOLIS_MAT* M) {
  release(M->n_dof_list);
  release(M->n_dof_index);
  release(M->n_dof_index2);
  monolis_mat_finalize_val_R(&M->R);
  monolis_mat_finalize_val_C(&M->C);
  monolis_mat_finalize_SCSR(&M->SCSR);
  monolis_mat_finalize_CSR(&M->CSR);
  monolis_mat_finalize_CSC(&M->CSC);
  M->N = M->NP = M->NDOF = 0;
}

void monolis_initialize(MONOLIS* mat) {
  monolis_prm_initialize(&mat->prm);
  monolis_mat_initialize(&mat->mat);
  monolis_mat_initialize(&mat->prec);
}

void monolis_finalize(MONOLIS* mat) {
  monolis_prm_finalize(&mat->prm);
  monolis_mat_finalize(&mat->mat);
  monolis_mat_finalize(&mat->prec);
}

void monolis_get_vec_size(int N, int NP, int NDOF, int NZ, int* n_dof_index, int* n_dof_index2, int* N_size,
                          int* NP_size, int* NZ_size) {
  // def_mat.f90:699-721 (monolis_def_mat_c.c)
  if (NDOF == -1) {
    *N_size = n_dof_index[N];
    *NP_size = n_dof_index[NP];
    *NZ_size = n_dof_index2[NZ];
  } else {
    *N_size = N * NDOF;
    *NP_size = NP * NDOF;
    *NZ_size = NZ * NDOF * NDOF;
  }
}

// ---- parameters -----------------------------------------------------------
void monolis_set_method(MONOLIS* mat, int p) { mat->prm.Iarray[MONOLIS_PRM_I_METHOD] = p; }
void monolis_set_precond(MONOLIS* mat, int p) { mat->prm.Iarray[MONOLIS_PRM_I_PRECOND] = p; }
void monolis_set_spmv_format(MONOLIS* mat, int p) { mat->prm.Iarray[MONOLIS_PRM_I_SPMV_FORMAT] = p; }
void monolis_set_maxiter(MONOLIS* mat, int p) { mat->prm.Iarray[MONOLIS_PRM_I_MAX_ITER] = p; }
void monolis_get_converge_iter(MONOLIS* mat, int* p) { *p = mat->prm.Iarray[MONOLIS_PRM_I_CUR_ITER]; }
void monolis_get_error_tag(MONOLIS* mat, int* p) { *p = mat->prm.Iarray[MONOLIS_PRM_I_IERR]; }
void monolis_set_init_x(MONOLIS* mat, bool p) { mat->prm.Iarray[MONOLIS_PRM_I_IS_INIT_X] = p ? 1 : 0; }
void monolis_set_sym_matrix(MONOLIS* mat, bool p) { mat->prm.Iarray[MONOLIS_PRM_I_IS_SYM_MATRIX] = p ? 1 : 0; }
void monolis_set_debug(MONOLIS* mat, bool p) { mat->prm.Iarray[MONOLIS_PRM_I_IS_DEBUG] = p ? 1 : 0; }
void monolis_set_performance_measurement(MONOLIS* mat, bool p) {
  mat->prm.Iarray[MONOLIS_PRM_I_IS_MEASUREMENT] = p ? 1 : 0;
}
void monolis_set_check_diag(MONOLIS* mat, bool p) { mat->prm.Iarray[MONOLIS_PRM_I_IS_CHECK_DIAG] = p ? 1 : 0; }
void monolis_set_prec_stored(MONOLIS* mat, bool p) { mat->prm.Iarray[MONOLIS_PRM_I_IS_PREC_STORED] = p ? 1 : 0; }
void monolis_set_error_abort(MONOLIS* mat, bool p) { mat->prm.Iarray[MONOLIS_PRM_I_IS_ERROR_ABORT] = p ? 1 : 0; }
void monolis_set_solver_IDRS_num_basis(MONOLIS* mat, int p) { mat->prm.Iarray[MONOLIS_PRM_I_IDRS_DIM] = p; }
void monolis_show_iterlog(MONOLIS* mat, bool p) { mat->prm.Iarray[MONOLIS_PRM_I_SHOW_ITERLOG] = p ? 1 : 0; }
void monolis_show_timelog(MONOLIS* mat, bool p) { mat->prm.Iarray[MONOLIS_PRM_I_SHOW_TIME] = p ? 1 : 0; }
void monolis_show_summary(MONOLIS* mat, bool p) { mat->prm.Iarray[MONOLIS_PRM_I_SHOW_SUMMARY] = p ? 1 : 0; }
void monolis_show_timelog_statistics(MONOLIS* mat, bool p) {
  mat->prm.Iarray[MONOLIS_PRM_I_SHOW_TIME_STATISTICS] = p ? 1 : 0;
}
void monolis_set_tolerance(MONOLIS* mat, double v) { mat->prm.Rarray[MONOLIS_PRM_R_TOL] = v; }
void monolis_get_converge_residual(MONOLIS* mat, double* v) { *v = mat->prm.Rarray[MONOLIS_PRM_R_CUR_RESID]; }
void monolis_get_time_solver(MONOLIS* mat, double* v) { *v = mat->prm.Rarray[MONOLIS_R_TIME_SOL]; }
void monolis_get_time_preparing(MONOLIS* mat, double* v) { *v = mat->prm.Rarray[MONOLIS_R_TIME_PREP]; }
void monolis_get_time_spmv(MONOLIS* mat, double* v) { *v = mat->prm.Rarray[MONOLIS_R_TIME_SPMV]; }
void monolis_get_time_inner_product(MONOLIS* mat, double* v) { *v = mat->prm.Rarray[MONOLIS_R_TIME_DOTP]; }
void monolis_get_time_precondition(MONOLIS* mat, double* v) { *v = mat->prm.Rarray[MONOLIS_R_TIME_PREC]; }
void monolis_get_time_comm_inner_product(MONOLIS* mat, double* v) {
  *v = mat->prm.Rarray[MONOLIS_R_TIME_COMM_DOTP];
}
void monolis_get_time_comm_spmv(MONOLIS* mat, double* v) { *v = mat->prm.Rarray[MONOLIS_R_TIME_COMM_SPMV]; }

// ---- nonzero patterns -----------------------------------------------------
void monolis_get_nonzero_pattern_by_simple_mesh_R(MONOLIS* mat, int n_node, int n_base, int n_dof, int n_elem,
                                                  int** elem) {
  pattern_simple_mesh(mat, n_node, n_base, nullptr, n_dof, n_elem, elem, false);
}
void monolis_get_nonzero_pattern_by_simple_mesh_V_R(MONOLIS* mat, int n_node, int n_base, int* n_dof_list,
                                                    int n_elem, int** elem) {
  pattern_simple_mesh(mat, n_node, n_base, n_dof_list, -1, n_elem, elem, false);
}
void monolis_get_nonzero_pattern_by_connectivity_R(MONOLIS* mat, int n_node, int n_dof, int n_elem,
                                                   int* conn_index, int* conn_item) {
  pattern_connectivity(mat, n_node, n_dof, n_elem, conn_index, conn_item, false);
}
void monolis_get_nonzero_pattern_by_nodal_graph_R(MONOLIS* mat, int n_node, int n_dof, int* index, int* item) {
  pattern_graph(mat, n_node, nullptr, n_dof, index, item, false);
}
void monolis_get_nonzero_pattern_by_nodal_graph_V_R(MONOLIS* mat, int n_node, int* n_dof_list, int* index,
                                                    int* item) {
  pattern_graph(mat, n_node, n_dof_list, -1, index, item, false);
}
void monolis_get_nonzero_pattern_by_simple_mesh_C(MONOLIS* mat, int n_node, int n_base, int n_dof, int n_elem,
                                                  int** elem) {
  pattern_simple_mesh(mat, n_node, n_base, nullptr, n_dof, n_elem, elem, true);
}
void monolis_get_nonzero_pattern_by_simple_mesh_V_C(MONOLIS* mat, int n_node, int n_base, int* n_dof_list,
                                                    int n_elem, int** elem) {
  pattern_simple_mesh(mat, n_node, n_base, n_dof_list, -1, n_elem, elem, true);
}
void monolis_get_nonzero_pattern_by_connectivity_C(MONOLIS* mat, int n_node, int n_dof, int n_elem,
                                                   int* conn_index, int* conn_item) {
  pattern_connectivity(mat, n_node, n_dof, n_elem, conn_index, conn_item, true);
}
void monolis_get_nonzero_pattern_by_nodal_graph_C(MONOLIS* mat, int n_node, int n_dof, int* index, int* item) {
  pattern_graph(mat, n_node, nullptr, n_dof, index, item, true);
}
void monolis_get_nonzero_pattern_by_nodal_graph_V_C(MONOLIS* mat, int n_node, int* n_dof_list, int* index,
                                                    int* item) {
  pattern_graph(mat, n_node, n_dof_list, -1, index, item, true);
}

// ---- assembly -------------------------------------------------------------
void monolis_set_scalar_to_sparse_matrix_R(MONOLIS* mat, int i, int j, int si, int sj, double val) {
  vals_R(&mat->mat)[entry(&mat->mat, i, j, si, sj)] = val;
}
void monolis_add_scalar_to_sparse_matrix_R(MONOLIS* mat, int i, int j, int si, int sj, double val) {
  double& a = vals_R(&mat->mat)[entry(&mat->mat, i, j, si, sj)];
  a = a + val;
}
void monolis_get_scalar_from_sparse_matrix_R(MONOLIS* mat, int i, int j, int si, int sj, double* val,
                                             bool* is_find) {
  // spmat_handler_util.f90 get_scalar: not found -> 0 and is_find = false
  *val = 0.0;
  *is_find = false;
  const MONOLIS_MAT* M = &mat->mat;
  check_node(M, i);
  check_node(M, j);
  if (si < 0 || si >= block_ndof(M, i) || sj < 0 || sj >= block_ndof(M, j))
    stop("DOF index outside its block (monolis_stop_by_submatrix_access)");
  const int e = find_block(M, i, j);
  if (e < 0) return;
  *val = vals_R(M)[(long long)M->n_dof_index2[e] + (long long)block_ndof(M, j) * si + sj];
  *is_find = true;
}
void monolis_add_matrix_to_sparse_matrix_R(MONOLIS* mat, int n_base, int* connectivity, double** val) {
  add_matrix_main<double>(&mat->mat, vals_R(&mat->mat), n_base, n_base, connectivity, connectivity, val);
}
void monolis_add_matrix_to_sparse_matrix_offdiag_R(MONOLIS* mat, int n_base1, int n_base2, int* connectivity1,
                                                   int* connectivity2, double** val) {
  add_matrix_main<double>(&mat->mat, vals_R(&mat->mat), n_base1, n_base2, connectivity1, connectivity2, val);
}
void monolis_set_matrix_BCSR_R(MONOLIS* mat, int n, int np, int n_dof, int nz, double* A, int* index, int* item) {
  set_bcsr<double>(mat, n, np, n_dof, nz, A, index, item, false);
}
void monolis_set_matrix_BCSR_mat_val_R(MONOLIS* mat, int n_dof, int nz, double* A) {
  check_mat_val(&mat->mat, n_dof, nz);
  std::copy(A, A + (size_t)nz * n_dof * n_dof, vals_R(&mat->mat));
}
void monolis_set_Dirichlet_bc_R(MONOLIS* mat, double* b, int node_id, int n_dof_bc, double val) {
  dirichlet_main<double>(&mat->mat, vals_R(&mat->mat), b, node_id, n_dof_bc, val);
}

void monolis_set_scalar_to_sparse_matrix_C(MONOLIS* mat, int i, int j, int si, int sj, double _Complex val) {
  vals_C(&mat->mat)[entry(&mat->mat, i, j, si, sj)] = reinterpret_cast<const zc&>(val);
}
void monolis_add_scalar_to_sparse_matrix_C(MONOLIS* mat, int i, int j, int si, int sj, double _Complex val) {
  zc& a = vals_C(&mat->mat)[entry(&mat->mat, i, j, si, sj)];
  a = a + reinterpret_cast<const zc&>(val);
}
void monolis_get_scalar_from_sparse_matrix_C(MONOLIS* mat, int i, int j, int si, int sj, double _Complex* val,
                                             bool* is_find) {
  zc* out = zptr(val);
  *out = zc(0.0, 0.0);
  *is_find = false;
  const MONOLIS_MAT* M = &mat->mat;
  check_node(M, i);
  check_node(M, j);
  if (si < 0 || si >= block_ndof(M, i) || sj < 0 || sj >= block_ndof(M, j))
    stop("DOF index outside its block (monolis_stop_by_submatrix_access)");
  const int e = find_block(M, i, j);
  if (e < 0) return;
  *out = vals_C(M)[(long long)M->n_dof_index2[e] + (long long)block_ndof(M, j) * si + sj];
  *is_find = true;
}
void monolis_add_matrix_to_sparse_matrix_C(MONOLIS* mat, int n_base, int* connectivity, double _Complex** val) {
  add_matrix_main<zc>(&mat->mat, vals_C(&mat->mat), n_base, n_base, connectivity, connectivity,
                      reinterpret_cast<zc**>(val));
}
void monolis_add_matrix_to_sparse_matrix_offdiag_C(MONOLIS* mat, int n_base1, int n_base2, int* connectivity1,
                                                   int* connectivity2, double _Complex** val) {
  add_matrix_main<zc>(&mat->mat, vals_C(&mat->mat), n_base1, n_base2, connectivity1, connectivity2,
                      reinterpret_cast<zc**>(val));
}
void monolis_set_matrix_BCSR_C(MONOLIS* mat, int n, int np, int n_dof, int nz, double _Complex* A, int* index,
                               int* item) {
  set_bcsr<zc>(mat, n, np, n_dof, nz, zptr(A), index, item, true);
}
void monolis_set_matrix_BCSR_mat_val_C(MONOLIS* mat, int n_dof, int nz, double _Complex* A) {
  check_mat_val(&mat->mat, n_dof, nz);
  std::copy(zptr(A), zptr(A) + (size_t)nz * n_dof * n_dof, vals_C(&mat->mat));
}
void monolis_set_Dirichlet_bc_C(MONOLIS* mat, double _Complex* b, int node_id, int n_dof_bc, double _Complex val) {
  dirichlet_main<zc>(&mat->mat, vals_C(&mat->mat), zptr(b), node_id, n_dof_bc, reinterpret_cast<const zc&>(val));
}

// ---- solver and linear algebra (GPU path) ---------------------------------
void monolis_solve_R(MONOLIS* mat, MONOLIS_COM* com, double* b, double* x) {
  int n, nz, nndof, npndof, nzndof2;
  sizes(mat, com, &n, &nz, &nndof, &npndof, &nzndof2);
  MONOLIS_MAT* M = &mat->mat;
  monolis_solve_R_c_main(n, M->NP, nz, M->NDOF, npndof, nzndof2, M->n_dof_list, M->n_dof_index, M->n_dof_index2,
                         vals_R(M), x, b, M->CSR.index, M->CSR.item, com->my_rank, com->comm, com->comm_size,
                         com->recv_n_neib, com->recv_index[com->recv_n_neib], com->recv_neib_pe, com->recv_index,
                         com->recv_item, com->send_n_neib, com->send_index[com->send_n_neib], com->send_neib_pe,
                         com->send_index, com->send_item, mat->prm.Iarray, mat->prm.Rarray);
}

void monolis_solve_C(MONOLIS* mat, MONOLIS_COM* com, double _Complex* b, double _Complex* x) {
  int n, nz, nndof, npndof, nzndof2;
  sizes(mat, com, &n, &nz, &nndof, &npndof, &nzndof2);
  MONOLIS_MAT* M = &mat->mat;
  monolis_solve_C_c_main(n, M->NP, nz, M->NDOF, npndof, nzndof2, M->n_dof_list, M->n_dof_index, M->n_dof_index2,
                         reinterpret_cast<double _Complex*>(vals_C(M)), x, b, M->CSR.index, M->CSR.item, com->my_rank, com->comm, com->comm_size,
                         com->recv_n_neib, com->recv_index[com->recv_n_neib], com->recv_neib_pe, com->recv_index,
                         com->recv_item, com->send_n_neib, com->send_index[com->send_n_neib], com->send_neib_pe,
                         com->send_index, com->send_item, mat->prm.Iarray, mat->prm.Rarray);
}

void monolis_matvec_product_R(MONOLIS* mat, MONOLIS_COM* com, double* x, double* y) {
  int n, nz, nndof, npndof, nzndof2;
  sizes(mat, com, &n, &nz, &nndof, &npndof, &nzndof2);
  MONOLIS_MAT* M = &mat->mat;
  monolis_matvec_product_R_c_main(n, M->NP, nz, M->NDOF, npndof, nzndof2, M->n_dof_list, M->n_dof_index,
                                  M->n_dof_index2, vals_R(M), x, y, M->CSR.index, M->CSR.item, com->my_rank, com->comm,
                                  com->comm_size, com->recv_n_neib, com->recv_index[com->recv_n_neib],
                                  com->recv_neib_pe, com->recv_index, com->recv_item, com->send_n_neib,
                                  com->send_index[com->send_n_neib], com->send_neib_pe, com->send_index,
                                  com->send_item);
}

void monolis_matvec_product_C(MONOLIS* mat, MONOLIS_COM* com, double _Complex* x, double _Complex* y) {
  int n, nz, nndof, npndof, nzndof2;
  sizes(mat, com, &n, &nz, &nndof, &npndof, &nzndof2);
  MONOLIS_MAT* M = &mat->mat;
  monolis_matvec_product_C_c_main(n, M->NP, nz, M->NDOF, npndof, nzndof2, M->n_dof_list, M->n_dof_index,
                                  M->n_dof_index2, reinterpret_cast<double _Complex*>(vals_C(M)), x, y, M->CSR.index, M->CSR.item, com->my_rank, com->comm,
                                  com->comm_size, com->recv_n_neib, com->recv_index[com->recv_n_neib],
                                  com->recv_neib_pe, com->recv_index, com->recv_item, com->send_n_neib,
                                  com->send_index[com->send_n_neib], com->send_neib_pe, com->send_index,
                                  com->send_item);
}

// monolis_inner_product_c.c: m = (internal rows) x n_dof
void monolis_inner_product_I(MONOLIS* mat, MONOLIS_COM* com, int n_dof, int* x, int* y, int* sum) {
  monolis_inner_productV_I(com, rows_of(mat, com) * n_dof, x, y, sum);
}
void monolis_inner_productV_I(MONOLIS_COM* com, int m, int* x, int* y, int* sum) {
  (void)com;
  int s = 0;
  for (int i = 0; i < m; ++i) s += x[i] * y[i];  // inner_product.f90 monolis_inner_product_main_I
  *sum = s;
}
void monolis_inner_product_R(MONOLIS* mat, MONOLIS_COM* com, int n_dof, double* x, double* y, double* sum) {
  monolis_inner_product_R_c_main(rows_of(mat, com) * n_dof, x, y, sum, com->comm);
}
void monolis_inner_productV_R(MONOLIS_COM* com, int m, double* x, double* y, double* sum) {
  monolis_inner_product_R_c_main(m, x, y, sum, com->comm);
}
void monolis_inner_product_C(MONOLIS* mat, MONOLIS_COM* com, int n_dof, double _Complex* x, double _Complex* y,
                             double _Complex* sum) {
  monolis_inner_product_C_c_main(rows_of(mat, com) * n_dof, x, y, sum, com->comm);
}
void monolis_inner_productV_C(MONOLIS_COM* com, int m, double _Complex* x, double _Complex* y,
                              double _Complex* sum) {
  monolis_inner_product_C_c_main(m, x, y, sum, com->comm);
}

void monolis_vec_copy_I(int m, int* x, int* y) { std::copy(x, x + m, y); }
void monolis_vec_copy_R(int m, double* x, double* y) { std::copy(x, x + m, y); }
void monolis_vec_copy_C(int m, double _Complex* x, double _Complex* y) { std::copy(x, x + m, y); }

}  // extern "C"
