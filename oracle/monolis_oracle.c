/*
 * monolis_oracle.c -- CPU restatement of monolis's FP64 Krylov path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * library under monolis_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker or as the
 * reported CPU baseline.  The product path never calls it.
 *
 * Every function follows one reference routine loop for loop; the citation is
 * given per function (paths relative to the monolis source tree).  Indices
 * follow the reference: `index` holds 0-based row offsets of length NP+1 and
 * `item` holds 1-based block columns, diagonal included, ascending per row.
 * Block values are row-major inside a block (src/linalg/matvec.f90:392).
 *
 * Threading: loops carry OpenMP pragmas with a static schedule, mirroring the
 * reference's `!$omp do` regions, so the same code serves as the multi-core
 * CPU baseline.  With one thread every reduction is a plain left-to-right sum.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   orc_matvec_*      reference KATs (wrapper_test/linalg/monolis_matvec_c_test.c
 *                     :63-67, :141-150, :238-252) and dense brute force.
 *   orc_inner_product closed form 10 (src_test/linalg/inner_product_test.f90:63-73).
 *   orc_vec_axpby     closed form 6 (src_test/linalg/vec_util_test.f90:108-118).
 *   orc_check_converge  resid 2.0 (src_test/linalg/mat_converge_test.f90:44-55).
 *   orc_diag_*        per-block numpy.linalg.solve, apply(D_i v) = v, zero pivot.
 *   orc_solve         x = 1 closed forms of the reference's solver tests and
 *                     scipy cg/bicgstab iteration counts.
 *   orc_sorted_dot    brute force: sum of products ordered by |x_i y_i|.
 *                     The order among products of equal magnitude is parity
 *                     unpinned: monolis takes it from monolis_qsort_R_2d
 *                     (monolis_utils, absent); this restatement keeps index
 *                     order, as the GPU's stable radix sort does.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_ITER_CG 1
#define ORC_ITER_GROPCG 2
#define ORC_ITER_PIPECG 3
#define ORC_ITER_PIPECR 4
#define ORC_ITER_BICGSTAB 5
#define ORC_ITER_PIPEBICGSTAB 6
#define ORC_ITER_BICGSTAB_NOPREC 7
#define ORC_ITER_PIPEBICGSTAB_NOPREC 8
#define ORC_PREC_NONE 0
#define ORC_PREC_DIAG 1

/* Iarray / Rarray slots, C index k == Fortran index k
 * (src/define/def_solver_prm.f90:90-163, wrapper/solver/solver_wrapper.f90:69-75). */
#define I_METHOD 1
#define I_PRECOND 2
#define I_MAX_ITER 3
#define I_CUR_ITER 4
#define I_IERR 5
#define I_IS_INIT_X 10
#define I_IS_MEASUREMENT 13
#define I_IS_ERROR_ABORT 16
#define R_TOL 1
#define R_CUR_RESID 2
#define R_TIME_SOL 3
#define R_TIME_PREP 4

static double orc_now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* ---------------------------------------------------------------------------
 * SpMV.  src/linalg/matvec.f90:707-741 (11), :783-826 (33), :355-402 (nn).
 * y(1..N*ndof) = A x for block rows 1..N.
 * ------------------------------------------------------------------------- */
void orc_matvec_11(int N, const int* index, const int* item, const double* A,
                   const double* X, double* Y) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    double y1 = 0.0;
    for (int j = index[i]; j < index[i + 1]; ++j) y1 += A[j] * X[item[j] - 1];
    Y[i] = y1;
  }
}

void orc_matvec_33(int N, const int* index, const int* item, const double* A,
                   const double* X, double* Y) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    double y1 = 0.0, y2 = 0.0, y3 = 0.0;
    for (int j = index[i]; j < index[i + 1]; ++j) {
      const int in = item[j] - 1;
      const double x1 = X[3 * in], x2 = X[3 * in + 1], x3 = X[3 * in + 2];
      const double* a = A + 9 * (size_t)j;
      y1 = y1 + a[0] * x1 + a[1] * x2 + a[2] * x3;
      y2 = y2 + a[3] * x1 + a[4] * x2 + a[5] * x3;
      y3 = y3 + a[6] * x1 + a[7] * x2 + a[8] * x3;
    }
    Y[3 * i] = y1;
    Y[3 * i + 1] = y2;
    Y[3 * i + 2] = y3;
  }
}

void orc_matvec_nn(int N, int ndof, const int* index, const int* item,
                   const double* A, const double* X, double* Y) {
  const int nd2 = ndof * ndof;
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    double yt[64];
    for (int k = 0; k < ndof; ++k) yt[k] = 0.0;
    for (int j = index[i]; j < index[i + 1]; ++j) {
      const int in = item[j] - 1;
      for (int k = 0; k < ndof; ++k)
        for (int l = 0; l < ndof; ++l)
          yt[k] += A[(size_t)nd2 * j + ndof * k + l] * X[(size_t)ndof * in + l];
    }
    for (int k = 0; k < ndof; ++k) Y[(size_t)ndof * i + k] = yt[k];
  }
}

/* Variable-block SpMV, src/linalg/matvec.f90:231-289 (NDOF = -1).
 * n_dof_index (NP+1) and n_dof_index2 (NZ+1) are 0-based prefix sums. */
void orc_matvec_V(int N, const int* n_dof_list, const int* n_dof_index,
                  const int* n_dof_index2, const int* index, const int* item,
                  const double* A, const double* X, double* Y) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    double yt[64];
    const int n1 = n_dof_list[i];
    for (int k = 0; k < n1; ++k) yt[k] = 0.0;
    for (int j = index[i]; j < index[i + 1]; ++j) {
      const int in = item[j] - 1;
      const int n2 = n_dof_list[in];
      const int kx = n_dof_index[in];
      const int ka = n_dof_index2[j];
      for (int k = 0; k < n1; ++k)
        for (int l = 0; l < n2; ++l) yt[k] += A[ka + n2 * k + l] * X[kx + l];
    }
    for (int k = 0; k < n1; ++k) Y[n_dof_index[i] + k] = yt[k];
  }
}

/* Dispatch of src/linalg/matvec.f90:173-184. */
void orc_matvec(int N, int ndof, const int* n_dof_list, const int* n_dof_index,
                const int* n_dof_index2, const int* index, const int* item,
                const double* A, const double* X, double* Y) {
  if (ndof == 3)
    orc_matvec_33(N, index, item, A, X, Y);
  else if (ndof == 1)
    orc_matvec_11(N, index, item, A, X, Y);
  else if (ndof == -1)
    orc_matvec_V(N, n_dof_list, n_dof_index, n_dof_index2, index, item, A, X, Y);
  else
    orc_matvec_nn(N, ndof, index, item, A, X, Y);
}

/* ---------------------------------------------------------------------------
 * BLAS-1.  src/linalg/inner_product.f90:148-190 (+ no_comm :503-528),
 * src/linalg/vec_util.f90:42-66 (copy), :132-162 (AXPBY).
 * ------------------------------------------------------------------------- */
double orc_inner_product(int m, const double* X, const double* Y) {
  double sum = 0.0;
  int i;
#pragma omp parallel for schedule(static) reduction(+ : sum)
  for (i = 0; i < m; ++i) sum = sum + X[i] * Y[i];
  return sum;
}

void orc_vec_copy(int m, const double* X, double* Y) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < m; ++i) Y[i] = X[i];
}

void orc_vec_axpby(int m, double alpha, const double* X, double beta,
                   const double* Y, double* Z) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < m; ++i) Z[i] = alpha * X[i] + beta * Y[i];
}

typedef struct {
  double v;
  int i;
} orc_prod;

/* ascending |v|; equal magnitudes keep index order */
static int cmp_abs(const void* a, const void* b) {
  const orc_prod* p = (const orc_prod*)a;
  const orc_prod* q = (const orc_prod*)b;
  const double x = fabs(p->v), y = fabs(q->v);
  if (x < y) return -1;
  if (x > y) return 1;
  return (p->i > q->i) - (p->i < q->i);
}

/* Breakdown fallback of BiCGSTAB: products sorted by |x_i y_i| ascending, then
 * summed.  src/linalg/inner_product.f90:262-331 (one rank). */
double orc_sorted_inner_product(int n, const double* X, const double* Y) {
  orc_prod* xy = (orc_prod*)malloc(sizeof(orc_prod) * (size_t)(n > 0 ? n : 1));
  for (int i = 0; i < n; ++i) {
    xy[i].v = X[i] * Y[i];
    xy[i].i = i;
  }
  qsort(xy, (size_t)n, sizeof(orc_prod), cmp_abs);
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += xy[i].v;
  free(xy);
  return s;
}

/* ---------------------------------------------------------------------------
 * Convergence test.  src/linalg/mat_converge.f90:90-141.
 * Returns 1 when converged (resid < tol).  Writes cur_iter / cur_resid.
 * ------------------------------------------------------------------------- */
int orc_check_converge(int m, const double* R, double B2, int iter, int* Iarray,
                       double* Rarray) {
  const double R2 = orc_inner_product(m, R, R);
  const double resid = sqrt(R2 / B2);
  Iarray[I_CUR_ITER] = iter;
  Rarray[R_CUR_RESID] = resid;
  return resid < Rarray[R_TOL];
}

/* ---------------------------------------------------------------------------
 * Diagonal (block-Jacobi) preconditioner.
 * src/prec/diag/diag_33.f90:16-99 (setup), :103-139 (apply);
 * src/prec/diag/diag_nn.f90:16-86 (setup), :185-231 (apply).
 * D holds, per block row, the in-place no-pivot Doolittle LU of the diagonal
 * block: unit L strictly below, U above, pivots stored as reciprocals, and a
 * zero pivot replaced by 1 before inversion.
 * ------------------------------------------------------------------------- */
void orc_diag_nn_setup(int N, int ndof, const int* index, const int* item,
                       const double* A, double* D) {
  const int nd2 = ndof * ndof;
  memset(D, 0, sizeof(double) * (size_t)nd2 * (size_t)N);
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    double LU[4096];
    double T[64];
    for (int ii = index[i]; ii < index[i + 1]; ++ii) {
      if (item[ii] - 1 != i) continue;
      for (int e = 0; e < nd2; ++e) LU[e] = A[(size_t)nd2 * ii + e];
      for (int k = 0; k < ndof; ++k) {
        if (LU[k * ndof + k] == 0.0) LU[k * ndof + k] = 1.0;
        LU[k * ndof + k] = 1.0 / LU[k * ndof + k];
        for (int l = k + 1; l < ndof; ++l) {
          LU[l * ndof + k] = LU[l * ndof + k] * LU[k * ndof + k];
          for (int j = k + 1; j < ndof; ++j)
            T[j] = LU[l * ndof + j] - LU[l * ndof + k] * LU[k * ndof + j];
          for (int j = k + 1; j < ndof; ++j) LU[l * ndof + j] = T[j];
        }
      }
      for (int e = 0; e < nd2; ++e) D[(size_t)nd2 * i + e] = LU[e];
    }
  }
}

void orc_diag_nn_apply(int N, int ndof, const double* D, const double* X,
                       double* Y) {
  const int nd2 = ndof * ndof;
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    double T[64];
    const double* L = D + (size_t)nd2 * i;
    for (int j = 0; j < ndof; ++j) T[j] = X[(size_t)ndof * i + j];
    for (int j = 1; j < ndof; ++j)
      for (int k = 0; k < j; ++k) T[j] = T[j] - L[ndof * j + k] * T[k];
    for (int j = ndof - 1; j >= 0; --j) {
      for (int k = ndof - 1; k > j; --k) T[j] = T[j] - L[ndof * j + k] * T[k];
      T[j] = L[(ndof + 1) * j] * T[j];
    }
    for (int k = 0; k < ndof; ++k) Y[(size_t)ndof * i + k] = T[k];
  }
}

/* diag_33 setup (diag_33.f90:40-96): the diagonal block is copied out first
 * (rows without one keep zeros), then EVERY row is LU-factored, so a missing
 * diagonal block becomes the identity through the zero-pivot rule.  This is
 * where diag_33 differs from diag_nn, which factors only rows that have a
 * diagonal entry. */
void orc_diag_33_setup(int N, const int* index, const int* item, const double* A,
                       double* D) {
  memset(D, 0, sizeof(double) * 9 * (size_t)N);
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    for (int j = index[i]; j < index[i + 1]; ++j)
      if (item[j] - 1 == i)
        for (int e = 0; e < 9; ++e) D[9 * (size_t)i + e] = A[9 * (size_t)j + e];
    double* T = D + 9 * (size_t)i;
    double P[3];
    for (int k = 0; k < 3; ++k) {
      if (T[4 * k] == 0.0) T[4 * k] = 1.0;
      T[4 * k] = 1.0 / T[4 * k];
      for (int r = k + 1; r < 3; ++r) {
        T[3 * r + k] = T[3 * r + k] * T[4 * k];
        for (int c = k + 1; c < 3; ++c) P[c] = T[3 * r + c] - T[3 * r + k] * T[3 * k + c];
        for (int c = k + 1; c < 3; ++c) T[3 * r + c] = P[c];
      }
    }
  }
}

/* The 3x3 specialisation; same arithmetic order as diag_33.f90:123-135. */
void orc_diag_33_apply(int N, const double* D, const double* X, double* Y) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    const double* L = D + 9 * (size_t)i;
    double x1 = X[3 * i], x2 = X[3 * i + 1], x3 = X[3 * i + 2];
    x2 = x2 - L[3] * x1;
    x3 = x3 - L[6] * x1 - L[7] * x2;
    x3 = L[8] * x3;
    x2 = L[4] * (x2 - L[5] * x3);
    x1 = L[0] * (x1 - L[2] * x3 - L[1] * x2);
    Y[3 * i] = x1;
    Y[3 * i + 1] = x2;
    Y[3 * i + 2] = x3;
  }
}

/* Variable-block diagonal preconditioner, src/prec/diag/diag_V.f90:13-204:
 * same LU per diagonal block of size n_dof_list[i]; D is laid out with the
 * offsets sum_{k<i} n_dof_list[k]^2. */
void orc_diag_V_setup(int N, const int* n_dof_list, const int* n_dof_index2,
                      const int* index, const int* item, const double* A,
                      long long* doff, double* D) {
  long long off = 0;
  for (int i = 0; i < N; ++i) {
    doff[i] = off;
    off += (long long)n_dof_list[i] * n_dof_list[i];
  }
  doff[N] = off;
  memset(D, 0, sizeof(double) * (size_t)off);
  for (int i = 0; i < N; ++i) {
    const int nd = n_dof_list[i];
    double T[64];
    for (int ii = index[i]; ii < index[i + 1]; ++ii) {
      if (item[ii] - 1 != i) continue;
      double* LU = D + doff[i];
      for (int e = 0; e < nd * nd; ++e) LU[e] = A[n_dof_index2[ii] + e];
      for (int k = 0; k < nd; ++k) {
        if (LU[k * nd + k] == 0.0) LU[k * nd + k] = 1.0;
        LU[k * nd + k] = 1.0 / LU[k * nd + k];
        for (int l = k + 1; l < nd; ++l) {
          LU[l * nd + k] = LU[l * nd + k] * LU[k * nd + k];
          for (int j = k + 1; j < nd; ++j)
            T[j] = LU[l * nd + j] - LU[l * nd + k] * LU[k * nd + j];
          for (int j = k + 1; j < nd; ++j) LU[l * nd + j] = T[j];
        }
      }
    }
  }
}

void orc_diag_V_apply(int N, const int* n_dof_list, const int* n_dof_index,
                      const long long* doff, const double* D, const double* X,
                      double* Y) {
  for (int i = 0; i < N; ++i) {
    const int nd = n_dof_list[i];
    const double* L = D + doff[i];
    double T[64];
    for (int j = 0; j < nd; ++j) T[j] = X[n_dof_index[i] + j];
    for (int j = 1; j < nd; ++j)
      for (int k = 0; k < j; ++k) T[j] -= L[nd * j + k] * T[k];
    for (int j = nd - 1; j >= 0; --j) {
      for (int k = nd - 1; k > j; --k) T[j] -= L[nd * j + k] * T[k];
      T[j] = L[(nd + 1) * j] * T[j];
    }
    for (int k = 0; k < nd; ++k) Y[n_dof_index[i] + k] = T[k];
  }
}

/* ---------------------------------------------------------------------------
 * Solver context: matrix + preconditioner, as monolis_solve_main_R sees them.
 * ------------------------------------------------------------------------- */
typedef struct {
  int N, ndof, m;
  const int *n_dof_list, *n_dof_index, *n_dof_index2, *index, *item;
  const double* A;
  int precond;
  double* D;
  long long* doff;
} orc_ctx;

static void ctx_matvec(const orc_ctx* c, const double* X, double* Y) {
  orc_matvec(c->N, c->ndof, c->n_dof_list, c->n_dof_index, c->n_dof_index2,
             c->index, c->item, c->A, X, Y);
}

/* src/linalg/matvec.f90:877-903: R = B - A X. */
static void ctx_residual(const orc_ctx* c, const double* X, const double* B,
                         double* R) {
  ctx_matvec(c, X, R);
  orc_vec_axpby(c->m, 1.0, B, -1.0, R, R);
}

/* src/prec/precond.f90:147-231 (DIAG or NONE). */
static void ctx_prec_apply(const orc_ctx* c, const double* X, double* Y) {
  if (c->precond == ORC_PREC_DIAG) {
    if (c->ndof == -1)
      orc_diag_V_apply(c->N, c->n_dof_list, c->n_dof_index, c->doff, c->D, X, Y);
    else if (c->ndof == 3)
      orc_diag_33_apply(c->N, c->D, X, Y);
    else
      orc_diag_nn_apply(c->N, c->ndof, c->D, X, Y);
  } else {
    orc_vec_copy(c->m, X, Y);
  }
}

/* Outcome codes returned by the solvers (and written to Iarray[I_IERR]). */
#define ORC_OK 0
#define ORC_NOT_CONVERGED 1
#define ORC_BAD_METHOD -2
#define ORC_BAD_PRECOND -3

/* src/iterative/CG.f90:18-119. */
static int orc_cg(const orc_ctx* c, const double* B, double* X, int* Iarray,
                  double* Rarray) {
  const int m = c->m;
  const int iter_RR = 200;
  double rho = 0.0, rho1 = 0.0, B2;
  if (Iarray[I_IS_INIT_X]) memset(X, 0, sizeof(double) * (size_t)m);
  double* R = (double*)calloc((size_t)m + 1, sizeof(double));
  double* Z = (double*)calloc((size_t)m + 1, sizeof(double));
  double* Q = (double*)calloc((size_t)m + 1, sizeof(double));
  double* P = (double*)calloc((size_t)m + 1, sizeof(double));
  int status = ORC_NOT_CONVERGED;
  ctx_residual(c, X, B, R);
  /* monolis_set_converge_R (mat_converge.f90:14-49) */
  B2 = orc_inner_product(m, R, R);
  if (B2 == 0.0) {
    memset(X, 0, sizeof(double) * (size_t)m);
    status = ORC_OK;
    goto done;
  }
  for (int iter = 1; iter <= Iarray[I_MAX_ITER]; ++iter) {
    ctx_prec_apply(c, R, Z);
    rho = orc_inner_product(m, R, Z);
    if (iter > 1) {
      const double beta = rho / rho1;
      orc_vec_axpby(m, beta, P, 1.0, Z, P);
    } else {
      orc_vec_copy(m, Z, P);
    }
    ctx_matvec(c, P, Q);
    const double omega = orc_inner_product(m, P, Q);
    const double alpha = rho / omega;
    orc_vec_axpby(m, alpha, P, 1.0, X, X);
    if (iter % iter_RR == 0)
      ctx_residual(c, X, B, R);
    else
      orc_vec_axpby(m, -alpha, Q, 1.0, R, R);
    if (orc_check_converge(m, R, B2, iter, Iarray, Rarray)) {
      status = ORC_OK;
      break;
    }
    rho1 = rho;
  }
done:
  free(R);
  free(Z);
  free(Q);
  free(P);
  return status;
}

/* src/iterative/GropCG.f90 (monolis method 2): Gropp's CG, the precondition
 * of A p overlapped with the dot products; no residual replacement. */
static int orc_gropcg(const orc_ctx* c, const double* B, double* X, int* Iarray,
                      double* Rarray) {
  const int m = c->m;
  double B2;
  if (Iarray[I_IS_INIT_X]) memset(X, 0, sizeof(double) * (size_t)m);
  double* V6[6];
  for (int k = 0; k < 6; ++k) V6[k] = (double*)calloc((size_t)m + 1, sizeof(double));
  double *R = V6[0], *U = V6[1], *V = V6[2], *Q = V6[3], *P = V6[4], *S = V6[5];
  int status = ORC_NOT_CONVERGED;
  ctx_residual(c, X, B, R);
  B2 = orc_inner_product(m, R, R);
  if (B2 == 0.0) {
    memset(X, 0, sizeof(double) * (size_t)m);
    status = ORC_OK;
    goto done;
  }
  ctx_prec_apply(c, R, U);
  orc_vec_copy(m, U, P);
  ctx_matvec(c, P, S);
  double gamma = orc_inner_product(m, R, U);
  for (int iter = 1; iter <= Iarray[I_MAX_ITER]; ++iter) {
    const double delta = orc_inner_product(m, P, S);
    ctx_prec_apply(c, S, Q);
    const double alpha = gamma / delta;
    orc_vec_axpby(m, alpha, P, 1.0, X, X);
    orc_vec_axpby(m, -alpha, S, 1.0, R, R);
    orc_vec_axpby(m, -alpha, Q, 1.0, U, U);
    const double gamma1 = orc_inner_product(m, R, U);
    ctx_matvec(c, U, V);
    const double beta = gamma1 / gamma;
    gamma = gamma1;
    orc_vec_axpby(m, beta, P, 1.0, U, P);
    orc_vec_axpby(m, beta, S, 1.0, V, S);
    if (orc_check_converge(m, R, B2, iter, Iarray, Rarray)) {
      status = ORC_OK;
      break;
    }
  }
done:
  for (int k = 0; k < 6; ++k) free(V6[k]);
  return status;
}

/* src/iterative/PipeCG.f90 (monolis method 3): pipelined PCG (Ghysels and
 * Vanroose), three dots per iteration, true residual every 50 iterations.
 * The scalars follow the reference literally: gamma1 = 1/gamma,
 * alpha1 = 1/alpha, beta = gamma*gamma1, alpha = gamma/(delta - beta*gamma*alpha1). */
static int orc_pipecg(const orc_ctx* c, const double* B, double* X, int* Iarray,
                      double* Rarray) {
  const int m = c->m;
  const int iter_RR = 50;
  double B2;
  if (Iarray[I_IS_INIT_X]) memset(X, 0, sizeof(double) * (size_t)m);
  double* V9[9];
  for (int k = 0; k < 9; ++k) V9[k] = (double*)calloc((size_t)m + 1, sizeof(double));
  double *R = V9[0], *U = V9[1], *V = V9[2], *Q = V9[3], *P = V9[4], *Z = V9[5], *L = V9[6],
         *M = V9[7], *S = V9[8];
  int status = ORC_NOT_CONVERGED;
  ctx_residual(c, X, B, R);
  B2 = orc_inner_product(m, R, R);
  if (B2 == 0.0) {
    memset(X, 0, sizeof(double) * (size_t)m);
    status = ORC_OK;
    goto done;
  }
  ctx_prec_apply(c, R, U);
  ctx_matvec(c, U, V);
  double gamma = 1.0, alpha = 1.0;
  for (int iter = 1; iter <= Iarray[I_MAX_ITER]; ++iter) {
    const double g = orc_inner_product(m, R, U);
    const double d = orc_inner_product(m, V, U);
    ctx_prec_apply(c, V, M);
    ctx_matvec(c, M, L);
    const double gamma1 = 1.0 / gamma;
    const double alpha1 = 1.0 / alpha;
    gamma = g;
    const double delta = d;
    double beta;
    if (1 < iter) {
      beta = gamma * gamma1;
      alpha = gamma / (delta - beta * gamma * alpha1);
    } else {
      alpha = gamma / delta;
      beta = 0.0;
    }
    orc_vec_axpby(m, beta, Z, 1.0, L, Z);
    orc_vec_axpby(m, beta, Q, 1.0, M, Q);
    orc_vec_axpby(m, beta, S, 1.0, V, S);
    orc_vec_axpby(m, beta, P, 1.0, U, P);
    if (iter % iter_RR == 0) {
      orc_vec_axpby(m, alpha, P, 1.0, X, X);
      ctx_residual(c, X, B, R);
      ctx_prec_apply(c, R, U);
      ctx_matvec(c, U, V);
    } else {
      orc_vec_axpby(m, -alpha, S, 1.0, R, R);
      orc_vec_axpby(m, -alpha, Q, 1.0, U, U);
      orc_vec_axpby(m, -alpha, Z, 1.0, V, V);
      orc_vec_axpby(m, alpha, P, 1.0, X, X);
    }
    if (orc_check_converge(m, R, B2, iter, Iarray, Rarray)) {
      status = ORC_OK;
      break;
    }
  }
done:
  for (int k = 0; k < 9; ++k) free(V9[k]);
  return status;
}

/* src/iterative/PipeCR.f90:18-146 (monolis method 4): pipelined preconditioned
 * conjugate residuals, three dots per iteration -- gamma = (v, u),
 * delta = (v, m), U2 = (u, u) -- and the true residual every 50 iterations.
 * The test is on the preconditioned residual, before the updates of the
 * iteration: sqrt(U2/B2) <= tol/phi with phi = sqrt(R2/U2_0) (PipeCR.f90:72-97).
 * The reference writes neither cur_iter nor cur_resid (no
 * monolis_check_converge call); here they hold the loop counter at the test
 * and resid*phi, the quantity of the commented log line PipeCR.f90:96.
 * Reaching max_iter is silent in the reference (ierr stays as it was); here
 * it returns ORC_NOT_CONVERGED like the other methods.  A zero initial
 * residual makes the reference divide 0/0 (phi, then alpha) and return NaN
 * in x; here x = 0 as mat_converge.f90:40-48 does for the other methods. */
static int orc_pipecr(const orc_ctx* c, const double* B, double* X, int* Iarray,
                      double* Rarray) {
  const int m = c->m;
  const int iter_RR = 50;
  if (Iarray[I_IS_INIT_X]) memset(X, 0, sizeof(double) * (size_t)m);
  double* V8[8];
  for (int k = 0; k < 8; ++k) V8[k] = (double*)calloc((size_t)m + 1, sizeof(double));
  double *R = V8[0], *U = V8[1], *V = V8[2], *Q = V8[3], *P = V8[4], *Z = V8[5], *L = V8[6],
         *M = V8[7];
  int status = ORC_NOT_CONVERGED;
  ctx_residual(c, X, B, R);
  const double B2 = orc_inner_product(m, R, R);
  const double R2 = orc_inner_product(m, R, R);
  if (B2 == 0.0) {
    memset(X, 0, sizeof(double) * (size_t)m);
    status = ORC_OK;
    goto done;
  }
  ctx_prec_apply(c, R, U);
  double U2 = orc_inner_product(m, U, U);
  ctx_matvec(c, U, V);
  const double phi = sqrt(R2 / U2);
  const double utol = Rarray[R_TOL] / phi;
  double gamma1 = 0.0, alpha1 = 0.0;
  for (int iter = 1; iter <= Iarray[I_MAX_ITER]; ++iter) {
    ctx_prec_apply(c, V, M);
    const double gamma = orc_inner_product(m, V, U);
    const double delta = orc_inner_product(m, V, M);
    U2 = orc_inner_product(m, U, U);
    ctx_matvec(c, M, L);
    const double resid = sqrt(U2 / B2);
    Iarray[I_CUR_ITER] = iter;
    Rarray[R_CUR_RESID] = resid * phi;
    if (resid <= utol) {
      status = ORC_OK;
      break;
    }
    double alpha, beta;
    if (1 < iter) {
      beta = gamma * gamma1;
      alpha = gamma / (delta - beta * gamma * alpha1);
    } else {
      beta = 0.0;
      alpha = gamma / delta;
    }
    orc_vec_axpby(m, beta, Z, 1.0, L, Z);
    orc_vec_axpby(m, beta, Q, 1.0, M, Q);
    orc_vec_axpby(m, beta, P, 1.0, U, P);
    orc_vec_axpby(m, alpha, P, 1.0, X, X);
    if (iter % iter_RR == 0) {
      ctx_residual(c, X, B, R);
      ctx_prec_apply(c, R, U);
      ctx_matvec(c, U, V);
    } else {
      orc_vec_axpby(m, -alpha, Q, 1.0, U, U);
      orc_vec_axpby(m, -alpha, Z, 1.0, V, V);
    }
    gamma1 = 1.0 / gamma;
    alpha1 = 1.0 / alpha;
  }
done:
  for (int k = 0; k < 8; ++k) free(V8[k]);
  return status;
}

/* src/iterative/BiCGSTAB.f90:18-192 (right preconditioned). */
/* noprec = 1: src/iterative/BiCGSTAB_noprec.f90 (method 7): no preconditioner
 * (whatever Iarray[2] says), no sorted-sum fallback, omega = (t,s)/(t,t) as is,
 * true residual every 50 iterations. */
static int orc_bicgstab(const orc_ctx* c, const double* B, double* X,
                        int* Iarray, double* Rarray, int noprec) {
  const int m = c->m;
  const int iter_RR = noprec ? 50 : 200;
  double alpha = 0.0, beta, rho = 0.0, rho1 = 0.0, c2, omega = 0.0, B2;
  double CG1, CG2;
  if (Iarray[I_IS_INIT_X]) memset(X, 0, sizeof(double) * (size_t)m);
  const size_t sz = (size_t)m + 1;
  double* R = (double*)calloc(sz, sizeof(double));
  double* RT = (double*)calloc(sz, sizeof(double));
  double* P = (double*)calloc(sz, sizeof(double));
  double* PT = (double*)calloc(sz, sizeof(double));
  double* S = (double*)calloc(sz, sizeof(double));
  double* ST = (double*)calloc(sz, sizeof(double));
  double* T = (double*)calloc(sz, sizeof(double));
  double* V = (double*)calloc(sz, sizeof(double));
  int status = ORC_NOT_CONVERGED;
  ctx_residual(c, X, B, R);
  B2 = orc_inner_product(m, R, R);
  if (B2 == 0.0) {
    memset(X, 0, sizeof(double) * (size_t)m);
    status = ORC_OK;
    goto done;
  }
  orc_vec_copy(m, R, RT);
  for (int iter = 1; iter <= Iarray[I_MAX_ITER]; ++iter) {
    rho = orc_inner_product(m, R, RT);
    if (!noprec && fabs(rho) == 0.0) rho = orc_sorted_inner_product(m, R, RT);
    if (iter > 1) {
      beta = (rho / rho1) * (alpha / omega);
      int i;
#pragma omp parallel for schedule(static)
      for (i = 0; i < m; ++i) P[i] = R[i] + beta * (P[i] - omega * V[i]);
    } else {
      orc_vec_copy(m, R, P);
    }
    if (noprec) orc_vec_copy(m, P, PT);
    else ctx_prec_apply(c, P, PT);
    ctx_matvec(c, PT, V);
    c2 = orc_inner_product(m, RT, V);
    if (!noprec && fabs(c2) == 0.0) c2 = orc_sorted_inner_product(m, RT, V);
    alpha = rho / c2;
    orc_vec_axpby(m, -alpha, V, 1.0, R, S);
    if (noprec) orc_vec_copy(m, S, ST);
    else ctx_prec_apply(c, S, ST);
    ctx_matvec(c, ST, T);
    CG1 = orc_inner_product(m, T, S);
    CG2 = orc_inner_product(m, T, T);
    if (noprec) {
      omega = CG1 / CG2;
    } else {
      if (fabs(CG2) == 0.0) CG2 = orc_sorted_inner_product(m, T, T);
      omega = (fabs(CG2) == 0.0) ? 0.0 : CG1 / CG2;
    }
    {
      int i;
#pragma omp parallel for schedule(static)
      for (i = 0; i < m; ++i) X[i] = X[i] + alpha * PT[i] + omega * ST[i];
    }
    if (iter % iter_RR == 0)
      ctx_residual(c, X, B, R);
    else
      orc_vec_axpby(m, -omega, T, 1.0, S, R);
    if (orc_check_converge(m, R, B2, iter, Iarray, Rarray)) {
      status = ORC_OK;
      break;
    }
    rho1 = rho;
  }
done:
  free(R);
  free(RT);
  free(P);
  free(PT);
  free(S);
  free(ST);
  free(T);
  free(V);
  return status;
}

/* ---------------------------------------------------------------------------
 * monolis_solve_main_R (src/solver/solver.f90:77-185) on one rank, with the
 * same argument list as the flat C entry point monolis_solve_R_c_main
 * (wrapper/solver/monolis_solver_c.h:51-80) minus the communication tables.
 * Instead of stopping the process on non-convergence it returns
 * ORC_NOT_CONVERGED and writes it to Iarray[5].
 * ------------------------------------------------------------------------- */
/* src/iterative/PipeBiCGSTAB.f90 (method 6) and, with noprec = 1,
 * PipeBiCGSTAB_noprec.f90 (method 8: M = I and omega1 = QY/YY without the
 * zero guard).  Pipelined BiCGSTAB with one fused reduction block per half
 * step; true residual every 200 iterations.  The residual-replacement branch
 * updates X with omega1*T as the reference does. */
static int orc_pipebicgstab(const orc_ctx* c, const double* B, double* X, int* Iarray,
                            double* Rarray, int noprec) {
  const int m = c->m;
  const int iter_RR = 200;
  double B2;
  if (Iarray[I_IS_INIT_X]) memset(X, 0, sizeof(double) * (size_t)m);
  double* W15[15];
  for (int k = 0; k < 15; ++k) W15[k] = (double*)calloc((size_t)m + 1, sizeof(double));
  double *R = W15[0], *R0 = W15[1], *RT = W15[2], *W0 = W15[3], *WT = W15[4], *T = W15[5],
         *PT = W15[6], *S = W15[7], *ST = W15[8], *Z = W15[9], *ZT = W15[10], *Q = W15[11],
         *QT = W15[12], *Y = W15[13], *V = W15[14];
  int status = ORC_NOT_CONVERGED;
#define PB_PREC(src, dst) \
  do {                    \
    if (noprec)           \
      orc_vec_copy(m, src, dst); \
    else                  \
      ctx_prec_apply(c, src, dst); \
  } while (0)
  ctx_residual(c, X, B, R);
  B2 = orc_inner_product(m, R, R);
  if (B2 == 0.0) {
    memset(X, 0, sizeof(double) * (size_t)m);
    status = ORC_OK;
    goto done;
  }
  orc_vec_copy(m, R, R0);
  PB_PREC(R, RT);
  ctx_matvec(c, RT, W0);
  PB_PREC(W0, WT);
  ctx_matvec(c, WT, T);
  double RR = orc_inner_product(m, R, R);
  const double RW0 = orc_inner_product(m, R, W0);
  double alpha = RR / RW0, beta = 0.0, omega = 0.0;
  for (int iter = 1; iter <= Iarray[I_MAX_ITER]; ++iter) {
    int i;
#pragma omp parallel for schedule(static)
    for (i = 0; i < m; ++i) {
      PT[i] = RT[i] + beta * (PT[i] - omega * ST[i]);
      S[i] = W0[i] + beta * (S[i] - omega * Z[i]);
      ST[i] = WT[i] + beta * (ST[i] - omega * ZT[i]);
      Z[i] = T[i] + beta * (Z[i] - omega * V[i]);
      Q[i] = R[i] - alpha * S[i];
      QT[i] = RT[i] - alpha * ST[i];
      Y[i] = W0[i] - alpha * Z[i];
    }
    const double QY = orc_inner_product(m, Q, Y);
    const double YY = orc_inner_product(m, Y, Y);
    PB_PREC(Z, ZT);
    ctx_matvec(c, ZT, V);
    const double omega1 = (noprec || YY != 0.0) ? QY / YY : 0.0;
    if (iter % iter_RR == 0) {
#pragma omp parallel for schedule(static)
      for (i = 0; i < m; ++i) X[i] = X[i] + alpha * PT[i] + omega1 * T[i];
      ctx_residual(c, X, B, R);
      PB_PREC(R, RT);
      ctx_matvec(c, RT, W0);
    } else {
#pragma omp parallel for schedule(static)
      for (i = 0; i < m; ++i) {
        X[i] = X[i] + alpha * PT[i] + omega1 * QT[i];
        R[i] = Q[i] - omega1 * Y[i];
        RT[i] = QT[i] - omega1 * (WT[i] - alpha * ZT[i]);
        W0[i] = Y[i] - omega1 * (T[i] - alpha * V[i]);
      }
    }
    const double RR1 = orc_inner_product(m, R0, R);
    const double RW = orc_inner_product(m, R0, W0);
    const double RS = orc_inner_product(m, R0, S);
    const double RZ = orc_inner_product(m, R0, Z);
    PB_PREC(W0, WT);
    ctx_matvec(c, WT, T);
    if (orc_check_converge(m, R, B2, iter, Iarray, Rarray)) {
      status = ORC_OK;
      break;
    }
    beta = (alpha * RR1) / (RR * omega1);
    alpha = RR1 / (RW + beta * RS - beta * omega1 * RZ);
    omega = omega1;
    RR = RR1;
  }
#undef PB_PREC
done:
  for (int k = 0; k < 15; ++k) free(W15[k]);
  return status;
}

int orc_solve(int n, int np, int nz, int n_dof, const int* n_dof_list,
              const int* n_dof_index, const int* n_dof_index2, const double* A,
              double* x, const double* b, const int* index, const int* item,
              int* Iarray, double* Rarray) {
  (void)np;
  (void)nz;
  const double t0 = orc_now();
  orc_ctx c;
  memset(&c, 0, sizeof(c));
  c.N = n;
  c.ndof = n_dof;
  c.m = (n_dof == -1) ? n_dof_index[n] : n * n_dof;
  c.n_dof_list = n_dof_list;
  c.n_dof_index = n_dof_index;
  c.n_dof_index2 = n_dof_index2;
  c.index = index;
  c.item = item;
  c.A = A;
  c.precond = Iarray[I_PRECOND];
  for (int k = R_TIME_SOL; k <= 9; ++k) Rarray[k] = 0.0;
  if (c.precond != ORC_PREC_NONE && c.precond != ORC_PREC_DIAG) {
    Iarray[I_IERR] = ORC_BAD_PRECOND;
    return ORC_BAD_PRECOND;
  }
  const double tp = orc_now();
  if (c.precond == ORC_PREC_DIAG) {
    if (n_dof == -1) {
      long long total = 0;
      for (int i = 0; i < n; ++i) total += (long long)n_dof_list[i] * n_dof_list[i];
      c.doff = (long long*)malloc(sizeof(long long) * ((size_t)n + 1));
      c.D = (double*)malloc(sizeof(double) * (size_t)(total > 0 ? total : 1));
      orc_diag_V_setup(n, n_dof_list, n_dof_index2, index, item, A, c.doff, c.D);
    } else {
      c.D = (double*)malloc(sizeof(double) * (size_t)n_dof * n_dof * (n > 0 ? n : 1));
      if (n_dof == 3)
        orc_diag_33_setup(n, index, item, A, c.D);
      else
        orc_diag_nn_setup(n, n_dof, index, item, A, c.D);
    }
  }
  Rarray[R_TIME_PREP] = orc_now() - tp;
  int st;
  if (Iarray[I_METHOD] == ORC_ITER_CG)
    st = orc_cg(&c, b, x, Iarray, Rarray);
  else if (Iarray[I_METHOD] == ORC_ITER_GROPCG)
    st = orc_gropcg(&c, b, x, Iarray, Rarray);
  else if (Iarray[I_METHOD] == ORC_ITER_PIPECG)
    st = orc_pipecg(&c, b, x, Iarray, Rarray);
  else if (Iarray[I_METHOD] == ORC_ITER_PIPECR)
    st = orc_pipecr(&c, b, x, Iarray, Rarray);
  else if (Iarray[I_METHOD] == ORC_ITER_BICGSTAB)
    st = orc_bicgstab(&c, b, x, Iarray, Rarray, 0);
  else if (Iarray[I_METHOD] == ORC_ITER_BICGSTAB_NOPREC)
    st = orc_bicgstab(&c, b, x, Iarray, Rarray, 1);
  else if (Iarray[I_METHOD] == ORC_ITER_PIPEBICGSTAB)
    st = orc_pipebicgstab(&c, b, x, Iarray, Rarray, 0);
  else if (Iarray[I_METHOD] == ORC_ITER_PIPEBICGSTAB_NOPREC)
    st = orc_pipebicgstab(&c, b, x, Iarray, Rarray, 1);
  else
    st = ORC_BAD_METHOD;
  free(c.D);
  free(c.doff);
  Iarray[I_IERR] = st;
  Rarray[R_TIME_SOL] = orc_now() - t0;
  return st;
}

/* Timing helper for the CPU baseline: `reps` SpMVs, returns seconds. */
double orc_time_matvec(int reps, int N, int ndof, const int* index,
                       const int* item, const double* A, const double* X,
                       double* Y) {
  const double t0 = orc_now();
  for (int r = 0; r < reps; ++r) orc_matvec(N, ndof, NULL, NULL, NULL, index, item, A, X, Y);
  return orc_now() - t0;
}
