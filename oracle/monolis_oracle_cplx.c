/*
 * monolis_oracle_cplx.c -- CPU restatement of monolis's complex (COCG) path.
 *
 * TEST INFRASTRUCTURE ONLY, like monolis_oracle.c: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Complex values are C99 `double complex` (two doubles, real then imaginary:
 * the layout of Fortran complex(kdouble) and of numpy complex128).  Matrix
 * layout is that of the real path: `index` 0-based row offsets (NP+1), `item`
 * 1-based block columns, A row-major per block.
 *
 * Pinning (tests/test_oracle_cocg.py):
 *   orc_matvec_C          dense complex brute force; real-valued input equals
 *                         the real oracle; i*A x = A (i x) (complex linearity).
 *   orc_inner_product_C   closed form (unconjugated: (i, i) = -1 per entry,
 *                         which a conjugating dot would report as +1).
 *   orc_diag_nn_setup_C / apply_C   per-block numpy.linalg.solve; zero pivot
 *                         reported, as the reference stops on it.
 *   orc_check_converge_C  resid = sqrt(R2 / B2) closed forms incl. a negative
 *                         R2/B2, whose principal root is purely imaginary.
 *   orc_solve_C (COCG)    the reference's COCG_test chain (x = 1 + i), dense
 *                         numpy solves, and the real-valued case equal to CG.
 *   orc_matvec_V_C, orc_diag_V_setup_C / apply_C, orc_solve_V_C (variable
 *                         blocks, n_dof = -1): dense brute force; every block
 *                         of one size b reproduces the fixed-b routines bit for
 *                         bit (same sums in the same order); real-valued data
 *                         reproduces the real variable-block oracle.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define I_METHOD 1
#define I_PRECOND 2
#define I_MAX_ITER 3
#define I_CUR_ITER 4
#define I_IERR 5
#define I_IS_INIT_X 10
#define R_TOL 1
#define R_CUR_RESID 2
#define R_TIME_SOL 3
#define R_TIME_PREP 4

#define ORC_ITER_COCG 13
#define ORC_PREC_NONE 0
#define ORC_PREC_DIAG 1

#define ORC_OK 0
#define ORC_NOT_CONVERGED 1
#define ORC_BAD_METHOD -2
#define ORC_BAD_PRECOND -3
#define ORC_ZERO_DIAG -6

typedef double complex zc;

static double orc_now_c(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* SpMV, src/linalg/matvec.f90:656-703 (matvec_nn_C; matvec_11_C and
 * matvec_33_C, :745-779 and :830-873, are the same sums unrolled; dispatch
 * at :192-227):
 * Y(k) = sum over the row's blocks, then l = 1..NDOF, of A(k, l) * X(l). */
void orc_matvec_C(int N, int ndof, const int* index, const int* item, const zc* A,
                  const zc* X, zc* Y) {
  const int nd2 = ndof * ndof;
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    zc yt[64];
    for (int k = 0; k < ndof; ++k) yt[k] = 0.0;
    for (int j = index[i]; j < index[i + 1]; ++j) {
      const int in = item[j] - 1;
      for (int k = 0; k < ndof; ++k)
        for (int l = 0; l < ndof; ++l)
          yt[k] = yt[k] + A[(size_t)nd2 * j + ndof * k + l] * X[(size_t)ndof * in + l];
    }
    for (int k = 0; k < ndof; ++k) Y[(size_t)ndof * i + k] = yt[k];
  }
}

/* src/linalg/inner_product.f90:457-499: sum = sum + X(i)*Y(i) -- the
 * unconjugated bilinear form COCG is built on. */
void orc_inner_product_C(int m, const zc* X, const zc* Y, zc* out) {
  double sr = 0.0, si = 0.0;
  int i;
#pragma omp parallel for schedule(static) reduction(+ : sr, si)
  for (i = 0; i < m; ++i) {
    const zc p = X[i] * Y[i];
    sr += creal(p);
    si += cimag(p);
  }
  *out = sr + si * I;
}

/* src/linalg/vec_util.f90:166-196 (AXPBY_C): Z = alpha*X + beta*Y. */
static void axpby_c(int m, zc alpha, const zc* X, zc beta, const zc* Y, zc* Z) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < m; ++i) Z[i] = alpha * X[i] + beta * Y[i];
}

/* src/linalg/mat_converge.f90:145-196: resid = sqrt(R2/B2) in complex
 * arithmetic (principal root), cur_resid = resid stored into a real slot
 * (Fortran assignment keeps the real part), converged when |resid| < tol. */
int orc_check_converge_C(int m, const zc* R, const zc* B2p, int iter, int* Iarray, double* Rarray) {
  const zc B2 = *B2p;
  zc R2;
  orc_inner_product_C(m, R, R, &R2);
  const zc resid = csqrt(R2 / B2);
  Iarray[I_CUR_ITER] = iter;
  Rarray[R_CUR_RESID] = creal(resid);
  return cabs(resid) < Rarray[R_TOL];
}

/* src/prec/diag/diag_nn.f90:90-159 (monolis_precond_diag_nn_setup_C; every
 * block size, 3 included, goes through the nn form for complex values):
 * in-place no-pivot LU of each row's diagonal block with reciprocal pivots;
 * rows without a diagonal block keep zeros.  A zero pivot stops the reference;
 * here it returns ORC_ZERO_DIAG. */
int orc_diag_nn_setup_C(int N, int ndof, const int* index, const int* item, const zc* A, zc* D) {
  const int nd2 = ndof * ndof;
  memset(D, 0, sizeof(zc) * (size_t)nd2 * (size_t)(N > 0 ? N : 1));
  for (int i = 0; i < N; ++i) {
    for (int jj = index[i]; jj < index[i + 1]; ++jj) {
      if (item[jj] - 1 != i) continue;
      zc* LU = D + (size_t)nd2 * i;
      for (int e = 0; e < nd2; ++e) LU[e] = A[(size_t)nd2 * jj + e];
      for (int k = 0; k < ndof; ++k) {
        if (LU[ndof * k + k] == 0.0) return ORC_ZERO_DIAG;
        LU[ndof * k + k] = 1.0 / LU[ndof * k + k];
        for (int l = k + 1; l < ndof; ++l) {
          LU[ndof * l + k] = LU[ndof * l + k] * LU[ndof * k + k];
          for (int j = k + 1; j < ndof; ++j) LU[ndof * l + j] = LU[ndof * l + j] - LU[ndof * l + k] * LU[ndof * k + j];
        }
      }
    }
  }
  return ORC_OK;
}

/* src/prec/diag/diag_nn.f90:235-284 (monolis_precond_diag_nn_apply_C). */
void orc_diag_nn_apply_C(int N, int ndof, const zc* D, const zc* X, zc* Y) {
  const int nd2 = ndof * ndof;
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    zc t[64];
    const zc* LU = D + (size_t)nd2 * i;
    for (int j = 0; j < ndof; ++j) t[j] = X[(size_t)ndof * i + j];
    for (int j = 1; j < ndof; ++j)
      for (int k = 0; k < j; ++k) t[j] = t[j] - LU[ndof * j + k] * t[k];
    for (int j = ndof - 1; j >= 0; --j) {
      for (int k = ndof - 1; k > j; --k) t[j] = t[j] - LU[ndof * j + k] * t[k];
      t[j] = LU[(ndof + 1) * j] * t[j];
    }
    for (int k = 0; k < ndof; ++k) Y[(size_t)ndof * i + k] = t[k];
  }
}

/* src/linalg/matvec.f90:293-351 (monolis_matvec_V_C): block (i, j) of size
 * n_i x n_j row-major at A(n_dof_index2(e)); Y(n_dof_index(i) + k) = sum over
 * the row's blocks, then l = 1..n_j, of A * X. */
void orc_matvec_V_C(int N, const int* n_dof_list, const int* n_dof_index, const int* n_dof_index2,
                    const int* index, const int* item, const zc* A, const zc* X, zc* Y) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    zc yt[64];
    const int n1 = n_dof_list[i];
    for (int k = 0; k < n1; ++k) yt[k] = 0.0;
    for (int j = index[i]; j < index[i + 1]; ++j) {
      const int in = item[j] - 1;
      const int n2 = n_dof_list[in];
      const int kn = n_dof_index[in];
      const long long an = n_dof_index2[j];
      for (int k = 0; k < n1; ++k)
        for (int l = 0; l < n2; ++l) yt[k] = yt[k] + A[an + (long long)n2 * k + l] * X[kn + l];
    }
    for (int k = 0; k < n1; ++k) Y[n_dof_index[i] + k] = yt[k];
  }
}

/* src/prec/diag/diag_V.f90:88-158 (monolis_precond_diag_V_setup_C): the
 * no-pivot LU of each diagonal block with reciprocal pivots, stored at
 * doff[i] (doff = prefix sums of n_i^2).  A zero pivot stops the reference
 * (:123); here it returns ORC_ZERO_DIAG.  A row without a diagonal block keeps
 * zeros (the reference's packed offset does not advance for such a row, which
 * misaligns every later row's factors; such matrices are outside its contract). */
int orc_diag_V_setup_C(int N, const int* n_dof_list, const int* n_dof_index2, const int* index,
                       const int* item, const zc* A, long long* doff, zc* D) {
  doff[0] = 0;
  for (int i = 0; i < N; ++i) doff[i + 1] = doff[i] + (long long)n_dof_list[i] * n_dof_list[i];
  memset(D, 0, sizeof(zc) * (size_t)(doff[N] > 0 ? doff[N] : 1));
  for (int i = 0; i < N; ++i) {
    for (int jj = index[i]; jj < index[i + 1]; ++jj) {
      if (item[jj] - 1 != i) continue;
      const int nd = n_dof_list[i];
      zc* LU = D + doff[i];
      for (int e = 0; e < nd * nd; ++e) LU[e] = A[(long long)n_dof_index2[jj] + e];
      for (int k = 0; k < nd; ++k) {
        if (LU[nd * k + k] == 0.0) return ORC_ZERO_DIAG;
        LU[nd * k + k] = 1.0 / LU[nd * k + k];
        for (int l = k + 1; l < nd; ++l) {
          LU[nd * l + k] = LU[nd * l + k] * LU[nd * k + k];
          for (int j = k + 1; j < nd; ++j) LU[nd * l + j] = LU[nd * l + j] - LU[nd * l + k] * LU[nd * k + j];
        }
      }
    }
  }
  return ORC_OK;
}

/* src/prec/diag/diag_V.f90:208-250 (monolis_precond_diag_V_apply_C). */
void orc_diag_V_apply_C(int N, const int* n_dof_list, const int* n_dof_index, const long long* doff, const zc* D,
                        const zc* X, zc* Y) {
  int i;
#pragma omp parallel for schedule(static)
  for (i = 0; i < N; ++i) {
    zc t[64];
    const int nd = n_dof_list[i];
    const int in = n_dof_index[i];
    const zc* LU = D + doff[i];
    for (int j = 0; j < nd; ++j) t[j] = X[in + j];
    for (int j = 1; j < nd; ++j)
      for (int k = 0; k < j; ++k) t[j] = t[j] - LU[nd * j + k] * t[k];
    for (int j = nd - 1; j >= 0; --j) {
      for (int k = nd - 1; k > j; --k) t[j] = t[j] - LU[nd * j + k] * t[k];
      t[j] = LU[(nd + 1) * j] * t[j];
    }
    for (int j = 0; j < nd; ++j) Y[in + j] = t[j];
  }
}

typedef struct {
  int N, ndof, m, precond;
  const int *index, *item;
  const int *n_dof_list, *n_dof_index, *n_dof_index2;  /* n_dof = -1 only */
  const zc* A;
  zc* D;
  long long* doff;                                      /* n_dof = -1 only */
} orc_zctx;

static void zctx_matvec(const orc_zctx* c, const zc* X, zc* Y) {
  if (c->ndof == -1)
    orc_matvec_V_C(c->N, c->n_dof_list, c->n_dof_index, c->n_dof_index2, c->index, c->item, c->A, X, Y);
  else
    orc_matvec_C(c->N, c->ndof, c->index, c->item, c->A, X, Y);
}

static void zctx_residual(const orc_zctx* c, const zc* X, const zc* B, zc* R) {
  /* src/linalg/matvec.f90:907-933: R = B - A X */
  zctx_matvec(c, X, R);
  axpby_c(c->m, 1.0, B, -1.0, R, R);
}

static void zctx_prec(const orc_zctx* c, const zc* X, zc* Y) {
  /* src/prec/precond.f90:235-286 (DIAG or NONE) */
  if (c->precond == ORC_PREC_DIAG && c->ndof == -1)
    orc_diag_V_apply_C(c->N, c->n_dof_list, c->n_dof_index, c->doff, c->D, X, Y);
  else if (c->precond == ORC_PREC_DIAG)
    orc_diag_nn_apply_C(c->N, c->ndof, c->D, X, Y);
  else
    memcpy(Y, X, sizeof(zc) * (size_t)c->m);
}

/* src/iterative/COCG.f90:18-105: preconditioned conjugate orthogonal CG, the
 * CG recurrence with the unconjugated bilinear form; true residual every 200
 * iterations. */
static int orc_cocg(const orc_zctx* c, const zc* B, zc* X, int* Iarray, double* Rarray) {
  const int m = c->m;
  const int iter_RR = 200;
  zc rho = 0.0, rho1 = 0.0, B2;
  if (Iarray[I_IS_INIT_X]) memset(X, 0, sizeof(zc) * (size_t)m);
  zc* R = (zc*)calloc((size_t)m + 1, sizeof(zc));
  zc* Z = (zc*)calloc((size_t)m + 1, sizeof(zc));
  zc* Q = (zc*)calloc((size_t)m + 1, sizeof(zc));
  zc* P = (zc*)calloc((size_t)m + 1, sizeof(zc));
  int status = ORC_NOT_CONVERGED;
  zctx_residual(c, X, B, R);
  /* monolis_set_converge_C (mat_converge.f90:53-86), called with R */
  orc_inner_product_C(m, R, R, &B2);
  if (B2 == 0.0) {
    memset(X, 0, sizeof(zc) * (size_t)m);
    status = ORC_OK;
    goto done;
  }
  for (int iter = 1; iter <= Iarray[I_MAX_ITER]; ++iter) {
    zctx_prec(c, R, Z);
    orc_inner_product_C(m, R, Z, &rho);
    if (iter > 1) {
      const zc beta = rho / rho1;
      axpby_c(m, beta, P, 1.0, Z, P);
    } else {
      memcpy(P, Z, sizeof(zc) * (size_t)m);
    }
    zctx_matvec(c, P, Q);
    zc omega;
    orc_inner_product_C(m, P, Q, &omega);
    const zc alpha = rho / omega;
    axpby_c(m, alpha, P, 1.0, X, X);
    if (iter % iter_RR == 0)
      zctx_residual(c, X, B, R);
    else
      axpby_c(m, -alpha, Q, 1.0, R, R);
    if (orc_check_converge_C(m, R, &B2, iter, Iarray, Rarray)) {
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

/* monolis_solve_main_C (src/solver/solver.f90:189-213) on one rank: the
 * preconditioner setup of precond.f90:105-143 (diag.f90:40-58: diag_V for
 * n_dof = -1, diag_nn otherwise), then monolis_solver_select_C, which accepts
 * COCG only. */
static int solve_C_ctx(orc_zctx* c, zc* x, const zc* b, int* Iarray, double* Rarray) {
  const double t0 = orc_now_c();
  for (int k = R_TIME_SOL; k <= 9; ++k) Rarray[k] = 0.0;
  if (Iarray[I_METHOD] != ORC_ITER_COCG) {
    Iarray[I_IERR] = ORC_BAD_METHOD;
    return ORC_BAD_METHOD;
  }
  if (Iarray[I_PRECOND] != ORC_PREC_NONE && Iarray[I_PRECOND] != ORC_PREC_DIAG) {
    Iarray[I_IERR] = ORC_BAD_PRECOND;
    return ORC_BAD_PRECOND;
  }
  c->precond = Iarray[I_PRECOND];
  int st = ORC_OK;
  const double tp = orc_now_c();
  if (c->precond == ORC_PREC_DIAG) {
    if (c->ndof == -1) {
      long long tot = 0;
      for (int i = 0; i < c->N; ++i) tot += (long long)c->n_dof_list[i] * c->n_dof_list[i];
      c->doff = (long long*)malloc(sizeof(long long) * ((size_t)c->N + 1));
      c->D = (zc*)malloc(sizeof(zc) * (size_t)(tot > 0 ? tot : 1));
      st = orc_diag_V_setup_C(c->N, c->n_dof_list, c->n_dof_index2, c->index, c->item, c->A, c->doff, c->D);
    } else {
      c->D = (zc*)malloc(sizeof(zc) * (size_t)c->ndof * c->ndof * (size_t)(c->N > 0 ? c->N : 1));
      st = orc_diag_nn_setup_C(c->N, c->ndof, c->index, c->item, c->A, c->D);
    }
  }
  Rarray[R_TIME_PREP] = orc_now_c() - tp;
  if (st == ORC_OK) st = orc_cocg(c, b, x, Iarray, Rarray);
  free(c->D);
  free(c->doff);
  Iarray[I_IERR] = st;
  Rarray[R_TIME_SOL] = orc_now_c() - t0;
  return st;
}

int orc_solve_C(int n, int n_dof, const int* index, const int* item, const zc* A, zc* x, const zc* b,
                int* Iarray, double* Rarray) {
  orc_zctx c;
  memset(&c, 0, sizeof(c));
  c.N = n;
  c.ndof = n_dof;
  c.m = n * n_dof;
  c.index = index;
  c.item = item;
  c.A = A;
  return solve_C_ctx(&c, x, b, Iarray, Rarray);
}

/* The same for variable block sizes (monolis n_dof = -1). */
int orc_solve_V_C(int n, const int* n_dof_list, const int* n_dof_index, const int* n_dof_index2, const int* index,
                  const int* item, const zc* A, zc* x, const zc* b, int* Iarray, double* Rarray) {
  orc_zctx c;
  memset(&c, 0, sizeof(c));
  c.N = n;
  c.ndof = -1;
  c.m = n_dof_index[n];
  c.index = index;
  c.item = item;
  c.n_dof_list = n_dof_list;
  c.n_dof_index = n_dof_index;
  c.n_dof_index2 = n_dof_index2;
  c.A = A;
  return solve_C_ctx(&c, x, b, Iarray, Rarray);
}
