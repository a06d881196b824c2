/*
 * A C program written against monolis's C API (include/monolis.h), linked
 * against libmonolis_front.so only: the way an unmodified monolis C
 * application uses the GPU path.
 *
 *   test_front host   pattern / assembly / Dirichlet / parameter checks (no GPU):
 *                     the reference's unit-test answers
 *                     (src_test/matrix/spmat_nzpattern_test.f90:38-73,
 *                     spmat_handler_test.f90:46-132) and index invariants
 *   test_front gpu    matrix-vector products against a host loop, the chain
 *                     solves of wrapper_test/solver/monolis_solver_c_test.c
 *                     (x = 1, x = 1 + i) for every method, variable blocks
 *   test_front err_*  one error monolis stops on (the process exits 1 with a
 *                     message): err_pattern (entry outside the pattern),
 *                     err_dof (DOF outside its block), err_dbc (Dirichlet
 *                     condition on a row without its diagonal block), and two
 *                     that would write outside monolis's arrays: err_kind
 *                     (complex entry into a real pattern), err_matval
 *                     (set_matrix_BCSR_mat_val with the wrong nz)
 *
 * Prints one line per failed check and exits 1 if any failed.
 */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "monolis.h"

static int fails = 0;
#define CHECK(cond, ...)                      \
  do {                                        \
    if (!(cond)) {                            \
      printf("FAIL %s:%d: ", __FILE__, __LINE__); \
      printf(__VA_ARGS__);                    \
      printf("\n");                           \
      ++fails;                                \
    }                                         \
  } while (0)

static unsigned long long lcg = 88172645463325252ull;
static double urand(void) {
  lcg ^= lcg << 13;
  lcg ^= lcg >> 7;
  lcg ^= lcg << 17;
  return (double)(lcg >> 11) / 9007199254740992.0;
}

/* elements of the n-node chain: (k, k+1) */
static int** chain_elems(int n_node) {
  int** e = (int**)malloc(sizeof(int*) * (size_t)(n_node - 1));
  for (int k = 0; k < n_node - 1; ++k) {
    e[k] = (int*)malloc(sizeof(int) * 2);
    e[k][0] = k;
    e[k][1] = k + 1;
  }
  return e;
}
static void free_elems(int** e, int n) {
  for (int k = 0; k < n; ++k) free(e[k]);
  free(e);
}

static void host_checks(void) {
  MONOLIS mat;
  monolis_initialize(&mat);
  /* defaults of def_solver_prm.f90:183-222 */
  CHECK(mat.prm.Iarray[MONOLIS_PRM_I_METHOD] == 1 && mat.prm.Iarray[MONOLIS_PRM_I_PRECOND] == 1 &&
            mat.prm.Iarray[MONOLIS_PRM_I_MAX_ITER] == 1000 && mat.prm.Iarray[MONOLIS_PRM_I_IERR] == -1 &&
            mat.prm.Iarray[MONOLIS_PRM_I_IS_INIT_X] == 1 && mat.prm.Iarray[MONOLIS_PRM_I_IS_ERROR_ABORT] == 1 &&
            mat.prm.Rarray[MONOLIS_PRM_R_TOL] == 1.0e-8,
        "prm defaults");
  monolis_set_method(&mat, MONOLIS_ITER_BICGSTAB);
  monolis_set_tolerance(&mat, 1e-10);
  monolis_set_init_x(&mat, false);
  CHECK(mat.prm.Iarray[1] == 5 && mat.prm.Rarray[1] == 1e-10 && mat.prm.Iarray[10] == 0, "setters");

  /* spmat_nzpattern_test.f90:38-73: 4-node chain, n_dof 2 */
  int** e = chain_elems(4);
  monolis_get_nonzero_pattern_by_simple_mesh_R(&mat, 4, 2, 2, 3, e);
  const int idx[5] = {0, 2, 5, 8, 10}, itm[10] = {1, 2, 1, 2, 3, 2, 3, 4, 3, 4};
  CHECK(mat.mat.N == 4 && mat.mat.NP == 4 && mat.mat.NDOF == 2, "sizes");
  for (int i = 0; i < 5; ++i) CHECK(mat.mat.CSR.index[i] == idx[i], "index[%d] = %d", i, mat.mat.CSR.index[i]);
  for (int i = 0; i < 10; ++i) CHECK(mat.mat.CSR.item[i] == itm[i], "item[%d] = %d", i, mat.mat.CSR.item[i]);
  for (int i = 0; i <= 10; ++i) CHECK(mat.mat.n_dof_index2[i] == 4 * i, "n_dof_index2");
  /* CSC view: column c lists its rows ascending, perm points back into CSR */
  for (int c = 0; c < 4; ++c)
    for (int p = mat.mat.CSC.index[c]; p < mat.mat.CSC.index[c + 1]; ++p) {
      const int r = mat.mat.CSC.item[p] - 1, pos = mat.mat.CSC.perm[p] - 1;
      CHECK(mat.mat.CSR.item[pos] == c + 1 && pos >= mat.mat.CSR.index[r] && pos < mat.mat.CSR.index[r + 1], "csc");
      if (p > mat.mat.CSC.index[c]) CHECK(mat.mat.CSC.item[p - 1] < mat.mat.CSC.item[p], "csc order");
    }

  /* spmat_handler_test.f90:46-104: set, add, then a one-node element on node 1 */
  monolis_set_scalar_to_sparse_matrix_R(&mat, 1, 1, 0, 0, 1.0);
  monolis_set_scalar_to_sparse_matrix_R(&mat, 1, 1, 0, 1, 2.0);
  monolis_set_scalar_to_sparse_matrix_R(&mat, 1, 1, 1, 0, 3.0);
  monolis_set_scalar_to_sparse_matrix_R(&mat, 1, 1, 1, 1, 4.0);
  monolis_add_scalar_to_sparse_matrix_R(&mat, 1, 1, 0, 0, 1.0);
  int conn[1] = {1};
  double r0[2] = {1.0, 2.0}, r1[2] = {3.0, 4.0};
  double* ke[2] = {r0, r1};
  monolis_add_matrix_to_sparse_matrix_R(&mat, 1, conn, ke);
  const double want[4] = {3.0, 4.0, 6.0, 8.0};
  for (int k = 0; k < 4; ++k) CHECK(mat.mat.R.A[12 + k] == want[k], "A(%d) = %g", 13 + k, mat.mat.R.A[12 + k]);
  double v = 0.0;
  bool found = false;
  monolis_get_scalar_from_sparse_matrix_R(&mat, 1, 1, 1, 1, &v, &found);
  CHECK(found && v == 8.0, "get_scalar");
  monolis_get_scalar_from_sparse_matrix_R(&mat, 0, 3, 0, 0, &v, &found);
  CHECK(!found && v == 0.0, "get_scalar outside the pattern");

  /* spmat_handler_test.f90:119-132: A = 2, B = 0, condition on node 1, DOF 0, value 1 */
  for (int k = 0; k < 40; ++k) mat.mat.R.A[k] = 2.0;
  double B[8] = {0};
  monolis_set_Dirichlet_bc_R(&mat, B, 1, 0, 1.0);
  const double wantB[8] = {-2.0, -2.0, 1.0, -2.0, -2.0, -2.0, 0.0, 0.0};
  for (int k = 0; k < 8; ++k) CHECK(B[k] == wantB[k], "B[%d] = %g", k, B[k]);

  /* the same pattern from the connectivity and from the nodal graph */
  MONOLIS m2, m3;
  monolis_initialize(&m2);
  monolis_initialize(&m3);
  int cidx[4] = {0, 2, 4, 6}, citm[6] = {0, 1, 1, 2, 2, 3};
  monolis_get_nonzero_pattern_by_connectivity_R(&m2, 4, 2, 3, cidx, citm);
  int gidx[5] = {0, 1, 3, 5, 6}, gitm[6] = {1, 0, 2, 1, 3, 2};
  monolis_get_nonzero_pattern_by_nodal_graph_R(&m3, 4, 2, gidx, gitm);
  for (int i = 0; i < 10; ++i) CHECK(m2.mat.CSR.item[i] == itm[i] && m3.mat.CSR.item[i] == itm[i], "graph pattern");
  monolis_finalize(&m2);
  monolis_finalize(&m3);

  /* variable blocks: node DOFs 1, 2, 1, 2 */
  MONOLIS mv;
  monolis_initialize(&mv);
  int lst[4] = {1, 2, 1, 2};
  monolis_get_nonzero_pattern_by_simple_mesh_V_R(&mv, 4, 2, lst, 3, e);
  const int di[5] = {0, 1, 3, 4, 6};
  for (int i = 0; i < 5; ++i) CHECK(mv.mat.n_dof_index[i] == di[i], "V n_dof_index");
  CHECK(mv.mat.NDOF == -1 && mv.mat.n_dof_index2[10] == 1 + 2 + 2 + 4 + 2 + 2 + 1 + 2 + 2 + 4, "V n_dof_index2");
  monolis_set_scalar_to_sparse_matrix_R(&mv, 1, 2, 1, 0, 5.0);  /* block (1,2): 2 x 1, entry (1,0) */
  monolis_get_scalar_from_sparse_matrix_R(&mv, 1, 2, 1, 0, &v, &found);
  CHECK(found && v == 5.0 && mv.mat.R.A[mv.mat.n_dof_index2[4] + 1] == 5.0, "V set/get");
  int ns, nps, nzs;
  monolis_get_vec_size(4, 4, -1, 10, mv.mat.n_dof_index, mv.mat.n_dof_index2, &ns, &nps, &nzs);
  CHECK(ns == 6 && nps == 6 && nzs == mv.mat.n_dof_index2[10], "V vec size");
  monolis_finalize(&mv);

  /* complex pattern and scalar access */
  MONOLIS mc;
  monolis_initialize(&mc);
  monolis_get_nonzero_pattern_by_simple_mesh_C(&mc, 4, 2, 2, 3, e);
  monolis_add_scalar_to_sparse_matrix_C(&mc, 2, 3, 1, 0, 1.0 + 2.0 * I);
  monolis_add_scalar_to_sparse_matrix_C(&mc, 2, 3, 1, 0, 0.5 - 1.0 * I);
  double _Complex z = 0;
  monolis_get_scalar_from_sparse_matrix_C(&mc, 2, 3, 1, 0, &z, &found);
  CHECK(found && creal(z) == 1.5 && cimag(z) == 1.0, "complex add/get");
  monolis_finalize(&mc);

  /* BCSR input: item arrives 0-based and is stored 1-based */
  MONOLIS mb;
  monolis_initialize(&mb);
  int bidx[3] = {0, 2, 3}, bitm[3] = {0, 1, 1};
  double bA[3] = {4.0, -1.0, 3.0};
  monolis_set_matrix_BCSR_R(&mb, 2, 2, 1, 3, bA, bidx, bitm);
  CHECK(mb.mat.CSR.item[0] == 1 && mb.mat.CSR.item[1] == 2 && mb.mat.CSR.item[2] == 2 && mb.mat.R.A[1] == -1.0,
        "set_matrix_BCSR");
  monolis_finalize(&mb);

  int xi[3] = {1, 2, 3}, yi[3] = {4, 5, 6}, s = 0;
  MONOLIS_COM com;
  monolis_com_initialize_by_self(&com);
  monolis_inner_productV_I(&com, 3, xi, yi, &s);
  CHECK(s == 32, "inner product I");
  monolis_com_finalize(&com);

  free_elems(e, 3);
  monolis_finalize(&mat);
}

/* monolis_solver_c_test.c shape: n-node chain, diagonal U + 2, off-diagonal
 * blocks U (symmetric), fixed n_dof. */
static void chain_matrix(MONOLIS* mat, int n_node, int n_dof) {
  int** e = chain_elems(n_node);
  monolis_get_nonzero_pattern_by_simple_mesh_R(mat, n_node, 2, n_dof, n_node - 1, e);
  for (int i = 0; i < n_node; ++i)
    for (int k = 0; k < n_dof; ++k) monolis_add_scalar_to_sparse_matrix_R(mat, i, i, k, k, urand() + 2.0);
  for (int k = 0; k < n_node - 1; ++k)
    for (int a = 0; a < n_dof; ++a)
      for (int b = 0; b < n_dof; ++b) {
        const double v = urand();
        monolis_add_scalar_to_sparse_matrix_R(mat, k, k + 1, a, b, v);
        monolis_add_scalar_to_sparse_matrix_R(mat, k + 1, k, b, a, v);
      }
  free_elems(e, n_node - 1);
}

/* y = A x over the internal rows with monolis's block order, on the host */
static void host_matvec(const MONOLIS_MAT* M, const double* x, double* y) {
  for (int i = 0; i < M->N; ++i) {
    const int r0 = M->n_dof_index[i], n1 = M->n_dof_list[i];
    for (int k = 0; k < n1; ++k) y[r0 + k] = 0.0;
    for (int e = M->CSR.index[i]; e < M->CSR.index[i + 1]; ++e) {
      const int j = M->CSR.item[e] - 1, c0 = M->n_dof_index[j], n2 = M->n_dof_list[j];
      for (int k = 0; k < n1; ++k)
        for (int l = 0; l < n2; ++l) y[r0 + k] += M->R.A[M->n_dof_index2[e] + n2 * k + l] * x[c0 + l];
    }
  }
}

static void gpu_checks(void) {
  MONOLIS_COM com;
  monolis_com_initialize_by_self(&com);
  /* every real method and both preconditioners on the 10-node chain, x = 1 */
  for (int n_dof = 1; n_dof <= 3; ++n_dof)
    for (int method = 1; method <= 8; ++method)
      for (int prec = 0; prec <= 1; ++prec) {
        MONOLIS mat;
        monolis_initialize(&mat);
        chain_matrix(&mat, 10, n_dof);
        const int m = 10 * n_dof;
        double one[30], b[30], x[30], yh[30];
        for (int k = 0; k < m; ++k) one[k] = 1.0;
        monolis_matvec_product_R(&mat, &com, one, b);
        host_matvec(&mat.mat, one, yh);
        double err = 0.0;
        for (int k = 0; k < m; ++k) err = fmax(err, fabs(b[k] - yh[k]));
        CHECK(err <= 1e-12, "matvec n_dof %d: %g", n_dof, err);
        monolis_set_method(&mat, method);
        monolis_set_precond(&mat, prec);
        monolis_set_tolerance(&mat, 1e-10);
        monolis_show_iterlog(&mat, false);
        monolis_show_timelog(&mat, false);
        monolis_show_summary(&mat, false);
        for (int k = 0; k < m; ++k) x[k] = 0.0;
        monolis_solve_R(&mat, &com, b, x);
        int it = 0, tag = 0;
        double res = 1.0;
        monolis_get_converge_iter(&mat, &it);
        monolis_get_error_tag(&mat, &tag);
        monolis_get_converge_residual(&mat, &res);
        double xe = 0.0;
        for (int k = 0; k < m; ++k) xe = fmax(xe, fabs(x[k] - 1.0));
        /* converged to 1e-10 in the relative residual; x within 1e-6 (the random chain's
           condition number reaches a few 1e3) */
        CHECK(tag == 0 && it >= 1 && res <= 1e-10 && xe <= 1e-6, "solve method %d prec %d n_dof %d: tag %d it %d "
              "res %g err %g", method, prec, n_dof, tag, it, res, xe);
        double dot = 0.0;
        monolis_inner_product_R(&mat, &com, n_dof, x, x, &dot);
        CHECK(fabs(dot - m) <= 1e-6, "inner product R %g", dot);
        monolis_finalize(&mat);
      }

  /* COCG on the complex chain (wrapper_test/solver/monolis_solver_c_test.c:83-158), x = 1 + i */
  for (int prec = 0; prec <= 1; ++prec) {
    MONOLIS mc;
    monolis_initialize(&mc);
    int** e = chain_elems(10);
    monolis_get_nonzero_pattern_by_simple_mesh_C(&mc, 10, 2, 2, 9, e);
    for (int i = 0; i < 10; ++i)
      for (int k = 0; k < 2; ++k)
        monolis_add_scalar_to_sparse_matrix_C(&mc, i, i, k, k, (urand() + 2.0) + (urand() + 2.0) * I);
    for (int k = 0; k < 9; ++k)
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          const double _Complex v = urand() + urand() * I;
          monolis_add_scalar_to_sparse_matrix_C(&mc, k, k + 1, a, b, v);
          monolis_add_scalar_to_sparse_matrix_C(&mc, k + 1, k, b, a, v);
        }
    free_elems(e, 9);
    double _Complex one[20], b[20], x[20];
    for (int k = 0; k < 20; ++k) one[k] = 1.0 + 1.0 * I;
    monolis_matvec_product_C(&mc, &com, one, b);
    monolis_set_method(&mc, MONOLIS_ITER_COCG);
    monolis_set_precond(&mc, prec);
    monolis_set_tolerance(&mc, 1e-10);
    monolis_show_iterlog(&mc, false);
    monolis_show_timelog(&mc, false);
    monolis_show_summary(&mc, false);
    for (int k = 0; k < 20; ++k) x[k] = 0.0;
    monolis_solve_C(&mc, &com, b, x);
    int tag = 0;
    monolis_get_error_tag(&mc, &tag);
    double xe = 0.0;
    for (int k = 0; k < 20; ++k) xe = fmax(xe, cabs(x[k] - (1.0 + 1.0 * I)));
    CHECK(tag == 0 && xe <= 1e-7, "COCG prec %d: tag %d err %g", prec, tag, xe);
    double _Complex dot = 0;
    monolis_inner_product_C(&mc, &com, 2, one, one, &dot);
    CHECK(cabs(dot - 40.0 * I) <= 1e-12, "inner product C (unconjugated)");
    monolis_finalize(&mc);
  }

  /* variable blocks through the front-end: SpMV against the host loop, CG solve */
  {
    MONOLIS mv;
    monolis_initialize(&mv);
    int** e = chain_elems(12);
    int lst[12];
    for (int i = 0; i < 12; ++i) lst[i] = 1 + i % 3;
    monolis_get_nonzero_pattern_by_simple_mesh_V_R(&mv, 12, 2, lst, 11, e);
    free_elems(e, 11);
    for (int i = 0; i < 12; ++i)
      for (int k = 0; k < lst[i]; ++k) monolis_add_scalar_to_sparse_matrix_R(&mv, i, i, k, k, 4.0 + urand());
    for (int k = 0; k < 11; ++k)
      for (int a = 0; a < lst[k]; ++a)
        for (int b = 0; b < lst[k + 1]; ++b) {
          const double v = 0.5 * urand();
          monolis_add_scalar_to_sparse_matrix_R(&mv, k, k + 1, a, b, v);
          monolis_add_scalar_to_sparse_matrix_R(&mv, k + 1, k, b, a, v);
        }
    const int m = mv.mat.n_dof_index[12];
    double one[24], b[24], yh[24], x[24];
    for (int k = 0; k < m; ++k) one[k] = 1.0;
    monolis_matvec_product_R(&mv, &com, one, b);
    host_matvec(&mv.mat, one, yh);
    double err = 0.0;
    for (int k = 0; k < m; ++k) err = fmax(err, fabs(b[k] - yh[k]));
    CHECK(err <= 1e-12, "V matvec %g", err);
    monolis_set_tolerance(&mv, 1e-10);
    monolis_show_iterlog(&mv, false);
    monolis_show_timelog(&mv, false);
    monolis_show_summary(&mv, false);
    for (int k = 0; k < m; ++k) x[k] = 0.0;
    monolis_solve_R(&mv, &com, b, x);
    double xe = 0.0;
    for (int k = 0; k < m; ++k) xe = fmax(xe, fabs(x[k] - 1.0));
    int tag = 0;
    monolis_get_error_tag(&mv, &tag);
    CHECK(tag == 0 && xe <= 1e-7, "V solve: tag %d err %g", tag, xe);
    monolis_finalize(&mv);
  }
  monolis_com_finalize(&com);
}

static void error_case(const char* which) {
  MONOLIS mat;
  monolis_initialize(&mat);
  int** e = chain_elems(4);
  monolis_get_nonzero_pattern_by_simple_mesh_R(&mat, 4, 2, 2, 3, e);
  free_elems(e, 3);
  if (!strcmp(which, "err_pattern")) {
    monolis_add_scalar_to_sparse_matrix_R(&mat, 0, 3, 0, 0, 1.0);  /* nodes 0 and 3 share no element */
  } else if (!strcmp(which, "err_dof")) {
    monolis_set_scalar_to_sparse_matrix_R(&mat, 1, 1, 2, 0, 1.0);  /* n_dof is 2 */
  } else if (!strcmp(which, "err_dbc")) {
    MONOLIS mb;
    monolis_initialize(&mb);
    int idx[3] = {0, 1, 2}, itm[2] = {1, 0};                         /* no diagonal blocks */
    double A[2] = {1.0, 1.0}, B[2] = {0.0, 0.0};
    monolis_set_matrix_BCSR_R(&mb, 2, 2, 1, 2, A, idx, itm);
    monolis_set_Dirichlet_bc_R(&mb, B, 0, 0, 1.0);
  } else if (!strcmp(which, "err_kind")) {
    monolis_add_scalar_to_sparse_matrix_C(&mat, 0, 0, 0, 0, 1.0 + 1.0 * I);
  } else if (!strcmp(which, "err_matval")) {
    double A[64] = {0.0};
    monolis_set_matrix_BCSR_mat_val_R(&mat, 2, 16, A);  /* the pattern has 10 blocks */
  }
  printf("no error raised\n");
  exit(0);
}

int main(int argc, char** argv) {
  const char* mode = argc > 1 ? argv[1] : "host";
  monolis_global_initialize();
  if (!strncmp(mode, "err_", 4))
    error_case(mode);
  if (!strcmp(mode, "host"))
    host_checks();
  else if (!strcmp(mode, "gpu"))
    gpu_checks();
  else {
    fprintf(stderr, "usage: %s host|gpu\n", argv[0]);
    return 2;
  }
  monolis_global_finalize();
  printf("%s: %s (%d failed)\n", argv[0], fails ? "FAILED" : "ok", fails);
  return fails ? 1 : 0;
}
