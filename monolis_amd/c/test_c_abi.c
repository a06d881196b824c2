/*
 * C caller of the GPU path through monolis's flat C-ABI, in the shape of
 * monolis's own C tests:
 *   - the R33 matvec known answer (wrapper_test/linalg/monolis_matvec_c_test.c:161-252),
 *   - the 2-DOF chain solve with diagonal rand()%1001+5000 and symmetric
 *     off-diagonal blocks rand()%1001, b = A 1, tol 1e-10, x = 1 expected
 *     (wrapper_test/solver/monolis_solver_c_test.c:8-81).
 * Prints "C_ABI_OK" when both pass, exits 1 otherwise.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "monolis_hip.h"

/* chain of n nodes: rows hold {i-1, i, i+1}; 1-based item as monolis stores it */
static int chain_pattern(int n, int* index, int* item) {
  int nz = 0;
  index[0] = 0;
  for (int i = 0; i < n; ++i) {
    if (i > 0) item[nz++] = i;      /* column i-1, 1-based */
    item[nz++] = i + 1;
    if (i < n - 1) item[nz++] = i + 2;
    index[i + 1] = nz;
  }
  return nz;
}

static int find_block(const int* index, const int* item, int i, int j) {
  for (int k = index[i]; k < index[i + 1]; ++k)
    if (item[k] == j + 1) return k;
  return -1;
}

static void set_entry(double* A, const int* index, const int* item, int nd, int i, int j, int si, int sj,
                      double v, int add) {
  const int k = find_block(index, item, i, j);
  if (add) A[nd * nd * k + nd * si + sj] += v;
  else A[nd * nd * k + nd * si + sj] = v;
}

static void prm_default(int* I, double* R) {
  memset(I, 0, 100 * sizeof(int));
  memset(R, 0, 100 * sizeof(double));
  I[1] = 1; I[2] = 1; I[3] = 1000; I[5] = -1; I[10] = 1; I[16] = 1; I[24] = 4;
  R[1] = 1.0e-8;
}

int main(void) {
  int ok = 1;
  int empty[1] = {0};
  /* --- R33 known answer --- */
  {
    const int n = 5, nd = 3;
    int index[6], item[15];
    const int nz = chain_pattern(n, index, item);
    double* A = calloc((size_t)nz * 9, sizeof(double));
    const int e[13][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}, {1, 2}, {2, 1}, {2, 2},
                          {2, 3}, {3, 2}, {3, 3}, {3, 4}, {4, 3}, {4, 4}};
    const double v[13] = {2, 1, 1, 2, 3, 1, 2, 4, 1, 2, 5, 1, 2};
    for (int d = 0; d < nd; ++d)
      for (int k = 0; k < 13; ++k) set_entry(A, index, item, nd, e[k][0], e[k][1], d, d, v[k] * (1 << d), 1);
    double x[15], y[15];
    for (int k = 0; k < 15; ++k) { x[k] = 1.0; y[k] = 0.0; }
    monolis_matvec_product_R_c_main(n, n, nz, nd, n * nd, nz * 9, empty, empty, empty, A, x, y, index, item, 0, 0,
                                    1, 0, 0, empty, empty, empty, 0, 0, empty, empty, empty);
    const double expect[15] = {3, 6, 12, 6, 12, 24, 7, 14, 28, 8, 16, 32, 3, 6, 12};
    for (int k = 0; k < 15; ++k)
      if (y[k] != expect[k]) ok = 0;
    printf("R33 matvec: %s\n", ok ? "ok" : "MISMATCH");
    free(A);
  }
  /* --- 2-DOF chain solve, x = 1 --- */
  {
    const int n = 10, nd = 2;
    int index[11], item[30];
    const int nz = chain_pattern(n, index, item);
    double* A = calloc((size_t)nz * 4, sizeof(double));
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < nd; ++j) set_entry(A, index, item, nd, i, i, j, j, rand() % 1001 + 5000.0, 0);
    for (int i = 0; i < n - 1; ++i)
      for (int j = 0; j < nd; ++j)
        for (int k = 0; k < nd; ++k) {
          const double val = rand() % 1001;
          set_entry(A, index, item, nd, i, i + 1, j, k, val, 0);
          set_entry(A, index, item, nd, i + 1, i, j, k, val, 0);
        }
    double a[20], b[20];
    for (int k = 0; k < 20; ++k) a[k] = 1.0;
    monolis_matvec_product_R_c_main(n, n, nz, nd, n * nd, nz * 4, empty, empty, empty, A, a, b, index, item, 0, 0,
                                    1, 0, 0, empty, empty, empty, 0, 0, empty, empty, empty);
    for (int k = 0; k < 20; ++k) a[k] = 0.0;
    int I[100];
    double R[100];
    prm_default(I, R);
    R[1] = 1.0e-10;
    monolis_solve_R_c_main(n, n, nz, nd, n * nd, nz * 4, empty, empty, empty, A, a, b, index, item, 0, 0, 1, 0, 0,
                           empty, empty, empty, 0, 0, empty, empty, empty, I, R);
    double err = 0.0;
    for (int k = 0; k < 20; ++k) err = fmax(err, fabs(a[k] - 1.0));
    printf("chain solve: iter %d resid %.3e max|x-1| %.3e status %d\n", I[4], R[2], err, I[5]);
    if (err > 1e-8 || I[5] != 0) ok = 0;
    double s = 0.0;
    monolis_inner_product_R_c_main(20, a, a, &s, 0);
    if (fabs(s - 20.0) > 1e-6) ok = 0;
    free(A);
  }
  if (!ok) return 1;
  printf("C_ABI_OK\n");
  return 0;
}
