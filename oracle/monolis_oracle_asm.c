/*
 * monolis_oracle_asm.c -- CPU restatement of monolis's matrix assembly
 * (fixed block size b, one rank, node ids 1-based inside as in the
 * reference, 0-based at this interface as in monolis's C API).
 *
 * TEST INFRASTRUCTURE ONLY, like monolis_oracle.c.
 *
 * Pinning (tests/test_oracle_assembly.py): the reference's own unit tests
 * (src_test/matrix/spmat_nzpattern_test.f90:24-73 pattern,
 * src_test/matrix/spmat_handler_test.f90:18-133 element add and Dirichlet
 * values), dense brute force (pair sets, dense element sums, the solution of
 * the reduced system after Dirichlet conditions).  The complex conditions:
 * the reference's unit test (spmat_handler_test.f90:236-248) and the closed
 * forms of the complex product (real values reduce to the real conditions;
 * values i*v move the real update into the imaginary parts).
 */
#include <stdlib.h>
#include <string.h>

static int cmp_int(const void* a, const void* b) {
  const int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

/* monolis_get_nonzero_pattern_by_simple_mesh_R (src/matrix/spmat_nzpattern.f90:15-49):
 * the nodal graph of the mesh (nodes sharing an element, gedatsu's
 * connectivity -> nodal graph conversion), then every row gets its diagonal
 * and is sorted (spmat_nzpattern_util.f90:33-49).  Duplicate couplings count
 * once.  index: n_node + 1 entries; item: capacity `cap`, 1-based columns.
 * Returns nz, or -1 when cap is too small / a node id is out of range. */
int orc_pattern_simple_mesh(int n_node, int n_base, int n_elem, const int* elem, int* index, int* item,
                            long long cap) {
  long long* cnt = (long long*)calloc((size_t)n_node + 1, sizeof(long long));
  for (long long e = 0; e < n_elem; ++e)
    for (int a = 0; a < n_base; ++a) {
      const int na = elem[e * n_base + a];
      if (na < 0 || na >= n_node) {
        free(cnt);
        return -1;
      }
      cnt[na + 1] += n_base;  /* upper bound: every node of the element, self included */
    }
  for (int i = 0; i < n_node; ++i) cnt[i + 1] += cnt[i] + 1;  /* + diagonal */
  int* buf = (int*)malloc(sizeof(int) * (size_t)(cnt[n_node] > 0 ? cnt[n_node] : 1));
  long long* fill = (long long*)malloc(sizeof(long long) * (size_t)(n_node > 0 ? n_node : 1));
  for (int i = 0; i < n_node; ++i) {
    buf[cnt[i]] = i;
    fill[i] = cnt[i] + 1;
  }
  for (long long e = 0; e < n_elem; ++e)
    for (int a = 0; a < n_base; ++a) {
      const int na = elem[e * n_base + a];
      for (int c = 0; c < n_base; ++c) buf[fill[na]++] = elem[e * n_base + c];
    }
  long long nz = 0;
  index[0] = 0;
  for (int i = 0; i < n_node; ++i) {
    int* row = buf + cnt[i];
    const long long len = fill[i] - cnt[i];
    qsort(row, (size_t)len, sizeof(int), cmp_int);
    for (long long k = 0; k < len; ++k) {
      if (k > 0 && row[k] == row[k - 1]) continue;
      if (nz >= cap) {
        free(cnt);
        free(buf);
        free(fill);
        return -1;
      }
      item[nz++] = row[k] + 1;
    }
    index[i + 1] = (int)nz;
  }
  free(cnt);
  free(buf);
  free(fill);
  return (int)nz;
}

/* monolis_add_matrix_to_sparse_matrix_main_R (src/matrix/spmat_handler_util.f90:398-497)
 * for one element and a fixed block size b: the connectivity is sorted with
 * its permutation, the element matrix reordered into `temp`, and every
 * (row node, column node) block found by a forward search along the row
 * (the search start moves to the last found block), then
 * A(block)(k1, k2) = A(block)(k1, k2) + temp.  val is the element matrix as
 * monolis's C API passes it: row-major (n_base*b)^2.  Returns 0, or -1 when a
 * block is not in the pattern (monolis stops: monolis_stop_by_matrix_assemble). */
int orc_add_matrix(int b, const int* index, const int* item, double* A, int n_base, const int* conn,
                   const double* val) {
  const int nzm = n_base * b;
  int* et = (int*)malloc(sizeof(int) * (size_t)n_base);
  int* perm = (int*)malloc(sizeof(int) * (size_t)n_base);
  double* temp = (double*)malloc(sizeof(double) * (size_t)nzm * nzm);
  for (int i = 0; i < n_base; ++i) {
    et[i] = conn[i];
    perm[i] = i;
  }
  /* sort nodes ascending, carrying the original position (monolis_qsort_I_2d) */
  for (int i = 1; i < n_base; ++i)
    for (int j = i; j > 0 && et[j - 1] > et[j]; --j) {
      int t = et[j]; et[j] = et[j - 1]; et[j - 1] = t;
      t = perm[j]; perm[j] = perm[j - 1]; perm[j - 1] = t;
    }
  /* temp(sorted row node i1, sorted column node i2) = val(original positions) */
  for (int i1 = 0; i1 < n_base; ++i1)
    for (int i2 = 0; i2 < n_base; ++i2)
      for (int k1 = 0; k1 < b; ++k1)
        for (int k2 = 0; k2 < b; ++k2)
          temp[(i1 * b + k1) * nzm + i2 * b + k2] = val[(perm[i1] * b + k1) * nzm + perm[i2] * b + k2];
  int st = 0;
  for (int i1 = 0; i1 < n_base && !st; ++i1) {
    const int in = et[i1];
    int jS = index[in];
    const int jE = index[in + 1];
    for (int i2 = 0; i2 < n_base; ++i2) {
      int found = 0;
      for (int j1 = jS; j1 < jE; ++j1) {
        if (item[j1] - 1 == et[i2]) {
          for (int k1 = 0; k1 < b; ++k1)
            for (int k2 = 0; k2 < b; ++k2) {
              const long long im = (long long)j1 * b * b + b * k1 + k2;
              A[im] = A[im] + temp[(i1 * b + k1) * nzm + i2 * b + k2];
            }
          jS = j1;
          found = 1;
          break;
        }
      }
      if (!found) {
        st = -1;
        break;
      }
    }
  }
  free(et);
  free(perm);
  free(temp);
  return st;
}

/* n_elem calls of orc_add_matrix, in element order. */
int orc_add_elements(int b, const int* index, const int* item, double* A, int n_elem, int n_base,
                     const int* elem, const double* Ke) {
  const long long per = (long long)n_base * b * n_base * b;
  for (long long e = 0; e < n_elem; ++e)
    if (orc_add_matrix(b, index, item, A, n_base, elem + e * n_base, Ke + e * per)) return -1;
  return 0;
}

/* monolis_set_Dirichlet_bc_main_R (src/matrix/spmat_handler_util.f90:604-677)
 * for one condition, fixed block size b, node_id and ndof_bc 0-based.  The
 * column of the condition DOF is walked through the CSC view
 * (monolis_get_CSC_format, rows ascending): B(row) = B(row) - val*A, A = 0;
 * then the row is zeroed, its diagonal set to 1 and B(dof) = val.
 * Returns 0, or -1 when the row has no diagonal block (monolis stops). */
int orc_dirichlet(int n, int b, const int* index, const int* item, double* A, double* B, int node_id, int ndof_bc,
                  double val) {
  /* column pass: rows jn holding block (jn, node_id), ascending jn */
  for (int jn = 0; jn < n; ++jn) {
    for (int j = index[jn]; j < index[jn + 1]; ++j) {
      if (item[j] - 1 != node_id) continue;
      for (int k = 0; k < b; ++k) {
        const long long m = (long long)j * b * b + (long long)b * k + ndof_bc;
        const long long nn = (long long)jn * b + k;
        B[nn] = B[nn] - val * A[m];
        A[m] = 0.0;
      }
    }
  }
  int is_add = 0;
  for (int j = index[node_id]; j < index[node_id + 1]; ++j) {
    const int jn = item[j] - 1;
    for (int k = 0; k < b; ++k) A[(long long)j * b * b + (long long)b * ndof_bc + k] = 0.0;
    if (jn == node_id) {
      A[(long long)j * b * b + (long long)(b + 1) * ndof_bc] = 1.0;
      is_add = 1;
    }
  }
  if (!is_add) return -1;
  B[(long long)node_id * b + ndof_bc] = val;
  return 0;
}

/* The conditions of a list applied one after the other. */
int orc_dirichlet_list(int n, int b, const int* index, const int* item, double* A, double* B, int n_bc,
                       const int* node, const int* dof, const double* val) {
  for (int i = 0; i < n_bc; ++i)
    if (orc_dirichlet(n, b, index, item, A, B, node[i], dof[i], val[i])) return -1;
  return 0;
}

/* monolis_set_Dirichlet_bc_main_C (src/matrix/spmat_handler_util.f90:681-749):
 * orc_dirichlet on complex values (A and B interleaved re, im).  The right-hand
 * side update B(n) = B(n) - val*A(m) (:723) takes Fortran's complex product
 * (vr*ar - vi*ai, vr*ai + vi*ar), written out so that no contraction applies;
 * the condition row is zeroed, its diagonal set to (1, 0) and B(dof) = val. */
int orc_dirichlet_C(int n, int b, const int* index, const int* item, double* A, double* B, int node_id, int ndof_bc,
                    double vr, double vi) {
  for (int jn = 0; jn < n; ++jn) {
    for (int j = index[jn]; j < index[jn + 1]; ++j) {
      if (item[j] - 1 != node_id) continue;
      for (int k = 0; k < b; ++k) {
        const long long m = (long long)j * b * b + (long long)b * k + ndof_bc;
        const long long nn = (long long)jn * b + k;
        const double ar = A[2 * m], ai = A[2 * m + 1];
        const double pr = vr * ar, pi_ = vi * ai, qr = vr * ai, qi = vi * ar;
        const double tr = pr - pi_, ti = qr + qi;
        B[2 * nn] = B[2 * nn] - tr;
        B[2 * nn + 1] = B[2 * nn + 1] - ti;
        A[2 * m] = 0.0;
        A[2 * m + 1] = 0.0;
      }
    }
  }
  int is_add = 0;
  for (int j = index[node_id]; j < index[node_id + 1]; ++j) {
    const int jn = item[j] - 1;
    for (int k = 0; k < b; ++k) {
      const long long m = (long long)j * b * b + (long long)b * ndof_bc + k;
      A[2 * m] = 0.0;
      A[2 * m + 1] = 0.0;
    }
    if (jn == node_id) {
      const long long m = (long long)j * b * b + (long long)(b + 1) * ndof_bc;
      A[2 * m] = 1.0;
      A[2 * m + 1] = 0.0;
      is_add = 1;
    }
  }
  if (!is_add) return -1;
  B[2 * ((long long)node_id * b + ndof_bc)] = vr;
  B[2 * ((long long)node_id * b + ndof_bc) + 1] = vi;
  return 0;
}

/* The complex conditions of a list applied one after the other (val: 2 * n_bc). */
int orc_dirichlet_list_C(int n, int b, const int* index, const int* item, double* A, double* B, int n_bc,
                         const int* node, const int* dof, const double* val) {
  for (int i = 0; i < n_bc; ++i)
    if (orc_dirichlet_C(n, b, index, item, A, B, node[i], dof[i], val[2 * i], val[2 * i + 1])) return -1;
  return 0;
}
