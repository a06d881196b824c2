// Device-side assembly: nonzero pattern of a simple mesh, batched element
// matrix assembly, Dirichlet conditions and value read-back.
#pragma once

#include "mh_internal.hpp"

namespace mh {

// Pattern of monolis_get_nonzero_pattern_by_simple_mesh_R: nodes sharing an
// element are coupled, every row holds its diagonal, columns ascending.
// elem: n_elem * n_base 0-based node ids (host or device).  On success the
// device arrays hold index (n_node + 1, 0-based offsets) and item (nz, 1-based).
int pattern_by_simple_mesh(int n_node, int n_base, int n_elem, const int* elem, DevArray<int>& index,
                           DevArray<int>& item, int* nz, hipStream_t s);

// A += element matrices, in element order (monolis_add_matrix_to_sparse_matrix_R
// called once per element).  Ke: n_elem blocks of (n_base*b)^2 row-major values,
// entry (a*b + k1, c*b + k2) adds to component (k1, k2) of block (elem[a], elem[c]).
int matrix_add_elements(Matrix* M, int n_elem, int n_base, const int* elem, const double* Ke, hipStream_t s);
// The same on a complex matrix (monolis_add_matrix_to_sparse_matrix_C per element).
int matrix_add_elements_C(Matrix* M, int n_elem, int n_base, const int* elem, const zd* Ke, hipStream_t s);

// Dirichlet conditions x(node*b + dof) = val, monolis_set_Dirichlet_bc_R applied
// for each listed condition in turn: the known values are moved to the right-hand
// side b (device vector, N*b), their rows and columns zeroed, diagonals set to 1.
int matrix_dirichlet(Matrix* M, double* b, int n_bc, const int* node, const int* dof, const double* val,
                     hipStream_t s);
// The same on a complex matrix (monolis_set_Dirichlet_bc_C per condition).
int matrix_dirichlet_C(Matrix* M, zd* b, int n_bc, const int* node, const int* dof, const zd* val, hipStream_t s);

// A in monolis's layout (nz * b^2, row-major blocks, CSR order); host or device.
int matrix_get_values(const Matrix& M, double* A, hipStream_t s);
int matrix_get_values_C(const Matrix& M, zd* A, hipStream_t s);
// index (np + 1) and item (nz, 1-based) as stored (variable blocks: the block
// pattern); host or device.
int matrix_get_pattern(const Matrix& M, int* index, int* item, hipStream_t s);

}  // namespace mh
