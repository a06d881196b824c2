/*
 * monolis.h -- monolis's user-facing C API over the GPU path.
 *
 * Source-compatible with the C front-ends of monolis (wrapper/monolis_solver.h
 * and the headers it includes): the same struct layouts, constants and
 * function signatures, so a C application written against monolis compiles
 * unchanged against this header and links against libmonolis_front.so (which
 * pulls in libmonolis_hip.so) with no monolis build.  Matrices live in host
 * memory as monolis keeps them (MONOLIS_MAT); pattern building and assembly
 * run on the host as monolis's do, and every solve / matrix-vector product /
 * inner product goes through the flat entry points of monolis_hip.h, i.e. the
 * GPU path.  Node, DOF and element indices are 0-based, as in monolis's C API.
 *
 * Not reproduced: MPI (one rank; MONOLIS_COM is defined here because
 * monolis_utils, which defines it in monolis, is not part of this build), the
 * eigen / optimisation / ScaLAPACK wrappers, matrix copies, the "with margin"
 * pattern builders and element matrices on variable-block patterns (stop with
 * a message).  Errors that stop monolis (an entry outside the pattern, a DOF
 * outside its block, a Dirichlet condition on a row without its diagonal
 * block) print a message and exit the process here too.
 */
#ifndef MONOLIS_H
#define MONOLIS_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- wrapper/define/monolis_def_solver_prm_c.h:12-154 ------------------- */
static const int MONOLIS_PRM_IARRAY_SIZE = 100;
static const int MONOLIS_PRM_RARRAY_SIZE = 100;
static const int MONOLIS_ITER_CG = 1;
static const int MONOLIS_ITER_GROPCG = 2;
static const int MONOLIS_ITER_PIPECG = 3;
static const int MONOLIS_ITER_PIPECR = 4;
static const int MONOLIS_ITER_BICGSTAB = 5;
static const int MONOLIS_ITER_PIPEBICGSTAB = 6;
static const int MONOLIS_ITER_BICGSTAB_NOPREC = 7;
static const int MONOLIS_ITER_PIPEBICGSTAB_NOPREC = 8;
static const int MONOLIS_ITER_BICGSTAB_N128 = 9;
static const int MONOLIS_ITER_SOR = 10;
static const int MONOLIS_ITER_BICGSAFE = 11;
static const int MONOLIS_ITER_IDRS = 12;
static const int MONOLIS_ITER_COCG = 13;
static const int MONOLIS_PREC_NONE = 0;
static const int MONOLIS_PREC_DIAG = 1;
static const int MONOLIS_PREC_SOR = 2;
static const int MONOLIS_PREC_LU = 3;
static const int MONOLIS_PRM_I_METHOD = 1;
static const int MONOLIS_PRM_I_PRECOND = 2;
static const int MONOLIS_PRM_I_MAX_ITER = 3;
static const int MONOLIS_PRM_I_CUR_ITER = 4;
static const int MONOLIS_PRM_I_IERR = 5;
static const int MONOLIS_PRM_I_IS_INIT_X = 10;
static const int MONOLIS_PRM_I_IS_SYM_MATRIX = 11;
static const int MONOLIS_PRM_I_IS_DEBUG = 12;
static const int MONOLIS_PRM_I_IS_MEASUREMENT = 13;
static const int MONOLIS_PRM_I_IS_CHECK_DIAG = 14;
static const int MONOLIS_PRM_I_IS_PREC_STORED = 15;
static const int MONOLIS_PRM_I_IS_ERROR_ABORT = 16;
static const int MONOLIS_PRM_I_SHOW_ITERLOG = 20;
static const int MONOLIS_PRM_I_SHOW_TIME = 21;
static const int MONOLIS_PRM_I_SHOW_SUMMARY = 22;
static const int MONOLIS_PRM_I_SHOW_TIME_STATISTICS = 23;
static const int MONOLIS_PRM_I_IDRS_DIM = 24;
static const int MONOLIS_PRM_I_SPMV_FORMAT = 25;
static const int MONOLIS_SPMV_ELL = 0;
static const int MONOLIS_SPMV_DIA = 1;
static const int MONOLIS_PRM_R_TOL = 1;
static const int MONOLIS_PRM_R_CUR_RESID = 2;
static const int MONOLIS_R_TIME_SOL = 3;
static const int MONOLIS_R_TIME_PREP = 4;
static const int MONOLIS_R_TIME_SPMV = 5;
static const int MONOLIS_R_TIME_DOTP = 6;
static const int MONOLIS_R_TIME_PREC = 7;
static const int MONOLIS_R_TIME_COMM_DOTP = 8;
static const int MONOLIS_R_TIME_COMM_SPMV = 9;

/* monolis_def_solver_prm_c.h:136-142 */
typedef struct {
  char com_top_dir_name[1024];
  char com_part_dir_name[1024];
  char com_file_name[1024];
  int Iarray[100];
  double Rarray[100];
} MONOLIS_PRM;

/* ---- wrapper/define/monolis_def_mat_c.h:16-117 -------------------------- */
typedef struct {
  double* A;
  double* U;
  double* D;
  double* L;
  double* X;
  double* B;
} MONOLIS_MAT_VAL_R;

typedef struct {
  double _Complex* A;
  double _Complex* U;
  double _Complex* D;
  double _Complex* L;
  double _Complex* X;
  double _Complex* B;
} MONOLIS_MAT_VAL_C;

typedef struct {
  int* indexU;
  int* indexL;
  int* itemU;
  int* itemL;
} MONOLIS_MAT_SEPARATED_CSR;

/* index: 0-based row offsets (NP + 1); item: 1-based block columns, ascending */
typedef struct {
  int* index;
  int* item;
} MONOLIS_MAT_CSR;

/* column view: index (NP + 1), item = 1-based rows, perm = 1-based CSR position */
typedef struct {
  int* index;
  int* item;
  int* perm;
} MONOLIS_MAT_CSC;

typedef struct {
  int N;
  int NP;
  int NDOF;
  int* n_dof_list;
  int* n_dof_index;
  int* n_dof_index2;
  MONOLIS_MAT_VAL_R R;
  MONOLIS_MAT_VAL_C C;
  MONOLIS_MAT_SEPARATED_CSR SCSR;
  MONOLIS_MAT_CSR CSR;
  MONOLIS_MAT_CSC CSC;
} MONOLIS_MAT;

/* ---- wrapper/define/monolis_def_struc_c.h:17-21 ------------------------- */
typedef struct {
  MONOLIS_PRM prm;
  MONOLIS_MAT mat;
  MONOLIS_MAT prec;
} MONOLIS;

/* Communication table.  monolis takes it from monolis_utils (absent from the
 * reference tree); the fields are those its C front-ends read
 * (wrapper/solver/monolis_solver_c.c:14-56).  One rank: see
 * monolis_com_initialize_by_self. */
typedef struct {
  int my_rank;
  int comm;
  int comm_size;
  int n_internal_vertex;
  int recv_n_neib;
  int* recv_neib_pe;
  int* recv_index;
  int* recv_item;
  int send_n_neib;
  int* send_neib_pe;
  int* send_index;
  int* send_item;
} MONOLIS_COM;

/* monolis_utils' monolis_com_initialize_by_self / monolis_com_finalize: a
 * one-rank table (no neighbours). */
void monolis_com_initialize_by_self(MONOLIS_COM* com);
void monolis_com_finalize(MONOLIS_COM* com);

/* ---- initialisation, wrapper/define/monolis_def_struc_c.h:26-52 --------- */
void monolis_global_initialize(void);
void monolis_global_finalize(void);
void monolis_initialize(MONOLIS* mat);
void monolis_finalize(MONOLIS* mat);
/* monolis_def_solver_prm_c.h:147-159 (defaults of def_solver_prm.f90:183-222) */
void monolis_prm_initialize(MONOLIS_PRM* prm);
void monolis_prm_finalize(MONOLIS_PRM* prm);
/* monolis_def_mat_c.h:122-181 */
void monolis_mat_initialize_val_R(MONOLIS_MAT_VAL_R* mat);
void monolis_mat_initialize_val_C(MONOLIS_MAT_VAL_C* mat);
void monolis_mat_initialize_SCSR(MONOLIS_MAT_SEPARATED_CSR* mat);
void monolis_mat_initialize_CSR(MONOLIS_MAT_CSR* mat);
void monolis_mat_initialize_CSC(MONOLIS_MAT_CSC* mat);
void monolis_mat_initialize(MONOLIS_MAT* mat);
void monolis_mat_finalize_val_R(MONOLIS_MAT_VAL_R* mat);
void monolis_mat_finalize_val_C(MONOLIS_MAT_VAL_C* mat);
void monolis_mat_finalize_SCSR(MONOLIS_MAT_SEPARATED_CSR* mat);
void monolis_mat_finalize_CSR(MONOLIS_MAT_CSR* mat);
void monolis_mat_finalize_CSC(MONOLIS_MAT_CSC* mat);
void monolis_mat_finalize(MONOLIS_MAT* mat);
/* monolis_def_mat_c.h:162-181 (monolis_get_vec_size, def_mat.f90:699-721) */
void monolis_get_vec_size(int N, int NP, int NDOF, int NZ, int* n_dof_index, int* n_dof_index2, int* N_size,
                          int* NP_size, int* NZ_size);

/* ---- parameters, wrapper/define/monolis_def_solver_prm_util_c.h:16-281 -- */
void monolis_set_method(MONOLIS* mat, int param);
void monolis_set_precond(MONOLIS* mat, int param);
void monolis_set_spmv_format(MONOLIS* mat, int param);
void monolis_set_maxiter(MONOLIS* mat, int param);
void monolis_get_converge_iter(MONOLIS* mat, int* param);
void monolis_get_error_tag(MONOLIS* mat, int* param);
void monolis_set_init_x(MONOLIS* mat, bool param);
void monolis_set_sym_matrix(MONOLIS* mat, bool param);
void monolis_set_debug(MONOLIS* mat, bool param);
void monolis_set_performance_measurement(MONOLIS* mat, bool param);
void monolis_set_check_diag(MONOLIS* mat, bool param);
void monolis_set_prec_stored(MONOLIS* mat, bool param);
void monolis_set_error_abort(MONOLIS* mat, bool param);
void monolis_set_solver_IDRS_num_basis(MONOLIS* mat, int param);
void monolis_show_iterlog(MONOLIS* mat, bool param);
void monolis_show_timelog(MONOLIS* mat, bool param);
void monolis_show_summary(MONOLIS* mat, bool param);
void monolis_show_timelog_statistics(MONOLIS* mat, bool param);
void monolis_set_tolerance(MONOLIS* mat, double val);
void monolis_get_converge_residual(MONOLIS* mat, double* val);
void monolis_get_time_solver(MONOLIS* mat, double* val);
void monolis_get_time_preparing(MONOLIS* mat, double* val);
void monolis_get_time_spmv(MONOLIS* mat, double* val);
void monolis_get_time_inner_product(MONOLIS* mat, double* val);
void monolis_get_time_precondition(MONOLIS* mat, double* val);
void monolis_get_time_comm_inner_product(MONOLIS* mat, double* val);
void monolis_get_time_comm_spmv(MONOLIS* mat, double* val);

/* ---- nonzero patterns, wrapper/matrix/monolis_spmat_nzpattern_c.h:20-186
 * (src/matrix/spmat_nzpattern.f90:15-327, spmat_nzpattern_util.f90:13-90):
 * nodes sharing an element (or adjacent in the given graph) are coupled, every
 * row holds its diagonal, columns ascending; values, X and B allocated zero.
 * elem[e][k]: node k of element e; conn_index / conn_item: element e's nodes
 * are conn_item[conn_index[e] .. conn_index[e+1]); index / item: 0-based
 * adjacency of the nodal graph without the node itself. */
void monolis_get_nonzero_pattern_by_simple_mesh_R(MONOLIS* mat, int n_node, int n_base, int n_dof, int n_elem,
                                                  int** elem);
void monolis_get_nonzero_pattern_by_simple_mesh_V_R(MONOLIS* mat, int n_node, int n_base, int* n_dof_list,
                                                    int n_elem, int** elem);
void monolis_get_nonzero_pattern_by_connectivity_R(MONOLIS* mat, int n_node, int n_dof, int n_elem,
                                                   int* conn_index, int* conn_item);
void monolis_get_nonzero_pattern_by_nodal_graph_R(MONOLIS* mat, int n_node, int n_dof, int* index, int* item);
void monolis_get_nonzero_pattern_by_nodal_graph_V_R(MONOLIS* mat, int n_node, int* n_dof_list, int* index,
                                                    int* item);
void monolis_get_nonzero_pattern_by_simple_mesh_C(MONOLIS* mat, int n_node, int n_base, int n_dof, int n_elem,
                                                  int** elem);
void monolis_get_nonzero_pattern_by_simple_mesh_V_C(MONOLIS* mat, int n_node, int n_base, int* n_dof_list,
                                                    int n_elem, int** elem);
void monolis_get_nonzero_pattern_by_connectivity_C(MONOLIS* mat, int n_node, int n_dof, int n_elem,
                                                   int* conn_index, int* conn_item);
void monolis_get_nonzero_pattern_by_nodal_graph_C(MONOLIS* mat, int n_node, int n_dof, int* index, int* item);
void monolis_get_nonzero_pattern_by_nodal_graph_V_C(MONOLIS* mat, int n_node, int* n_dof_list, int* index,
                                                    int* item);

/* ---- assembly, wrapper/matrix/monolis_spmat_handler_c.h:20-288
 * (src/matrix/spmat_handler_util.f90:1-677).  i, j: 0-based block row / column;
 * submat_i, submat_j: 0-based DOF inside the block.  val of add_matrix:
 * val[a*n_dof + k1][c*n_dof + k2] is component (k1, k2) of block
 * (connectivity[a], connectivity[c]) (fixed block sizes). */
void monolis_set_scalar_to_sparse_matrix_R(MONOLIS* mat, int i, int j, int submat_i, int submat_j, double val);
void monolis_add_scalar_to_sparse_matrix_R(MONOLIS* mat, int i, int j, int submat_i, int submat_j, double val);
void monolis_get_scalar_from_sparse_matrix_R(MONOLIS* mat, int i, int j, int submat_i, int submat_j, double* val,
                                             bool* is_find);
void monolis_add_matrix_to_sparse_matrix_R(MONOLIS* mat, int n_base, int* connectivity, double** val);
void monolis_add_matrix_to_sparse_matrix_offdiag_R(MONOLIS* mat, int n_base1, int n_base2, int* connectivity1,
                                                   int* connectivity2, double** val);
void monolis_set_matrix_BCSR_R(MONOLIS* mat, int n, int np, int n_dof, int nz, double* A, int* index, int* item);
void monolis_set_matrix_BCSR_mat_val_R(MONOLIS* mat, int n_dof, int nz, double* A);
void monolis_set_Dirichlet_bc_R(MONOLIS* mat, double* b, int node_id, int n_dof_bc, double val);
void monolis_set_scalar_to_sparse_matrix_C(MONOLIS* mat, int i, int j, int submat_i, int submat_j,
                                           double _Complex val);
void monolis_add_scalar_to_sparse_matrix_C(MONOLIS* mat, int i, int j, int submat_i, int submat_j,
                                           double _Complex val);
void monolis_get_scalar_from_sparse_matrix_C(MONOLIS* mat, int i, int j, int submat_i, int submat_j,
                                             double _Complex* val, bool* is_find);
void monolis_add_matrix_to_sparse_matrix_C(MONOLIS* mat, int n_base, int* connectivity, double _Complex** val);
void monolis_add_matrix_to_sparse_matrix_offdiag_C(MONOLIS* mat, int n_base1, int n_base2, int* connectivity1,
                                                   int* connectivity2, double _Complex** val);
void monolis_set_matrix_BCSR_C(MONOLIS* mat, int n, int np, int n_dof, int nz, double _Complex* A, int* index,
                               int* item);
void monolis_set_matrix_BCSR_mat_val_C(MONOLIS* mat, int n_dof, int nz, double _Complex* A);
void monolis_set_Dirichlet_bc_C(MONOLIS* mat, double _Complex* b, int node_id, int n_dof_bc, double _Complex val);

/* ---- solver / linear algebra on the GPU path -------------------------------
 * wrapper/solver/monolis_solver_c.h:22-45 (monolis_solver_c.c:8-114),
 * wrapper/linalg/monolis_matvec_c.h:20-37 (monolis_matvec_c.c:8-108),
 * wrapper/linalg/monolis_inner_product_c.h:20-137, monolis_vec_util_c.h:18-44.
 * solve: x in / out (zeroed first unless monolis_set_init_x(false)), b in;
 * status and statistics in mat->prm (monolis_get_converge_iter ...). */
void monolis_solve_R(MONOLIS* mat, MONOLIS_COM* com, double* b, double* x);
void monolis_solve_C(MONOLIS* mat, MONOLIS_COM* com, double _Complex* b, double _Complex* x);
void monolis_matvec_product_R(MONOLIS* mat, MONOLIS_COM* com, double* x, double* y);
void monolis_matvec_product_C(MONOLIS* mat, MONOLIS_COM* com, double _Complex* x, double _Complex* y);
void monolis_inner_product_I(MONOLIS* mat, MONOLIS_COM* com, int n_dof, int* x, int* y, int* sum);
void monolis_inner_productV_I(MONOLIS_COM* com, int m, int* x, int* y, int* sum);
void monolis_inner_product_R(MONOLIS* mat, MONOLIS_COM* com, int n_dof, double* x, double* y, double* sum);
void monolis_inner_productV_R(MONOLIS_COM* com, int m, double* x, double* y, double* sum);
void monolis_inner_product_C(MONOLIS* mat, MONOLIS_COM* com, int n_dof, double _Complex* x, double _Complex* y,
                             double _Complex* sum);
void monolis_inner_productV_C(MONOLIS_COM* com, int m, double _Complex* x, double _Complex* y,
                              double _Complex* sum);
void monolis_vec_copy_I(int m, int* x, int* y);
void monolis_vec_copy_R(int m, double* x, double* y);
void monolis_vec_copy_C(int m, double _Complex* x, double _Complex* y);

#ifdef __cplusplus
}
#endif

#endif /* MONOLIS_H */
