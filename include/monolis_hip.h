/*
 * monolis_hip.h -- C-ABI of libmonolis_hip.so, the MI355X (gfx950) Krylov path
 * of monolis: block-CSR SpMV, CG / BiCGSTAB and the diagonal (block-Jacobi)
 * preconditioner, on one GPU.
 *
 * Part 1 are monolis's own flat entry points, with identical names and
 * signatures, so a monolis C or Fortran caller links against this library
 * instead of libmonolis_solver.  Part 2 adds a device-resident interface (the
 * matrix is uploaded once and stays in HBM) and the precision, batching and
 * adjoint extensions.
 *
 * Conventions of part 1 (as in monolis):
 *  - index[NP+1] holds 0-based row offsets, item[NZ] 1-based block columns
 *    (diagonal included, ascending per row), A[NZ*ndof*ndof] row-major blocks;
 *  - Iarray[100] / Rarray[100]: slot k here is Fortran slot k (1-based names,
 *    e.g. Iarray[1] = method, Iarray[4] = cur_iter, Rarray[1] = tol);
 *  - pointers may be host or device memory (device memory is used in place).
 * One rank only: comm_size must be 1 (the halo tables are then empty).
 */
#ifndef MONOLIS_HIP_H
#define MONOLIS_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================== */
/* Part 1: monolis flat C-ABI                                               */
/* ======================================================================== */

/* Replaces monolis_solve_R_c_main, wrapper/solver/monolis_solver_c.h:51-80,
 * defined as the Fortran bind(c) routine wrapper/solver/solver_wrapper.f90:11-92.
 * Runs monolis_solve_main_R (src/solver/solver.f90:77-185) with the method in
 * Iarray[1] (CG = 1, GropCG = 2, PipeCG = 3, PipeCR = 4, BiCGSTAB = 5,
 * PipeBiCGSTAB = 6, BiCGSTAB_noprec = 7, PipeBiCGSTAB_noprec = 8) and
 * preconditioner in Iarray[2] (NONE = 0, DIAG = 1) on the GPU.  Fills Iarray[4] (cur_iter), Iarray[5] (status: 0 ok,
 * 1 not converged, < 0 error), Rarray[2] (cur_resid) and Rarray[3..4]
 * (solution / preparation time, s).  Non-convergence stops the process when
 * Iarray[16] (error_abort) is set and Iarray[13] (measurement) is not,
 * as src/linalg/mat_converge.f90:134-140 does -- except for PipeCR, whose
 * loop ends silently at max_iter (src/iterative/PipeCR.f90:81-123).  PipeCR
 * reports cur_iter (the iteration whose test passed) and cur_resid
 * (||M^-1 r|| / ||M^-1 r0||) although the reference leaves both slots
 * unwritten for it. */
void monolis_solve_R_c_main(int n, int np, int nz, int n_dof, int NPNDOF, int NZNDOF2, int* n_dof_list,
                            int* n_dof_index, int* n_dof_index2, double* A, double* x, double* b, int* index,
                            int* item, int my_rank, int comm, int comm_size, int recv_n_neib, int recv_nitem,
                            int* recv_neib_pe, int* recv_index, int* recv_item, int send_n_neib, int send_nitem,
                            int* send_neib_pe, int* send_index, int* send_item, int* Iarray, double* Rarray);

/* Replaces monolis_matvec_product_R_c_main, wrapper/linalg/monolis_matvec_c.h:39-66,
 * defined at wrapper/linalg/matvec_wrap.f90:13-108: y[0 .. n*n_dof) = A x. */
void monolis_matvec_product_R_c_main(int n, int np, int nz, int n_dof, int NPNDOF, int NZNDOF2, int* n_dof_list,
                                     int* n_dof_index, int* n_dof_index2, double* A, double* x, double* y,
                                     int* index, int* item, int my_rank, int comm, int comm_size,
                                     int recv_n_neib, int recv_nitem, int* recv_neib_pe, int* recv_index,
                                     int* recv_item, int send_n_neib, int send_nitem, int* send_neib_pe,
                                     int* send_index, int* send_item);

/* Replaces monolis_inner_product_R_c_main, wrapper/linalg/monolis_inner_product_c.h:139,
 * defined at wrapper/linalg/inner_product_wrap.f90:35-55: *sum = sum_i x_i y_i. */
void monolis_inner_product_R_c_main(int m, double* x, double* y, double* sum, int comm);

/* Complex path (double _Complex = two doubles, real part first; the layout of
 * Fortran complex(c_double) and C99 double _Complex).  Fixed block sizes
 * 1..8; variable blocks (n_dof = -1) are rejected on this path. */

/* Replaces monolis_solve_C_c_main, wrapper/solver/monolis_solver_c.h:124-153,
 * defined as the Fortran bind(c) routine wrapper/solver/solver_wrapper.f90:94-175.
 * Runs monolis_solve_main_C (src/solver/solver.f90:189-213): the method in
 * Iarray[1] must be COCG = 13 (monolis_solver_select_C accepts nothing else),
 * preconditioner NONE = 0 or DIAG = 1.  Fills Iarray[4] (cur_iter), Iarray[5]
 * (status as the real entry point; -6 = zero pivot in a diagonal block, where
 * monolis stops), Rarray[2] = real part of the complex sqrt(R2/B2) (what
 * monolis stores; convergence tests its modulus) and Rarray[3..4]. */
void monolis_solve_C_c_main(int n, int np, int nz, int n_dof, int NPNDOF, int NZNDOF2, int* n_dof_list,
                            int* n_dof_index, int* n_dof_index2, double _Complex* A, double _Complex* x,
                            double _Complex* b, int* index, int* item, int my_rank, int comm, int comm_size,
                            int recv_n_neib, int recv_nitem, int* recv_neib_pe, int* recv_index, int* recv_item,
                            int send_n_neib, int send_nitem, int* send_neib_pe, int* send_index, int* send_item,
                            int* Iarray, double* Rarray);

/* Replaces monolis_matvec_product_C_c_main, wrapper/linalg/monolis_matvec_c.h:68-95,
 * defined at wrapper/linalg/matvec_wrap.f90:112-207:
 * y[0 .. n*n_dof) = A x in complex arithmetic. */
void monolis_matvec_product_C_c_main(int n, int np, int nz, int n_dof, int NPNDOF, int NZNDOF2, int* n_dof_list,
                                     int* n_dof_index, int* n_dof_index2, double _Complex* A, double _Complex* x,
                                     double _Complex* y, int* index, int* item, int my_rank, int comm,
                                     int comm_size, int recv_n_neib, int recv_nitem, int* recv_neib_pe,
                                     int* recv_index, int* recv_item, int send_n_neib, int send_nitem,
                                     int* send_neib_pe, int* send_index, int* send_item);

/* Replaces monolis_inner_product_C_c_main, wrapper/linalg/monolis_inner_product_c.h:155-160,
 * defined at wrapper/linalg/inner_product_wrap.f90:59-77:
 * *sum = sum_i x_i y_i, unconjugated (src/linalg/inner_product.f90:457-499). */
void monolis_inner_product_C_c_main(int m, double _Complex* x, double _Complex* y, double _Complex* sum, int comm);

/* ======================================================================== */
/* Part 2: device-resident interface and extensions (no monolis counterpart  */
/* except where cited).  Every function returns 0 on success, < 0 on error;  */
/* monolis_hip_last_error() describes the last failure of the calling thread.*/
/* ======================================================================== */

typedef struct monolis_hip_matrix monolis_hip_matrix;

const char* monolis_hip_last_error(void);
/* Selects the GPU, creates the library stream (default priority). */
int monolis_hip_init(void);
/* Library stream as a hipStream_t (void* here to keep HIP out of this header). */
void* monolis_hip_stream(void);
int monolis_hip_synchronize(void);
int monolis_hip_malloc(void** ptr, size_t bytes);
int monolis_hip_free(void* ptr);
/* Copy in any direction (host/device detected from the pointers), synchronous. */
int monolis_hip_memcpy(void* dst, const void* src, size_t bytes);
int monolis_hip_memset(void* dst, int value, size_t bytes);

/* Uploads a monolis BCSR matrix (host or device arrays, same layout as part 1)
 * and builds the GPU layout.  The counterpart of monolis's device staging
 * (src/solver/solver.f90:102-156) done once instead of per solve.  n_dof 1..8
 * takes the fixed-block kernels; 9..16 is stored as the variable-block
 * expansion (the same sums in the same order as matvec_nn / diag_nn). */
int monolis_hip_matrix_create(monolis_hip_matrix** out, int n, int np, int nz, int n_dof, const int* index,
                              const int* item, const double* A);
/* Variable block sizes (monolis n_dof = -1): n_dof_list[NP], n_dof_index[NP+1]
 * (0-based DOF offsets) and n_dof_index2[NZ+1] (0-based value offsets) as
 * monolis builds them (src/matrix/spmat_nzpattern_util.f90:62-90); block (i, j)
 * is n_i x n_j row-major at A[n_dof_index2[e]].  Stored as the scalar expansion
 * (one row per DOF) with the diag_V block-Jacobi factors (blocks up to 16).
 * The flat entry points take this path when n_dof = -1. */
int monolis_hip_matrix_create_V(monolis_hip_matrix** out, int n, int np, int nz, const int* n_dof_list,
                                const int* n_dof_index, const int* n_dof_index2, const int* index, const int* item,
                                const double* A);
int monolis_hip_matrix_destroy(monolis_hip_matrix* M);
/* New values, same pattern (A(theta) updates). */
int monolis_hip_matrix_set_values(monolis_hip_matrix* M, const double* A);
/* info[0..7] = n, np, nz, n_dof, slices, slots, bytes of FP64 layout, bytes of FP32 layout */
int monolis_hip_matrix_info(const monolis_hip_matrix* M, long long* info);

/* y = A x on device vectors (FP64 / FP32 values and vectors). */
int monolis_hip_matvec_R(monolis_hip_matrix* M, const double* x, double* y);
int monolis_hip_matvec_S(monolis_hip_matrix* M, const float* x, float* y);
/* Deterministic device dot product; *sum is host memory. */
int monolis_hip_inner_product_R(int m, const double* x, const double* y, double* sum);
/* Allocation-free variant of monolis_hip_inner_product_R.  No counterpart in
 * monolis: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc / hipFree and initialises nothing, so
 * monolis_hip_init must have run first (it also sets the call up).  x and y
 * must be device arrays (staging host data would allocate); sum may be host
 * or device memory.  The reduction's partial sums live in a file-scope device
 * array of 4096 doubles per GPU (up to 4095 partial sets; a launch needs at
 * most 2 per CU, so sizes beyond that are rejected with an error only on
 * GPUs of more than 2047 CUs).  Every call on a GPU shares that array:
 * calls that overlap on one GPU race.  Same kernels and partial count as
 * monolis_hip_inner_product_R, so the two agree bit for bit. */
int monolis_hip_inner_product_R_noalloc(int m, const double* x, const double* y, double* sum);

/* monolis_solve_main_R on device-resident b and x (FP64), Iarray/Rarray host,
 * same slots and semantics as monolis_solve_R_c_main except that failures are
 * returned, never turned into a process stop.  If hist != NULL and
 * hist_len > 0, hist[k] = relative residual after iteration k (k >= 1). */
int monolis_hip_solve_R(monolis_hip_matrix* M, const double* b, double* x, int* Iarray, double* Rarray,
                        double* hist, int hist_len);

/* FP32-inner / FP64-refined solve.  The inner CG (Iarray[1] = 1) or BiCGSTAB
 * (5) runs its SpMV, vectors and preconditioner in FP32 to relative residual
 * inner_tol; FP64 is used for r = b - A x, the update x += d and the norms.
 * Stops when ||b - A x|| / ||b - A x0|| < Rarray[1].  stats[0..3] = outer
 * iterations, total inner iterations, final relative residual, converged. */
int monolis_hip_solve_refined_R(monolis_hip_matrix* M, const double* b, double* x, int* Iarray, double* Rarray,
                                double inner_tol, int max_outer, double* stats);

/* Kernel timing with HIP events on the library stream: `warmup` untimed and
 * `reps` timed SpMVs; *ms = average milliseconds per SpMV.  fp32 selects the
 * FP32 layout (x, y are float* then). */
int monolis_hip_time_matvec(monolis_hip_matrix* M, int fp32, const void* x, void* y, int warmup, int reps,
                            double* ms);
/* The same with a pair of events around each of the `reps` SpMVs:
 * ms_each[k] = milliseconds of SpMV k (for medians; the event pairs add a
 * little per-launch overhead that the batch average above does not have). */
int monolis_hip_time_matvec_each(monolis_hip_matrix* M, int fp32, const void* x, void* y, int warmup, int reps,
                                 double* ms_each);

/* Autotuner (no monolis counterpart): times each of the SELL SpMV kernel
 * variants (slots per step, plain or software-pipelined, x launch shape: one
 * wavefront per slice or a persistent grid of 4..24 waves per CU) on this
 * matrix with HIP events, `reps` launches each, and keeps the fastest one per
 * precision for every later SpMV and solve on the handle.  best[0] / best[1]
 * receive the FP64 / FP32 choice; ms (if not NULL) receives 2 x
 * monolis_hip_num_variants() average times, FP64 first.  fp32 = 0 tunes FP64
 * only. */
int monolis_hip_num_variants(void);
int monolis_hip_matrix_autotune(monolis_hip_matrix* M, int fp32, int reps, int* best, double* ms);
/* Solver-context variant tuning (no counterpart in monolis): times `reps`
 * solves of exactly `iters` iterations of `method` / `precond` (tolerance 0)
 * per SpMV variant in one precision (fp32 = 0 or 1, the FP32 copy that the
 * refined solve's inner iterations use) and keeps the fastest on the handle;
 * ms (if not NULL) gets monolis_hip_num_variants() milliseconds per iteration. */
int monolis_hip_matrix_autotune_solver(monolis_hip_matrix* M, int fp32, int method, int precond, int iters,
                                       int reps, int* best, double* ms);
/* Forces a variant for one precision (fp32 = 0 or 1); an untuned handle uses
 * variant 4 (one wavefront per slice, 3 slots per step). */
int monolis_hip_matrix_set_variant(monolis_hip_matrix* M, int fp32, int variant);

/* Batched solves (no monolis counterpart; equivalent to one monolis_solve_R
 * per system).  A batch is a block-diagonal scalar CSR matrix in monolis's
 * format whose system k owns rows sys_row[k] .. sys_row[k+1]-1 (sys_row has
 * nsys+1 entries, host memory) and couples only to its own rows.  Every system
 * is solved by CG (Iarray[1] = 1) with NONE / DIAG preconditioning
 * (Iarray[2]), its own scalars, convergence test (Rarray[1]) and iteration
 * count, in one persistent launch captured in a hipGraph and replayed. */
typedef struct monolis_hip_batch monolis_hip_batch;
int monolis_hip_batch_create(monolis_hip_batch** out, int nsys, const int* sys_row, int n, int nz,
                             const int* index, const int* item, const double* A);
int monolis_hip_batch_destroy(monolis_hip_batch* B);
/* b, x: device vectors of length n.  iters / resid / status: host arrays of
 * nsys entries (status 0 converged, 1 not converged, 2 zero right-hand side,
 * 3 abandoned: a team of workgroups gave up waiting for a member, in which case
 * the call returns an error). */
int monolis_hip_batch_solve(monolis_hip_batch* B, const double* b, double* x, int* Iarray, double* Rarray,
                            int* iters, double* resid, int* status);
/* Autotuner of the persistent launch: workgroup size x slots per step x team
 * split x vector streams (configurations 13 and 14 load and store the vectors
 * of the elementwise passes non-temporally).  Runs the solve `reps` times per
 * configuration, keeps the fastest; ms (if not NULL) gets
 * monolis_hip_batch_num_configs() average times. */
int monolis_hip_batch_num_configs(void);
int monolis_hip_batch_config(int cfg, int* threads, int* unroll);
/* Selects configuration cfg for the following solves (the autotuner's choice
 * otherwise; 0 before any tuning). */
int monolis_hip_batch_set_config(monolis_hip_batch* B, int cfg);
/* Team split of configuration cfg (0: one workgroup per system; s > 0: systems
 * whose estimated work exceeds s/4 of one workgroup's balanced share are solved
 * by a team of up to 32 workgroups) and the number of slots (one per workgroup
 * task) and teams its schedule for B has. */
int monolis_hip_batch_schedule(monolis_hip_batch* B, int cfg, int* split, int* nslot, int* nteam);
/* Longest time in seconds (default 5, at most 600) that one team barrier waits
 * for a member before the launch gives up on the system (status 3). */
int monolis_hip_batch_set_wait_limit(monolis_hip_batch* B, double seconds);
int monolis_hip_batch_autotune(monolis_hip_batch* B, const double* b, double* x, int* Iarray, double* Rarray,
                               int reps, int* best, double* ms);
/* Average time of `reps` graph replays of the batched solve (ms). */
int monolis_hip_batch_time(monolis_hip_batch* B, const double* b, double* x, int* Iarray, double* Rarray, int reps,
                           double* ms);

/* Multi-right-hand-side solves (no monolis counterpart; equivalent to one
 * monolis_solve_R per column).  A scalar matrix (n_dof = 1) and nrhs (1..64)
 * right-hand sides stored row-major: b and x are device arrays of NP x nrhs
 * doubles with entry (i, j) at i*nrhs + j.  Every column runs monolis's CG
 * (Iarray[1] = 1, NONE / DIAG preconditioning) with its own scalars,
 * convergence test (Rarray[1]) and iteration count; the SpMM that applies A to
 * all columns reads the matrix once.  iters / resid / status: host arrays of
 * nrhs entries (status 0 converged, 1 not converged, 2 zero right-hand side,
 * x = 0).  Iarray[4] receives the largest iteration count, Rarray[2] the
 * largest relative residual, Iarray[5] the status. */
int monolis_hip_solve_multi_R(monolis_hip_matrix* M, int nrhs, const double* b, double* x, int* Iarray,
                              double* Rarray, int* iters, double* resid, int* status);
/* FP32-inner / FP64-refined variant: FP32 multi-RHS CG corrections to relative
 * residual inner_tol; FP64 residual B - A X, update and norms; a column stops
 * when its ||b - A x|| / ||b - A x0|| < Rarray[1].  iters = inner iterations
 * per column summed over corrections; stats[0] = outer iterations. */
int monolis_hip_solve_multi_refined_R(monolis_hip_matrix* M, int nrhs, const double* b, double* x, int* Iarray,
                                      double* Rarray, double inner_tol, int max_outer, int* iters, double* resid,
                                      int* status, double* stats);

/* Y = A X for nrhs columns (SpMM, scalar matrices): x, y device arrays of
 * NP x nrhs (row-major, right-hand sides innermost), FP64 / FP32.  Equivalent
 * to nrhs monolis_matvec_product_R calls; the matrix is read once. */
int monolis_hip_matvec_multi_R(monolis_hip_matrix* M, int nrhs, const double* x, double* y);
int monolis_hip_matvec_multi_S(monolis_hip_matrix* M, int nrhs, const float* x, float* y);

/* Autotuner of the multi-RHS SpMM: row batches whose gathers are issued
 * together x workgroups per CU x right-hand sides per pass over the matrix
 * (chunk; 0 = all in one pass) x threads per workgroup x lds (1: the tile's
 * column ids and values staged through LDS, x read with 32-bit offsets).  Times `reps` SpMMs with nrhs columns per
 * configuration (fp32 = 1: the FP32 copy used by the refined solve's inner CG)
 * and keeps the fastest on the handle; ms (if not NULL) gets
 * monolis_hip_multi_num_configs() averages. */
int monolis_hip_multi_num_configs(void);
int monolis_hip_multi_config(int cfg, int* steps_per_batch, int* wg_per_cu, int* chunk, int* threads, int* lds);
int monolis_hip_multi_autotune(monolis_hip_matrix* M, int nrhs, int fp32, int reps, int* best, double* ms);
/* Forces a multi-RHS SpMM configuration for one precision (fp32 = 0 or 1). */
int monolis_hip_multi_set_config(monolis_hip_matrix* M, int fp32, int cfg);

/* Device-side scalar access to a stored entry, block (i, j), component
 * (sub_i, sub_j), all 0-based.  Counterparts of
 * monolis_{set,add}_scalar_to_sparse_matrix_R and
 * monolis_get_scalar_from_sparse_matrix_R (wrapper/matrix/monolis_spmat_handler_c.h:19-61,
 * src/matrix/spmat_handler_util.f90:32-74, 207-252, 306-348) on the handle's
 * device copy.  A block outside the pattern is an error (monolis stops the
 * process there); get returns *is_find = 0 and *val = 0 for it. */
int monolis_hip_matrix_set_scalar_R(monolis_hip_matrix* M, int i, int j, int sub_i, int sub_j, double val);
int monolis_hip_matrix_add_scalar_R(monolis_hip_matrix* M, int i, int j, int sub_i, int sub_j, double val);
int monolis_hip_matrix_get_scalar_R(monolis_hip_matrix* M, int i, int j, int sub_i, int sub_j, double* val,
                                    int* is_find);

/* Adjoint extension (no monolis counterpart).  For x = A^{-1} b and a loss
 * L(x) with g = dL/dx:
 *   A^T lambda = g,  dL/db = lambda,  dL/dA(i,j)[r][c] = -lambda[i*b+r] * x[j*b+c]
 * on the stored blocks only, in monolis's A layout (block e at e*b*b, row-major).
 * All vectors are device arrays; square local matrices (n == np) only.
 *
 * monolis_hip_solve_transpose_R: A^T x = b with the method / preconditioner /
 * tolerance of Iarray / Rarray; the transpose is built on the device on first
 * use (a radix sort of (column, row) keys) and refreshed after value changes. */
int monolis_hip_solve_transpose_R(monolis_hip_matrix* M, const double* b, double* x, int* Iarray, double* Rarray);
/* dA = -lambda x^T restricted to the pattern (nz * n_dof^2 values). */
int monolis_hip_outer_grad_R(monolis_hip_matrix* M, const double* lambda, const double* x, double* dA);
/* Adjoint solve plus gradient: lambda = A^{-T} g (A itself when symmetric != 0),
 * then dA as monolis_hip_outer_grad_R when dA != NULL.  Iarray / Rarray as in
 * monolis_hip_solve_R, for the adjoint solve. */
int monolis_hip_adjoint_R(monolis_hip_matrix* M, const double* x, const double* g, double* lambda, double* dA,
                          int symmetric, int* Iarray, double* Rarray);

/* Device assembly: counterparts of monolis's matrix assembly routines that
 * build and fill the handle's matrix on the GPU.  Node ids are 0-based, as in
 * monolis's C API; real fixed-block square matrices (n == np).
 *
 * monolis_hip_matrix_create_by_simple_mesh replaces the pattern build of
 * monolis_get_nonzero_pattern_by_simple_mesh_R (wrapper/matrix/monolis_spmat_nzpattern_c.h:21-27,
 * src/matrix/spmat_nzpattern.f90:15-49): nodes that share an element are
 * coupled, every row holds its diagonal, columns ascending; the matrix starts
 * at zero.  elem holds n_elem * n_base node ids (element-major). */
int monolis_hip_matrix_create_by_simple_mesh(monolis_hip_matrix** out, int n_node, int n_base, int n_dof, int n_elem,
                                             const int* elem);
/* index (np + 1, 0-based offsets) and item (nz, 1-based columns) of the handle (variable
 * blocks: the block pattern); host or device. */
int monolis_hip_matrix_get_pattern(const monolis_hip_matrix* M, int* index, int* item);
/* The values in monolis's layout (nz * n_dof^2, row-major blocks); host or device. */
int monolis_hip_matrix_get_values_R(const monolis_hip_matrix* M, double* A);
/* monolis_add_matrix_to_sparse_matrix_R (wrapper/matrix/monolis_spmat_handler_c.h:73-77,
 * src/matrix/spmat_handler_util.f90:398-497) for n_elem elements in one call:
 * Ke holds n_elem row-major (n_base*n_dof)^2 element matrices; entry
 * (a*n_dof + k1, c*n_dof + k2) of element e is added to component (k1, k2) of
 * block (elem[e*n_base + a], elem[e*n_base + c]).  The sums are those of one
 * call per element in element order, bit for bit.  Elements are processed
 * in chunks of 2^24 node pairs (environment MONOLIS_HIP_ASSEMBLY_CHUNK_PAIRS
 * overrides); a block outside the pattern is an error (monolis stops there),
 * and the chunks before the failing one stay added. */
int monolis_hip_matrix_add_elements_R(monolis_hip_matrix* M, int n_elem, int n_base, const int* elem,
                                      const double* Ke);
/* monolis_set_Dirichlet_bc_R (wrapper/matrix/monolis_spmat_handler_c.h:142-147,
 * src/matrix/spmat_handler_util.f90:604-677) for n_bc conditions at once:
 * x(node_id[i]*n_dof + n_dof_bc[i]) = val[i].  b (N*n_dof, host or device)
 * receives b - val * A(:, dof) for the other rows and val for the condition
 * rows; condition rows and columns are zeroed and their diagonals set to 1.
 * Equal to applying the conditions one by one (the right-hand-side updates
 * of one row are summed in column order, which is list order for a list
 * sorted by node and DOF).  A DOF listed twice, or a node without a diagonal
 * block (monolis stops there), is an error. */
int monolis_hip_set_Dirichlet_bc_R(monolis_hip_matrix* M, double* b, int n_bc, const int* node_id,
                                   const int* n_dof_bc, const double* val);
/* Variable block sizes (monolis n_dof = -1): monolis_get_nonzero_pattern_by_simple_mesh_V
 * (src/matrix/spmat_nzpattern.f90:53-87) with n_dof_list[n_node] (host or device) block
 * sizes of 1..16; the matrix starts at zero.  On such a handle add_elements_R takes
 * element e's matrix as (sum_a n_dof_list[elem[e*n_base + a]])^2 row-major values,
 * the elements' matrices back to back; set_Dirichlet_bc_R takes DOFs below the node's
 * block size; get_values_R returns monolis's variable-block layout (n_dof_index2[nz]
 * values).  The complex handle has the same (create_by_simple_mesh_V_C with the _C
 * calls). */
int monolis_hip_matrix_create_by_simple_mesh_V(monolis_hip_matrix** out, int n_node, int n_base,
                                               const int* n_dof_list, int n_elem, const int* elem);

/* Device-resident complex matrix (the complex counterpart of monolis_hip_matrix;
 * same BCSR arrays, A holds nz * n_dof^2 double _Complex values).  A separate
 * handle type keeps real and complex matrices apart at compile time. */
typedef struct monolis_hip_matrix_C monolis_hip_matrix_C;
int monolis_hip_matrix_create_C(monolis_hip_matrix_C** out, int n, int np, int nz, int n_dof, const int* index,
                                const int* item, const double _Complex* A);
/* Variable block sizes (n_dof = -1) for complex values: the arrays as
 * monolis_hip_matrix_create_V takes them, A of n_dof_index2[nz] complex
 * entries.  Replaces the complex variable-block path of
 * monolis_matvec_product_main_C (src/linalg/matvec.f90:218-220, matvec_V_C
 * :293-351) and of the DIAG preconditioner (src/prec/diag.f90:53-57,
 * diag_V.f90:88-158 setup, :208-250 apply).  Blocks of up to 16 DOFs.  A zero
 * pivot in a diagonal block is reported (status -6) by the first solve with
 * the DIAG preconditioner, where the reference stops. */
int monolis_hip_matrix_create_V_C(monolis_hip_matrix_C** out, int n, int np, int nz, const int* n_dof_list,
                                  const int* n_dof_index, const int* n_dof_index2, const int* index, const int* item,
                                  const double _Complex* A);
int monolis_hip_matrix_destroy_C(monolis_hip_matrix_C* M);
int monolis_hip_matrix_set_values_C(monolis_hip_matrix_C* M, const double _Complex* A);
/* info[0..7] as monolis_hip_matrix_info, value bytes at 16 per entry (info[7] = 0). */
/* Complex device assembly, the counterparts of the real calls above on a
 * complex handle: the pattern of a simple mesh (monolis_get_nonzero_pattern_by_simple_mesh_C,
 * wrapper/matrix/monolis_spmat_nzpattern_c.h, src/matrix/spmat_nzpattern.f90), element
 * matrices (monolis_add_matrix_to_sparse_matrix_C, src/matrix/spmat_handler_util.f90:501-600,
 * bit for bit one call per element in element order), Dirichlet conditions
 * (monolis_set_Dirichlet_bc_C, src/matrix/spmat_handler_util.f90:681-749; b -= val * A with
 * the complex product (ar*br - ai*bi, ar*bi + ai*br)) and read-back. */
int monolis_hip_matrix_create_by_simple_mesh_C(monolis_hip_matrix_C** out, int n_node, int n_base, int n_dof,
                                               int n_elem, const int* elem);
int monolis_hip_matrix_create_by_simple_mesh_V_C(monolis_hip_matrix_C** out, int n_node, int n_base,
                                                 const int* n_dof_list, int n_elem, const int* elem);
int monolis_hip_matrix_get_pattern_C(const monolis_hip_matrix_C* M, int* index, int* item);
int monolis_hip_matrix_get_values_C(const monolis_hip_matrix_C* M, double _Complex* A);
int monolis_hip_matrix_add_elements_C(monolis_hip_matrix_C* M, int n_elem, int n_base, const int* elem,
                                      const double _Complex* Ke);
int monolis_hip_set_Dirichlet_bc_C(monolis_hip_matrix_C* M, double _Complex* b, int n_bc, const int* node_id,
                                   const int* n_dof_bc, const double _Complex* val);
int monolis_hip_matrix_info_C(const monolis_hip_matrix_C* M, long long* info);
/* y = A x (monolis_matvec_product_main_C, src/linalg/matvec.f90:192-227). */
int monolis_hip_matvec_C(monolis_hip_matrix_C* M, const double _Complex* x, double _Complex* y);
/* *sum = sum_i x_i y_i, unconjugated (monolis_inner_product_main_C). */
int monolis_hip_inner_product_C(int m, const double _Complex* x, const double _Complex* y, double _Complex* sum);
/* COCG (src/iterative/COCG.f90:18-105) on device vectors; Iarray / Rarray as in
 * monolis_solve_C_c_main; hist[k] = |sqrt(R2/B2)| after iteration k. */
int monolis_hip_solve_C(monolis_hip_matrix_C* M, const double _Complex* b, double _Complex* x, int* Iarray,
                        double* Rarray, double* hist, int hist_len);
/* Average time of one complex SpMV over `reps` launches after `warmup`. */
int monolis_hip_time_matvec_C(monolis_hip_matrix_C* M, const double _Complex* x, double _Complex* y, int warmup,
                              int reps, double* ms);

/* Bandwidth probe (no monolis counterpart): a plain streaming copy kernel
 * dst[i] = src[i] over n doubles, `reps` timed launches on the library stream;
 * *ms = average per launch.  Used to calibrate HBM byte counters and to state
 * the copy bandwidth the SpMV is compared with. */
int monolis_hip_stream_copy(double* dst, const double* src, long long n, int reps, double* ms);

/* Read-bandwidth probe (no monolis counterpart): every one of the n doubles of
 * src is read once per launch (summed per wavefront so none is dropped; one
 * double per wavefront written), `reps` timed launches on the library stream
 * for each of six load shapes and grid sizes; *ms = the fastest average per
 * launch.  States the read ceiling that the read-dominated SpMV is compared with. */
int monolis_hip_stream_read(const double* src, long long n, int reps, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* MONOLIS_HIP_H */
