// Host-side structures shared by the translation units of libmonolis_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <vector>

#include "mh_common.hpp"

namespace mh {

template <typename TV>
struct SellView;

// Largest fixed block size with specialised kernels.
constexpr int kMaxB = 8;

// ---------------------------------------------------------------------------
// Per-process device context: one HIP stream at default priority, pinned host
// slots for polling solver flags, and a reusable pool of reduction scratch.
// ---------------------------------------------------------------------------
// What the host reads of a running solve: the last tested residual, its
// iteration and the done flag (laid out as in KState / ZState).
struct PollView {
  double resid;
  int iter;
  int done;
};

struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  int* h_flags = nullptr;            // pinned host, kFlagSlots ints
  PollView* h_poll = nullptr;        // pinned host, kFlagSlots solver progress views
  hipEvent_t ev[16] = {};
  hipStream_t up[8] = {};            // streams of the threaded pageable uploads
  bool ok = false;
};
constexpr int kFlagSlots = 16;
constexpr int kUploadThreads = 8;

// Returns 0 and fills *c on success.  Initialises HIP on first use.
int context(Context** c);
// The context if it is already initialised (never initialises: that allocates), else null.
Context* context_if_ready();

template <typename T>
struct DevArray {
  T* p = nullptr;
  size_t n = 0;
  int alloc(size_t count);
  void release();
  ~DevArray() { release(); }
  DevArray() = default;
  DevArray(const DevArray&) = delete;
  DevArray& operator=(const DevArray&) = delete;
};

// Copies `count` elements from a host or device pointer into device memory
// (the direction is detected from the pointer's attributes).
int is_device_pointer(const void* p);
int upload(void* dst, const void* src, size_t bytes, hipStream_t s);

// Device-side solver scalars, written by the last workgroup of a reduction and
// read by later kernels.  One instance per solve.
struct KState {
  double rho;       // (r, z) of the current iteration (CG) / (r, rt) (BiCGSTAB)
  double rho_prev;
  double alpha, beta, omega;
  double pq;        // (p, q) / (rt, v)
  double rr;        // (r, r)
  double B2;        // (r0, r0), denominator of the relative residual
  // resid, iter and done are adjacent: the host polls them as one 16-byte PollView
  double resid;     // sqrt(rr / B2)
  int iter;         // last iteration whose convergence test ran
  int done;
  double tol;
  double ts, tt;    // BiCGSTAB (t, s), (t, t)
  double gamma, delta;  // PipeCG (r, u), (v, u)
  double aux[4];
  int max_iter;
  int converged;
  int breakdown;    // bit mask of exact-zero dot products seen (BiCGSTAB)
  int bd_stage;     // != 0: the loop stopped for the sorted-sum fallback at this stage
  int bd_iter;      // ... in this iteration
  int pad;
  // rho of iteration k in rhok[k & 1]: a kernel that derives rho_k stores it
  // while its other workgroups still read rho_{k-1} from the other slot
  double rhok[2];
};

static_assert(offsetof(KState, iter) == offsetof(KState, resid) + 8 &&
                  offsetof(KState, done) == offsetof(KState, resid) + 12,
              "KState::resid, iter, done must form a PollView");

// ---------------------------------------------------------------------------
// Variable block sizes (monolis n_dof = -1): the matrix is stored as the scalar
// expansion (one SELL row per DOF, columns in monolis's block-then-entry order)
// and keeps the block structure for the diag_V preconditioner.
// ---------------------------------------------------------------------------
constexpr int kMaxVBlock = 16;  // largest variable block the diag_V kernels take
struct VBlock {
  int nb = 0, npb = 0, nzb = 0, maxb = 0;  // block rows (internal / total), stored blocks, largest block
  long long nzs = 0;                       // scalar entries = n_dof_index2[nz]
  DevArray<int> list, dof_index, index, item, index2;  // monolis arrays, device copies
  DevArray<long long> doff;                // offset of block row i's factor (prefix of n_i^2)
  DevArray<int> srowptr;                   // scalar row offsets of the expansion
  DevArray<double> D64;                    // diag_V factors (no-pivot LU, reciprocal pivots)
  DevArray<float> D32;
  bool d32_ready = false;
  DevArray<zd> Dz;                         // complex matrices: diag_V_C factors
  DevArray<double> A64;                    // the values in monolis's layout (nzs), kept for device assembly
  DevArray<zd> Az;                         // complex matrices: the same
  int zbad = -1;                           // first block row whose complex factor hit a zero pivot, or -1
};

// ---------------------------------------------------------------------------
// Device matrix: SELL-64 pattern + FP64 (and optionally FP32) values, plus the
// CSR arrays kept for the preconditioner setup, transposes and gradients.
// ---------------------------------------------------------------------------
struct Matrix {
  int n = 0, np = 0, nz = 0, b = 0;
  int nslices = 0;
  size_t nslots = 0;
  DevArray<int> slice_ptr, rowptr, col;     // SELL pattern (rowptr = monolis index)
  // 16-bit column deltas (col - row), two slots per 32-bit word (SellView::col16);
  // empty when some |col - row| exceeds 32767
  DevArray<unsigned short> col16;
  DevArray<int> csr_item;                   // 1-based, as given
  DevArray<double> csr_val;                 // AoS blocks, as given (kept for setup / gradients)
  DevArray<double> val64;
  DevArray<float> val32;
  // block-Jacobi factors, structure of arrays (entry e of row i at D[e*n + i])
  DevArray<double> D64;
  DevArray<float> D32;
  bool diag_ready = false, diag32_ready = false, val32_ready = false;
  // SELL kernel variant per precision (0 = default); set by the autotuner
  int variant64 = 0, variant32 = 0;
  // multi-right-hand-side SpMM launch configuration per precision (-1 = default)
  int mrhs_cfg64 = -1, mrhs_cfg32 = -1;
  // transpose (for adjoint solves of nonsymmetric systems), built lazily;
  // tperm[p] = block of A that becomes block p of A^T, tiperm its inverse
  Matrix* transpose = nullptr;
  DevArray<int> tperm, tiperm;
  bool transpose_stale = false;   // values of A changed since the transpose was filled
  VBlock* vb = nullptr;           // non-null for variable block sizes
  // complex matrices (monolis's *_C path): values in valz instead of val64,
  // complex block-Jacobi factors in Dz (same structure-of-arrays layout as D64)
  bool is_complex = false;
  DevArray<zd> valz;
  DevArray<zd> Dz;
  bool diagz_ready = false;
  ~Matrix();
};

int matrix_create(Matrix** out, int n, int np, int nz, int b, const int* index, const int* item,
                  const double* A, hipStream_t s);
int matrix_set_values(Matrix* M, const double* A, hipStream_t s);
// Complex matrix: A holds NZ*b*b complex values, interleaved (re, im) doubles;
// matrix_set_values refills it from the same layout.
int matrix_create_complex(Matrix** out, int n, int np, int nz, int b, const int* index, const int* item,
                          const double* A, hipStream_t s);
// Variable block sizes (monolis n_dof = -1): arrays as monolis stores them
// (host or device pointers); see VBlock.
int matrix_create_vblock(Matrix** out, int n, int np, int nz, const int* n_dof_list, const int* n_dof_index,
                         const int* n_dof_index2, const int* index, const int* item, const double* A,
                         hipStream_t s);
int matrix_set_values_vblock(Matrix* M, const double* A, hipStream_t s);
// The same for complex values (monolis_matvec_V_C / diag_V_*_C); the diag_V_C
// factors are built with the values, a zero pivot recorded in VBlock::zbad.
int matrix_create_vblock_complex(Matrix** out, int n, int np, int nz, const int* n_dof_list, const int* n_dof_index,
                                 const int* n_dof_index2, const int* index, const int* item, const zd* A,
                                 hipStream_t s);
int matrix_set_values_vblock_complex(Matrix* M, const zd* A, hipStream_t s);
// FP32 copy of the diag_V factors (the FP64 ones are built with the values).
int vdiag_make_fp32(Matrix& M, hipStream_t s);
// y = M_V^-1 x (diag_V apply) over all block rows.
template <typename T>
int vdiag_apply(const Matrix& M, const T* x, T* y, hipStream_t s);
int matrix_make_fp32(Matrix* M, hipStream_t s);
int matrix_transpose(Matrix* M, hipStream_t s);

template <typename TV>
SellView<TV> view(const Matrix& M);

// ---------------------------------------------------------------------------
// Kernel launchers (all asynchronous on `s`).
// ---------------------------------------------------------------------------
template <typename T>
int spmv(const Matrix& M, const T* x, T* y, hipStream_t s);
// SpMV with an explicit kernel variant (autotuner / sweeps).
template <typename T>
int spmv_variant(const Matrix& M, const T* x, T* y, int variant, hipStream_t s);
// r = b - A x; if rr_out != nullptr, rr_out = (r, r) (device scalar, deterministic).
template <typename T>
int residual(const Matrix& M, const T* x, const T* b, T* r, double* partials, double* rr_out, hipStream_t s);
template <typename T>
int diag_setup(Matrix& M, hipStream_t s);
template <typename T>
int diag_apply(const Matrix& M, const T* x, T* y, hipStream_t s);
// Device dot, deterministic; result written to *out (device pointer).
template <typename T>
int dot(long long m, const T* x, const T* y, double* out, double* partials, hipStream_t s);
template <typename T>
int axpby(long long m, double alpha, const T* x, double beta, const T* y, T* z, hipStream_t s);
// monolis_global_sorted_inner_product_main_R on the device: products sorted by
// magnitude, summed left to right; result written to *out (device).
template <typename T>
int sorted_dot(long long n, const T* x, const T* y, double* out, hipStream_t s);

// Number of reduction partials a launch over `rows` rows (or `slices`) may need.
size_t partials_needed(long long n);

// ---------------------------------------------------------------------------
// Solver driver.
// ---------------------------------------------------------------------------
struct SolveParams {
  int method = 1;        // 1 CG, 3 PipeCG, 5 BiCGSTAB
  int precond = 1;       // 0 NONE, 1 DIAG
  int max_iter = 1000;
  int init_x = 1;
  double tol = 1e-8;
  int chunk = 8;         // iterations enqueued between convergence polls
  int keep_hist = 1;     // record per-iteration residuals
  int rhs_nonzero = 0;   // caller guarantees b - A x0 != 0: skip the zero right-hand-side read-back
};

struct SolveResult {
  int iter = 0;
  double resid = 0.0;
  int converged = 0;
  int zero_rhs = 0;
  int breakdown = 0;     // bit mask of exact-zero dots seen
  int fallbacks = 0;     // sorted-sum recomputations performed
  double t_solve = 0.0;  // seconds, device-resident part
  std::vector<double> hist;
};

// Work buffers of one solve, reusable across solves on the same matrix.
struct Workspace {
  DevArray<unsigned char> raw;
  size_t bytes = 0;
  int ensure(size_t need);
};

template <typename T>
int krylov_solve(Matrix& M, const SolveParams& P, const T* b, T* x, Workspace& W, SolveResult* out,
                 hipStream_t s);

// FP32-inner / FP64-refined solve.  `x` and `b` are FP64 device vectors.
struct RefineParams {
  double tol = 1e-8;        // outer relative residual ||b - A x|| / ||b - A x0||
  double inner_tol = 1e-4;  // inner FP32 relative residual per correction
  int max_outer = 50;
  int inner_max_iter = 1000;
  int method = 1;
  int precond = 1;
  int init_x = 1;
};
struct RefineResult {
  int outer = 0;
  int inner_total = 0;
  double resid = 0.0;
  int converged = 0;
  std::vector<double> outer_resid;
  std::vector<int> inner_iters;
};
int refine_solve(Matrix& M, const RefineParams& P, const double* b, double* x, Workspace& W, Workspace& Wo,
                 RefineResult* out, hipStream_t s);

}  // namespace mh
