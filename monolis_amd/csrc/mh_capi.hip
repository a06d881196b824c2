// extern "C" entry points of libmonolis_hip.so (declared in include/monolis_hip.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <vector>

#include "../../include/monolis_hip.h"
#include "mh_internal.hpp"
#include "mh_kernels.hpp"
#include "mh_launch.hpp"
#include "mh_batch.hpp"
#include "mh_multirhs.hpp"
#include "mh_adjoint.hpp"
#include "mh_complex.hpp"
#include "mh_assembly.hpp"

using mh::Context;
using mh::Matrix;

struct monolis_hip_batch {
  mh::Batch* B = nullptr;
  ~monolis_hip_batch() { delete B; }
};

struct monolis_hip_matrix_C {
  Matrix* M = nullptr;  // is_complex
  mh::Workspace W;
  ~monolis_hip_matrix_C() { delete M; }
};

struct monolis_hip_matrix {
  Matrix* M = nullptr;
  mh::Workspace W;
  mh::Workspace W_outer;  // FP64 residual / FP32 correction of multi-RHS refinement
  ~monolis_hip_matrix() { delete M; }
};

namespace {

// Iarray / Rarray slots (src/define/def_solver_prm.f90:90-163)
enum {
  I_METHOD = 1, I_PRECOND = 2, I_MAX_ITER = 3, I_CUR_ITER = 4, I_IERR = 5, I_IS_INIT_X = 10,
  I_IS_MEASUREMENT = 13, I_IS_ERROR_ABORT = 16, I_SHOW_ITERLOG = 20, I_SHOW_TIME = 21, I_SHOW_SUMMARY = 22,
  I_SHOW_TIME_STATISTICS = 23
};
enum { R_TOL = 1, R_CUR_RESID = 2, R_TIME_SOL = 3, R_TIME_PREP = 4, R_TIME_SPMV = 5, R_TIME_DOTP = 6,
       R_TIME_PREC = 7, R_TIME_COMM_DOTP = 8, R_TIME_COMM_SPMV = 9 };

// Status codes written to Iarray[5].
enum { ST_OK = 0, ST_NOT_CONVERGED = 1, ST_HIP = -1, ST_ARGS = -2, ST_PRECOND = -3, ST_METHOD = -4,
       ST_MULTI_RANK = -5, ST_ZERO_PIVOT = -6 };

const char* str_iter(int m) {
  static const char* s[] = {"CG", "GropCG", "PipeCG", "PipeCR", "BiCGSTAB", "PipeBiCGSTAB", "BiCGSTAB_noprec",
                            "PipeBiCGSTAB_noprec", "BiCGSTAB_N128", "SOR", "BiCGSAFE", "IDR(s)", "COCG"};
  return (m >= 1 && m <= 13) ? s[m - 1] : "unknown";
}
const char* str_prec(int p) {
  static const char* s[] = {"None", "Diag", "SOR", "LU", "MUMPS", "MUMPSL"};
  return (p >= 0 && p <= 5) ? s[p] : "unknown";
}

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

mh::SolveParams params_from(const int* Iarray, const double* Rarray) {
  mh::SolveParams P;
  P.method = Iarray[I_METHOD];
  P.precond = Iarray[I_PRECOND];
  P.max_iter = Iarray[I_MAX_ITER];
  P.init_x = Iarray[I_IS_INIT_X] != 0;
  P.tol = Rarray[R_TOL];
  return P;
}

int check_method(int method, int precond) {
  if (method < 1 || method > 8) {
    mh::set_error("method %d (%s) is not implemented on the GPU path (CG = 1, GropCG = 2, PipeCG = 3, "
                  "PipeCR = 4, BiCGSTAB = 5, PipeBiCGSTAB = 6, BiCGSTAB_noprec = 7, PipeBiCGSTAB_noprec = 8)", method,
                  str_iter(method));
    return ST_METHOD;
  }
  if (precond != 0 && precond != 1) {
    mh::set_error("preconditioner %d (%s) is not implemented on the GPU path (NONE = 0, DIAG = 1)", precond,
                  str_prec(precond));
    return ST_PRECOND;
  }
  return 0;
}

void print_banner(const int* Iarray, int my_rank) {
  // src/solver/solver.f90:230-234
  if (Iarray[I_SHOW_SUMMARY] && my_rank == 0)
    printf("** MONOLIS solver: %s, prec: %s\n", str_iter(Iarray[I_METHOD]), str_prec(Iarray[I_PRECOND]));
}

void print_after(const int* Iarray, const double* Rarray, const std::vector<double>& hist, int my_rank) {
  if (my_rank != 0) return;
  // per-iteration log of src/linalg/mat_converge.f90:125-127, printed after the solve
  if (Iarray[I_SHOW_ITERLOG])
    for (size_t k = 1; k < hist.size(); ++k) printf("%7d%16.6E\n", (int)k, hist[k]);
  // src/define/def_solver_prm.f90:272-303
  if (Iarray[I_SHOW_SUMMARY]) {
    printf("** MONOLIS converge iter:%10d\n", Iarray[I_CUR_ITER]);
    printf("** MONOLIS rel. residual:%10.3E\n", Rarray[R_CUR_RESID]);
  }
  if (Iarray[I_SHOW_SUMMARY] || Iarray[I_SHOW_TIME]) printf("** MONOLIS solution time:%10.3E\n", Rarray[R_TIME_SOL]);
  if (Iarray[I_SHOW_TIME]) {
    printf(" - solution/prepost time:%10.3E\n", Rarray[R_TIME_PREP]);
    printf(" - solution/SpMV    time:%10.3E\n", Rarray[R_TIME_SPMV]);
    printf(" - solution/inner p time:%10.3E\n", Rarray[R_TIME_DOTP]);
    printf(" - solution/precond time:%10.3E\n", Rarray[R_TIME_PREC]);
    printf(" - (comm time/inner p)  :%10.3E\n", Rarray[R_TIME_COMM_DOTP]);
    printf(" - (comm time/spmv)     :%10.3E\n", Rarray[R_TIME_COMM_SPMV]);
  }
  fflush(stdout);
}

// Fixed blocks larger than the fixed-block kernels take (kMaxB) are stored as
// the variable-block expansion with every block n_dof: matvec_nn and diag_nn
// (src/linalg/matvec.f90:123-188 nn branch, src/prec/diag/diag_nn.f90) sum and
// factor in the same order as matvec_V and diag_V, and the zero-pivot rules
// agree (real: pivot 0 -> 1; complex: stop).
struct UniformBlocks {
  std::vector<int> list, dofi, dofi2;
  int make(int np, int nz, int b) {
    if (np < 0 || nz < 0) {
      mh::set_error("invalid matrix sizes np=%d nz=%d", np, nz);
      return ST_ARGS;
    }
    if ((long long)nz * b * b >= (1LL << 31) || (long long)np * b >= (1LL << 31)) {
      mh::set_error("block size %d with %d blocks exceeds 32-bit offsets", b, nz);
      return ST_ARGS;
    }
    list.assign((size_t)np + 1, b);
    dofi.resize((size_t)np + 1);
    dofi2.resize((size_t)nz + 1);
    for (int i = 0; i <= np; ++i) dofi[i] = i * b;
    for (int e = 0; e <= nz; ++e) dofi2[e] = e * b * b;
    return 0;
  }
};

[[noreturn]] void error_stop(const char* msg) {
  fprintf(stderr, "** monolis error: %s\n", msg);
  fflush(stderr);
  exit(EXIT_FAILURE);
}

// Solver time split of monolis's Rarray (time_spmv 5, time_dotp 6,
// time_prec 7, time_comm_dotp 8, time_comm_spmv 9; src/define/
// def_solver_prm.f90:151-163, zeroed per solve as in monolis_timer_initialize
// :237-251).  The split is measured when the caller asks for timing
// (Iarray is_measurement 13, show_time 21 or show_time_statistics 23): an event
// after every solver kernel charges the interval to the kernel's class (SpMV,
// vector update / dot product, preconditioner apply).  Without those flags the
// slots stay 0, so no events sit between the kernels of the hot loop.  One
// GPU exchanges no halo and reduces without MPI: the communication slots are 0.
struct SplitTimer {
  std::unique_ptr<mh::KTimer> t;
  SplitTimer(const int* Iarray, double* Rarray, hipStream_t s) {
    for (int k = R_TIME_SPMV; k <= R_TIME_COMM_SPMV; ++k) Rarray[k] = 0.0;
    if (Iarray[I_IS_MEASUREMENT] || Iarray[I_SHOW_TIME] || Iarray[I_SHOW_TIME_STATISTICS]) {
      t.reset(new mh::KTimer(s));
      mh::active_ktimer() = t.get();
    }
  }
  void finish(double* Rarray) {
    if (!t) return;
    mh::active_ktimer() = nullptr;
    t->harvest(true);
    if (t->err) return;
    Rarray[R_TIME_SPMV] = t->ms[mh::kClsSpmv] * 1e-3;
    Rarray[R_TIME_DOTP] = t->ms[mh::kClsDot] * 1e-3;
    Rarray[R_TIME_PREC] = t->ms[mh::kClsPrec] * 1e-3;
  }
  ~SplitTimer() {
    if (t) mh::active_ktimer() = nullptr;
  }
};

// Common device-resident solve with matrix A (the handle's matrix or its
// transpose).  Fills Iarray/Rarray; returns a status code.
int solve_device(Matrix& A, mh::Workspace& W, const double* b, double* x, int* Iarray, double* Rarray,
                 std::vector<double>* hist_out) {
  Context* c = nullptr;
  if (mh::context(&c)) return ST_HIP;
  const mh::SolveParams P = params_from(Iarray, Rarray);
  int rc = check_method(P.method, P.precond);
  if (rc) return rc;
  mh::SolveResult res;
  SplitTimer timer(Iarray, Rarray, c->stream);
  rc = mh::krylov_solve<double>(A, P, b, x, W, &res, c->stream);
  if (rc) return rc == -3 ? ST_PRECOND : rc == -4 ? ST_METHOD : rc == -2 ? ST_ARGS : ST_HIP;
  timer.finish(Rarray);
  if (!res.zero_rhs) {
    Iarray[I_CUR_ITER] = res.iter;
    Rarray[R_CUR_RESID] = res.resid;
  } else {
    fprintf(stderr, "** monolis error: norm of RHS vector B is 0.0\n");  // mat_converge.f90:41-43
  }
  if (hist_out) *hist_out = std::move(res.hist);
  return res.converged ? ST_OK : ST_NOT_CONVERGED;
}

// Complex counterpart of solve_device: COCG on the handle's complex matrix.
int solve_device_C(Matrix& A, mh::Workspace& W, const mh::zd* b, mh::zd* x, int* Iarray, double* Rarray,
                   std::vector<double>* hist_out) {
  Context* c = nullptr;
  if (mh::context(&c)) return ST_HIP;
  const mh::SolveParams P = params_from(Iarray, Rarray);
  mh::ZSolveResult res;
  SplitTimer timer(Iarray, Rarray, c->stream);
  const int rc = mh::cocg_solve(A, P, b, x, W, &res, c->stream);
  if (rc) return rc == -3 ? ST_PRECOND : rc == -4 ? ST_METHOD : rc == -2 ? ST_ARGS : rc == -6 ? ST_ZERO_PIVOT : ST_HIP;
  timer.finish(Rarray);
  if (!res.zero_rhs) {
    Iarray[I_CUR_ITER] = res.iter;
    Rarray[R_CUR_RESID] = res.resid_re;  // mat_converge.f90:178: complex resid stored in a real slot
  } else {
    fprintf(stderr, "** monolis error: norm of RHS vector B is 0.0\n");  // mat_converge.f90:79-82
  }
  if (hist_out) *hist_out = std::move(res.hist);
  return res.converged ? ST_OK : ST_NOT_CONVERGED;
}

// Device memory of the allocation-free inner product
// (monolis_hip_inner_product_R_noalloc): the partial sums of the reduction
// and its result live in this file-scope array instead of memory the library
// allocates.  Each GPU holds its own copy until the process ends.  In host
// code the name is a host-side stand-in, never a device address: the address
// hipGetSymbolAddress returns for a device is looked up once per device at
// set-up (monolis_hip_init, with that device current) and only that address is
// handed to kernels and copies.
constexpr int kScratchSlots = 4096;  // doubles: up to 4095 partial sums + the result
static __device__ __align__(256) double scratch[kScratchSlots];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_addr[kMaxDevices] = {};

// Looks up this device's address of `scratch` (current device, once).
int scratch_setup() {
  int dev = 0;
  MH_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    mh::set_error("device %d: the allocation-free entry points support devices 0..%d", dev, kMaxDevices - 1);
    return ST_ARGS;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void* p = nullptr;
    MH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_addr[dev] = static_cast<double*>(p);
  }
  return 0;
}

double* scratch_addr(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch_addr[dev];
}

// Read-bandwidth probe kernel: every element of src is read once per launch,
// UNR non-temporal loads of type V in flight per thread over a grid-stride
// loop, summed so that no load can be dropped; each wavefront writes one double.
template <typename V, int UNR>
__global__ void __launch_bounds__(256) k_stream_read(long long nv, const V* __restrict__ src,
                                                     double* __restrict__ out) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double acc = 0.0;
  for (; i + (UNR - 1) * stride < nv; i += UNR * stride) {
    V v[UNR];
#pragma unroll
    for (int u = 0; u < UNR; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
    for (int u = 0; u < UNR; ++u) {
      if constexpr (sizeof(V) == 16) acc += v[u].x + v[u].y;
      else acc += v[u];
    }
  }
  for (; i < nv; i += stride) {
    const V v = src[i];
    if constexpr (sizeof(V) == 16) acc += v.x + v.y;
    else acc += v;
  }
  acc = mh::wave_sum(acc);
  if ((threadIdx.x & (mh::kWave - 1)) == 0) out[(size_t)blockIdx.x * (256 / mh::kWave) + threadIdx.x / mh::kWave] = acc;
}

}  // namespace

// Pattern of a simple mesh with per-node block sizes (monolis_get_nonzero_pattern_by_simple_mesh_V,
// src/matrix/spmat_nzpattern.f90:53-87): the same nodal graph; n_dof_index / n_dof_index2
// follow from the block sizes; values start at zero.
template <typename V>
int create_by_simple_mesh_V(mh::Matrix** out, int n_node, int n_base, const int* n_dof_list, int n_elem,
                            const int* elem, hipStream_t s) {
  *out = nullptr;
  if (n_node < 0 || !n_dof_list) {
    mh::set_error("invalid node count or null n_dof_list");
    return ST_ARGS;
  }
  mh::DevArray<int> index, item;
  int nz = 0;
  MH_TRY(mh::pattern_by_simple_mesh(n_node, n_base, n_elem, elem, index, item, &nz, s));
  std::vector<int> idx((size_t)n_node + 1), itm((size_t)nz + 1), list((size_t)n_node + 1);
  MH_HIP(hipMemcpyAsync(idx.data(), index.p, sizeof(int) * ((size_t)n_node + 1), hipMemcpyDeviceToHost, s));
  if (nz) MH_HIP(hipMemcpyAsync(itm.data(), item.p, sizeof(int) * (size_t)nz, hipMemcpyDeviceToHost, s));
  if (n_node) MH_HIP(hipMemcpyAsync(list.data(), n_dof_list, sizeof(int) * (size_t)n_node, hipMemcpyDefault, s));
  MH_HIP(hipStreamSynchronize(s));
  std::vector<int> dofi((size_t)n_node + 1, 0), idx2((size_t)nz + 1, 0);
  for (int i = 0; i < n_node; ++i) {
    if (list[i] < 1 || list[i] > mh::kMaxVBlock) {
      mh::set_error("block size %d of node %d outside 1..%d", list[i], i, mh::kMaxVBlock);
      return ST_ARGS;
    }
    dofi[i + 1] = dofi[i] + list[i];
  }
  long long acc = 0;
  for (int i = 0; i < n_node; ++i)
    for (int e = idx[i]; e < idx[i + 1]; ++e) {
      idx2[e] = (int)acc;
      acc += (long long)list[i] * list[itm[e] - 1];
      if (acc >= (1LL << 31)) {
        mh::set_error("variable-block values exceed 2^31 entries");
        return ST_ARGS;
      }
    }
  idx2[nz] = (int)acc;
  std::vector<V> zeros((size_t)acc + 1, V(0.0));
  if constexpr (sizeof(V) == sizeof(mh::zd))
    return mh::matrix_create_vblock_complex(out, n_node, n_node, nz, list.data(), dofi.data(), idx2.data(),
                                            idx.data(), itm.data(), zeros.data(), s);
  else
    return mh::matrix_create_vblock(out, n_node, n_node, nz, list.data(), dofi.data(), idx2.data(), idx.data(),
                                    itm.data(), zeros.data(), s);
}

// Solver-context tuning of the SpMV variant: a variant's time inside a Krylov
// loop (between the row kernels that rewrite its input and consume its output)
// does not follow its back-to-back time (100^3 BCSR3, block-Jacobi CG: variant
// 4 runs 20.1 ms, the back-to-back winner 16 runs 20.6 ms).  Times `reps`
// solves of exactly `iters` iterations (tolerance 0: never converged) per
// variant with b = 1, x0 = 0, and keeps the fastest for that precision.
template <typename T>
int autotune_solver_T(Matrix& M, mh::Workspace& W, const mh::SolveParams& P, int reps, int* best,
                             double* ms, hipStream_t s) {
  const size_t vec = (size_t)M.np * M.b;
  mh::DevArray<T> b, x;
  MH_TRY(b.alloc(vec));
  MH_TRY(x.alloc(vec));
  std::vector<T> ones(vec, (T)1);
  MH_HIP(hipMemcpyAsync(b.p, ones.data(), sizeof(T) * vec, hipMemcpyHostToDevice, s));
  MH_HIP(hipStreamSynchronize(s));
  int& var = sizeof(T) == 8 ? M.variant64 : M.variant32;
  const int keep = var;
  double best_ms = 1e300;
  int best_v = keep;
  for (int v = 0; v < mh::kNumVariants; ++v) {
    var = v;
    mh::SolveResult res;
    int rc = mh::krylov_solve<T>(M, P, b.p, x.p, W, &res, s);  // warm-up
    const auto t0 = std::chrono::steady_clock::now();
    for (int i = 0; i < reps && !rc; ++i) rc = mh::krylov_solve<T>(M, P, b.p, x.p, W, &res, s);
    if (rc) {
      var = keep;
      return rc;
    }
    const double per_it =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() /
        ((double)reps * std::max(P.max_iter, 1));
    if (ms) ms[v] = per_it;
    if (per_it < best_ms) {
      best_ms = per_it;
      best_v = v;
    }
  }
  var = best_v;
  if (best) *best = best_v;
  return 0;
}

extern "C" {

const char* monolis_hip_last_error(void) { return mh::last_error(); }

int monolis_hip_init(void) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  return scratch_setup();  // the context's device is current here
}

void* monolis_hip_stream(void) {
  Context* c = nullptr;
  if (mh::context(&c)) return nullptr;
  return (void*)c->stream;
}

int monolis_hip_synchronize(void) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_malloc(void** ptr, size_t bytes) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_HIP(hipMalloc(ptr, bytes ? bytes : 1));
  return 0;
}

int monolis_hip_free(void* ptr) {
  if (!ptr) return 0;
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_HIP(hipStreamSynchronize(c->stream));
  MH_HIP(hipFree(ptr));
  return 0;
}

int monolis_hip_memcpy(void* dst, const void* src, size_t bytes) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (bytes == 0) return 0;
  if (mh::is_device_pointer(dst) && !mh::is_device_pointer(src)) {
    MH_TRY(mh::upload(dst, src, bytes, c->stream));  // large pageable sources: threaded
  } else {
    MH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, c->stream));
  }
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_memset(void* dst, int value, size_t bytes) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (bytes == 0) return 0;
  MH_HIP(hipMemsetAsync(dst, value, bytes, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_matrix_create(monolis_hip_matrix** out, int n, int np, int nz, int n_dof, const int* index,
                              const int* item, const double* A) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (!out) {
    mh::set_error("null output handle");
    return ST_ARGS;
  }
  *out = nullptr;
  if (n_dof == -1) {
    mh::set_error("variable block sizes (n_dof = -1) need monolis_hip_matrix_create_V");
    return ST_ARGS;
  }
  if (n_dof > mh::kMaxB) {
    UniformBlocks u;
    MH_TRY(u.make(np, nz, n_dof));
    return monolis_hip_matrix_create_V(out, n, np, nz, u.list.data(), u.dofi.data(), u.dofi2.data(), index, item, A);
  }
  Matrix* M = nullptr;
  MH_TRY(mh::matrix_create(&M, n, np, nz, n_dof, index, item, A, c->stream));
  auto* H = new monolis_hip_matrix();
  H->M = M;
  *out = H;
  return 0;
}

int monolis_hip_matrix_create_V(monolis_hip_matrix** out, int n, int np, int nz, const int* n_dof_list,
                                const int* n_dof_index, const int* n_dof_index2, const int* index, const int* item,
                                const double* A) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (!out) {
    mh::set_error("null output handle");
    return ST_ARGS;
  }
  *out = nullptr;
  Matrix* M = nullptr;
  MH_TRY(mh::matrix_create_vblock(&M, n, np, nz, n_dof_list, n_dof_index, n_dof_index2, index, item, A, c->stream));
  auto* H = new monolis_hip_matrix();
  H->M = M;
  *out = H;
  return 0;
}

int monolis_hip_matrix_destroy(monolis_hip_matrix* H) {
  if (!H) return 0;
  Context* c = nullptr;
  if (mh::context(&c) == 0) (void)hipStreamSynchronize(c->stream);
  delete H;
  return 0;
}

int monolis_hip_matrix_set_values(monolis_hip_matrix* H, const double* A) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (H->M->vb) MH_TRY(mh::matrix_set_values_vblock(H->M, A, c->stream));
  else MH_TRY(mh::matrix_set_values(H->M, A, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_matrix_info(const monolis_hip_matrix* H, long long* info) {
  const Matrix* M = H->M;
  info[0] = M->n;
  info[1] = M->np;
  info[2] = M->nz;
  info[3] = M->vb ? -1 : M->b;
  info[4] = M->nslices;
  info[5] = (long long)M->nslots;
  info[6] = (long long)(M->nslots * mh::kWave * (size_t)M->b * M->b * 8 + M->nslots * mh::kWave * 4);
  info[7] = (long long)(M->nslots * mh::kWave * (size_t)M->b * M->b * 4 + M->nslots * mh::kWave * 4);
  return 0;
}

int monolis_hip_matvec_R(monolis_hip_matrix* H, const double* x, double* y) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_TRY(mh::spmv<double>(*H->M, x, y, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_matvec_S(monolis_hip_matrix* H, const float* x, float* y) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_TRY(mh::matrix_make_fp32(H->M, c->stream));
  MH_TRY(mh::spmv<float>(*H->M, x, y, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_inner_product_R(int m, const double* x, const double* y, double* sum) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  mh::DevArray<double> parts;
  const size_t np = mh::partials_needed(m);
  MH_TRY(parts.alloc(np + 16));
  const double* dx = x;
  const double* dy = y;
  mh::DevArray<double> hx, hy;
  if (!mh::is_device_pointer(x)) {
    MH_TRY(hx.alloc((size_t)m));
    MH_TRY(mh::upload(hx.p, x, sizeof(double) * (size_t)m, c->stream));
    dx = hx.p;
  }
  if (!mh::is_device_pointer(y)) {
    MH_TRY(hy.alloc((size_t)m));
    MH_TRY(mh::upload(hy.p, y, sizeof(double) * (size_t)m, c->stream));
    dy = hy.p;
  }
  MH_TRY(mh::dot<double>(m, dx, dy, parts.p + np, parts.p, c->stream));
  MH_HIP(hipMemcpyAsync(sum, parts.p + np, sizeof(double), hipMemcpyDefault, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_inner_product_R_noalloc(int m, const double* x, const double* y, double* sum) {
  // no allocation here: the context must already exist (monolis_hip_init),
  // x and y must be device arrays, and the partials go to `scratch`
  Context* c = mh::context_if_ready();
  if (!c) {
    mh::set_error("monolis_hip_inner_product_R_noalloc: call monolis_hip_init first");
    return ST_ARGS;
  }
  double* sp = scratch_addr(c->device);
  if (!sp) {
    mh::set_error("monolis_hip_inner_product_R_noalloc: device %d was not set up by monolis_hip_init", c->device);
    return ST_ARGS;
  }
  if (m < 0 || !sum || (m > 0 && (!x || !y))) {
    mh::set_error("monolis_hip_inner_product_R_noalloc: invalid arguments (m = %d)", m);
    return ST_ARGS;
  }
  if (m > 0 && (!mh::is_device_pointer(x) || !mh::is_device_pointer(y))) {
    mh::set_error("monolis_hip_inner_product_R_noalloc: x and y must be device arrays (staging host data allocates)");
    return ST_ARGS;
  }
  const int nwg = mh::rows_grid<double, 1>(m);
  if (nwg + 1 > kScratchSlots) {
    mh::set_error("monolis_hip_inner_product_R_noalloc: %d partial sums exceed the %d-slot scratch array", nwg,
                  kScratchSlots - 1);
    return ST_ARGS;
  }
  // partials in sp[0 .. nwg), every one written by the launch before it is
  // read, so nothing needs clearing; the result in the last slot unless sum is
  // itself device memory
  const bool dev_sum = mh::is_device_pointer(sum) != 0;
  double* out = dev_sum ? sum : sp + (kScratchSlots - 1);
  MH_TRY(mh::dot<double>(m, x, y, out, sp, c->stream));
  if (!dev_sum) MH_HIP(hipMemcpyAsync(sum, out, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_solve_R(monolis_hip_matrix* H, const double* b, double* x, int* Iarray, double* Rarray,
                        double* hist, int hist_len) {
  const double t0 = now();
  std::vector<double> h;
  const int st = solve_device(*H->M, H->W, b, x, Iarray, Rarray, &h);
  Iarray[I_IERR] = st;
  Rarray[R_TIME_SOL] = now() - t0;
  if (hist && hist_len > 0) {
    for (int k = 0; k < hist_len; ++k) hist[k] = (k < (int)h.size()) ? h[k] : 0.0;
  }
  return st < 0 ? st : 0;
}

int monolis_hip_solve_refined_R(monolis_hip_matrix* H, const double* b, double* x, int* Iarray, double* Rarray,
                                double inner_tol, int max_outer, double* stats) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  const int rc0 = check_method(Iarray[I_METHOD], Iarray[I_PRECOND]);
  if (rc0) {
    Iarray[I_IERR] = rc0;
    return rc0;
  }
  const double t0 = now();
  mh::RefineParams P;
  P.tol = Rarray[R_TOL];
  P.inner_tol = inner_tol;
  P.max_outer = max_outer;
  P.inner_max_iter = Iarray[I_MAX_ITER];
  P.method = Iarray[I_METHOD];
  P.precond = Iarray[I_PRECOND];
  P.init_x = Iarray[I_IS_INIT_X] != 0;
  mh::RefineResult R;
  SplitTimer timer(Iarray, Rarray, c->stream);
  const int rc = mh::refine_solve(*H->M, P, b, x, H->W, H->W_outer, &R, c->stream);
  if (rc) {
    Iarray[I_IERR] = ST_HIP;
    return rc;
  }
  timer.finish(Rarray);
  Iarray[I_CUR_ITER] = R.inner_total;
  Rarray[R_CUR_RESID] = R.resid;
  Iarray[I_IERR] = R.converged ? ST_OK : ST_NOT_CONVERGED;
  Rarray[R_TIME_SOL] = now() - t0;
  if (stats) {
    stats[0] = R.outer;
    stats[1] = R.inner_total;
    stats[2] = R.resid;
    stats[3] = R.converged;
  }
  return 0;
}

int monolis_hip_time_matvec(monolis_hip_matrix* H, int fp32, const void* x, void* y, int warmup, int reps,
                            double* ms) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (fp32) MH_TRY(mh::matrix_make_fp32(H->M, c->stream));
  auto one = [&]() -> int {
    if (fp32) return mh::spmv<float>(*H->M, (const float*)x, (float*)y, c->stream);
    return mh::spmv<double>(*H->M, (const double*)x, (double*)y, c->stream);
  };
  for (int i = 0; i < warmup; ++i) MH_TRY(one());
  hipEvent_t e0, e1;
  MH_HIP(hipEventCreate(&e0));
  MH_HIP(hipEventCreate(&e1));
  MH_HIP(hipEventRecord(e0, c->stream));
  for (int i = 0; i < reps; ++i) MH_TRY(one());
  MH_HIP(hipEventRecord(e1, c->stream));
  MH_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  MH_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = reps > 0 ? (double)t / reps : 0.0;
  return 0;
}

int monolis_hip_time_matvec_each(monolis_hip_matrix* H, int fp32, const void* x, void* y, int warmup, int reps,
                                 double* ms_each) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (reps < 0) {
    mh::set_error("reps < 0");
    return ST_ARGS;
  }
  if (fp32) MH_TRY(mh::matrix_make_fp32(H->M, c->stream));
  auto one = [&]() -> int {
    if (fp32) return mh::spmv<float>(*H->M, (const float*)x, (float*)y, c->stream);
    return mh::spmv<double>(*H->M, (const double*)x, (double*)y, c->stream);
  };
  for (int i = 0; i < warmup; ++i) MH_TRY(one());
  std::vector<hipEvent_t> ev(2 * (size_t)reps);
  for (auto& e : ev) MH_HIP(hipEventCreate(&e));
  int rc = 0;
  for (int i = 0; i < reps && !rc; ++i) {
    if (hipEventRecord(ev[2 * i], c->stream) != hipSuccess) rc = ST_HIP;
    if (!rc) rc = one();
    if (!rc && hipEventRecord(ev[2 * i + 1], c->stream) != hipSuccess) rc = ST_HIP;
  }
  if (!rc && hipStreamSynchronize(c->stream) != hipSuccess) rc = ST_HIP;
  for (int i = 0; i < reps && !rc; ++i) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, ev[2 * i], ev[2 * i + 1]) != hipSuccess) rc = ST_HIP;
    ms_each[i] = (double)t;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  if (rc == ST_HIP) mh::set_error("event timing failed");
  return rc;
}

int monolis_hip_num_variants(void) { return mh::kNumVariants; }

int monolis_hip_matrix_set_variant(monolis_hip_matrix* H, int fp32, int variant) {
  if (variant < 0 || variant >= mh::kNumVariants) {
    mh::set_error("variant %d out of range [0, %d)", variant, mh::kNumVariants);
    return ST_ARGS;
  }
  if (fp32) H->M->variant32 = variant;
  else H->M->variant64 = variant;
  return 0;
}

int monolis_hip_matrix_autotune(monolis_hip_matrix* H, int fp32, int reps, int* best, double* ms) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  Matrix& M = *H->M;
  if (reps < 1) reps = 1;
  const size_t vec = (size_t)M.np * M.b;
  mh::DevArray<double> x, y;
  MH_TRY(x.alloc(vec));
  MH_TRY(y.alloc(vec));
  MH_HIP(hipMemsetAsync(x.p, 0, sizeof(double) * vec, c->stream));
  if (fp32) MH_TRY(mh::matrix_make_fp32(&M, c->stream));
  hipEvent_t e0, e1;
  MH_HIP(hipEventCreate(&e0));
  MH_HIP(hipEventCreate(&e1));
  int rc = 0;
  for (int prec = 0; prec < (fp32 ? 2 : 1) && !rc; ++prec) {
    double best_ms = 1e300;
    int best_v = 0;
    for (int v = 0; v < mh::kNumVariants && !rc; ++v) {
      auto one = [&]() -> int {
        if (prec == 0) return mh::spmv_variant<double>(M, x.p, y.p, v, c->stream);
        return mh::spmv_variant<float>(M, (const float*)x.p, (float*)y.p, v, c->stream);
      };
      for (int i = 0; i < 3 && !rc; ++i) rc = one();
      if (rc) break;
      if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = -1; break; }
      for (int i = 0; i < reps && !rc; ++i) rc = one();
      if (rc) break;
      float t = 0.f;
      if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
          hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
        mh::set_error("autotune timing failed");
        rc = -1;
        break;
      }
      const double avg = (double)t / reps;
      if (ms) ms[prec * mh::kNumVariants + v] = avg;
      if (avg < best_ms) {
        best_ms = avg;
        best_v = v;
      }
    }
    if (!rc) {
      if (prec == 0) M.variant64 = best_v;
      else M.variant32 = best_v;
      if (best) best[prec] = best_v;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

int monolis_hip_matrix_autotune_solver(monolis_hip_matrix* H, int fp32, int method, int precond, int iters,
                                       int reps, int* best, double* ms) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (!H || !H->M || iters < 1 || reps < 1) {
    mh::set_error("autotune_solver: needs a matrix handle, iters >= 1 and reps >= 1");
    return ST_ARGS;
  }
  int rc = check_method(method, precond);
  if (rc) return rc;
  mh::SolveParams P;
  P.method = method;
  P.precond = precond;
  P.max_iter = iters;
  P.tol = 0.0;
  P.keep_hist = 0;
  P.rhs_nonzero = 1;
  Matrix& M = *H->M;
  if (fp32) rc = autotune_solver_T<float>(M, H->W, P, reps, best, ms, c->stream);
  else rc = autotune_solver_T<double>(M, H->W, P, reps, best, ms, c->stream);
  return rc ? (rc == -3 ? ST_PRECOND : rc == -4 ? ST_METHOD : rc == -2 ? ST_ARGS : ST_HIP) : 0;
}

static mh::BatchParams batch_params(const int* Iarray, const double* Rarray) {
  mh::BatchParams P;
  P.max_iter = Iarray[I_MAX_ITER];
  P.tol = Rarray[R_TOL];
  P.init_x = Iarray[I_IS_INIT_X] != 0;
  P.precond = Iarray[I_PRECOND];
  return P;
}

static int batch_check(const int* Iarray) {
  if (Iarray[I_METHOD] != 1) {
    mh::set_error("batched solves implement CG (method 1), got %d", Iarray[I_METHOD]);
    return ST_METHOD;
  }
  if (Iarray[I_PRECOND] != 0 && Iarray[I_PRECOND] != 1) {
    mh::set_error("batched solves implement NONE / DIAG, got %d", Iarray[I_PRECOND]);
    return ST_PRECOND;
  }
  if (Iarray[I_MAX_ITER] < 1) {
    mh::set_error("max_iter must be >= 1");
    return ST_ARGS;
  }
  return 0;
}

int monolis_hip_batch_create(monolis_hip_batch** out, int nsys, const int* sys_row, int n, int nz, const int* index,
                             const int* item, const double* A) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  *out = nullptr;
  mh::Batch* B = nullptr;
  MH_TRY(mh::batch_create(&B, nsys, sys_row, n, nz, index, item, A, c->stream));
  auto* H = new monolis_hip_batch();
  H->B = B;
  *out = H;
  return 0;
}

int monolis_hip_batch_destroy(monolis_hip_batch* H) {
  if (!H) return 0;
  Context* c = nullptr;
  if (mh::context(&c) == 0) (void)hipStreamSynchronize(c->stream);
  delete H;
  return 0;
}

int monolis_hip_batch_num_configs(void) { return mh::batch_num_configs(); }

int monolis_hip_batch_config(int cfg, int* threads, int* unroll) {
  if (cfg < 0 || cfg >= mh::batch_num_configs()) {
    mh::set_error("batch config %d out of range", cfg);
    return ST_ARGS;
  }
  mh::batch_config_info(cfg, threads, unroll);
  return 0;
}

int monolis_hip_batch_set_config(monolis_hip_batch* H, int cfg) {
  if (!H || cfg < 0 || cfg >= mh::batch_num_configs()) {
    mh::set_error("batch config %d out of range", cfg);
    return ST_ARGS;
  }
  H->B->cfg = cfg;
  return 0;
}

int monolis_hip_batch_schedule(monolis_hip_batch* H, int cfg, int* split, int* nslot, int* nteam) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (!H) {
    mh::set_error("null batch");
    return ST_ARGS;
  }
  const int rc = mh::batch_schedule_info(*H->B, cfg, split, nslot, nteam, c->stream);
  return rc == -2 ? ST_ARGS : rc ? ST_HIP : 0;
}

int monolis_hip_batch_set_wait_limit(monolis_hip_batch* H, double seconds) {
  if (!H || !(seconds >= 0.0) || seconds > 600.0) {
    mh::set_error("team wait limit must lie in [0, 600] s");
    return ST_ARGS;
  }
  H->B->wait_s = seconds;
  return 0;
}

int monolis_hip_batch_solve(monolis_hip_batch* H, const double* b, double* x, int* Iarray, double* Rarray, int* iters,
                            double* resid, int* status) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  const int rc0 = batch_check(Iarray);
  if (rc0) return rc0;
  const double t0 = now();
  mh::Batch& B = *H->B;
  MH_TRY(mh::batch_solve(B, batch_params(Iarray, Rarray), b, x, B.cfg, 1, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  if (iters) MH_HIP(hipMemcpy(iters, B.iters.p, sizeof(int) * B.nsys, hipMemcpyDeviceToHost));
  if (resid) MH_HIP(hipMemcpy(resid, B.resid.p, sizeof(double) * B.nsys, hipMemcpyDeviceToHost));
  if (status) MH_HIP(hipMemcpy(status, B.status.p, sizeof(int) * B.nsys, hipMemcpyDeviceToHost));
  Rarray[R_TIME_SOL] = now() - t0;
  int aborted = 0;
  MH_TRY(mh::batch_aborted(B, B.cfg, &aborted, c->stream));
  if (aborted) {
    mh::set_error("batched solve: a team of workgroups gave up waiting for a member (systems with status 3)");
    return ST_HIP;
  }
  return 0;
}

int monolis_hip_batch_time(monolis_hip_batch* H, const double* b, double* x, int* Iarray, double* Rarray, int reps,
                           double* ms) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  const int rc0 = batch_check(Iarray);
  if (rc0) return rc0;
  mh::Batch& B = *H->B;
  const mh::BatchParams P = batch_params(Iarray, Rarray);
  MH_TRY(mh::batch_solve(B, P, b, x, B.cfg, 1, c->stream));  // capture + warm-up
  hipEvent_t e0, e1;
  MH_HIP(hipEventCreate(&e0));
  MH_HIP(hipEventCreate(&e1));
  MH_HIP(hipEventRecord(e0, c->stream));
  for (int i = 0; i < reps; ++i) MH_TRY(mh::batch_solve(B, P, b, x, B.cfg, 1, c->stream));
  MH_HIP(hipEventRecord(e1, c->stream));
  MH_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  MH_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = reps > 0 ? (double)t / reps : 0.0;
  int aborted = 0;
  MH_TRY(mh::batch_aborted(B, B.cfg, &aborted, c->stream));
  if (aborted) {
    mh::set_error("batched solve: a team of workgroups gave up waiting for a member");
    return ST_HIP;
  }
  return 0;
}

int monolis_hip_batch_autotune(monolis_hip_batch* H, const double* b, double* x, int* Iarray, double* Rarray, int reps,
                               int* best, double* ms) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  const int rc0 = batch_check(Iarray);
  if (rc0) return rc0;
  mh::Batch& B = *H->B;
  const mh::BatchParams P = batch_params(Iarray, Rarray);
  double best_ms = 1e300;
  int best_cfg = 0;
  for (int cfg = 0; cfg < mh::batch_num_configs(); ++cfg) {
    MH_TRY(mh::batch_solve(B, P, b, x, cfg, 0, c->stream));  // warm-up
    hipEvent_t e0, e1;
    MH_HIP(hipEventCreate(&e0));
    MH_HIP(hipEventCreate(&e1));
    MH_HIP(hipEventRecord(e0, c->stream));
    for (int i = 0; i < reps; ++i) MH_TRY(mh::batch_solve(B, P, b, x, cfg, 0, c->stream));
    MH_HIP(hipEventRecord(e1, c->stream));
    MH_HIP(hipEventSynchronize(e1));
    float t = 0.f;
    MH_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    int aborted = 0;
    MH_TRY(mh::batch_aborted(B, cfg, &aborted, c->stream));
    const double avg = aborted ? 1e30 : reps > 0 ? (double)t / reps : 0.0;  // a configuration that gave up is never chosen
    if (ms) ms[cfg] = avg;
    if (avg < best_ms) {
      best_ms = avg;
      best_cfg = cfg;
    }
  }
  B.cfg = best_cfg;
  if (best) *best = best_cfg;
  return 0;
}

static int multi_params(const int* Iarray, const double* Rarray, mh::MultiParams* P) {
  if (Iarray[I_METHOD] != 1) {
    mh::set_error("multi-RHS solves implement CG (method 1), got %d", Iarray[I_METHOD]);
    return ST_METHOD;
  }
  if (Iarray[I_PRECOND] != 0 && Iarray[I_PRECOND] != 1) {
    mh::set_error("multi-RHS solves implement NONE / DIAG, got %d", Iarray[I_PRECOND]);
    return ST_PRECOND;
  }
  P->max_iter = Iarray[I_MAX_ITER];
  P->precond = Iarray[I_PRECOND];
  P->init_x = Iarray[I_IS_INIT_X] != 0;
  P->tol = Rarray[R_TOL];
  return 0;
}

// Iarray[4] / Rarray[2] / Iarray[5] summarise the columns: largest iteration
// count, largest relative residual, not converged if any column is not.
static void multi_summary(int nrhs, const std::vector<int>& it, const std::vector<double>& res,
                          const std::vector<int>& st, int* Iarray, double* Rarray, int* iters, double* resid,
                          int* status) {
  int imax = 0, any_nc = 0;
  double rmax = 0.0;
  for (int j = 0; j < nrhs; ++j) {
    imax = std::max(imax, it[j]);
    if (st[j] != 2) rmax = std::max(rmax, res[j]);
    any_nc |= st[j] == 1;
    if (iters) iters[j] = it[j];
    if (resid) resid[j] = res[j];
    if (status) status[j] = st[j];
  }
  Iarray[I_CUR_ITER] = imax;
  Rarray[R_CUR_RESID] = rmax;
  Iarray[I_IERR] = any_nc ? ST_NOT_CONVERGED : ST_OK;
}

int monolis_hip_solve_multi_R(monolis_hip_matrix* H, int nrhs, const double* b, double* x, int* Iarray,
                              double* Rarray, int* iters, double* resid, int* status) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  mh::MultiParams P;
  int rc = multi_params(Iarray, Rarray, &P);
  if (rc) {
    Iarray[I_IERR] = rc;
    return rc;
  }
  const double t0 = now();
  mh::MultiResult R;
  rc = mh::multi_cg<double>(*H->M, P, nrhs, b, x, H->W, &R, c->stream);
  if (rc) {
    Iarray[I_IERR] = rc == -3 ? ST_PRECOND : rc == -2 ? ST_ARGS : ST_HIP;
    return Iarray[I_IERR];
  }
  multi_summary(nrhs, R.iters, R.resid, R.status, Iarray, Rarray, iters, resid, status);
  Rarray[R_TIME_SOL] = now() - t0;
  return 0;
}

int monolis_hip_solve_multi_refined_R(monolis_hip_matrix* H, int nrhs, const double* b, double* x, int* Iarray,
                                      double* Rarray, double inner_tol, int max_outer, int* iters, double* resid,
                                      int* status, double* stats) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  mh::MultiParams P;
  int rc = multi_params(Iarray, Rarray, &P);
  if (rc) {
    Iarray[I_IERR] = rc;
    return rc;
  }
  const double t0 = now();
  mh::MultiRefineResult R;
  rc = mh::multi_refine(*H->M, P, inner_tol, max_outer, nrhs, b, x, H->W, H->W_outer, &R, c->stream);
  if (rc) {
    Iarray[I_IERR] = rc == -3 ? ST_PRECOND : rc == -2 ? ST_ARGS : ST_HIP;
    return Iarray[I_IERR];
  }
  multi_summary(nrhs, R.inner_iters, R.resid, R.status, Iarray, Rarray, iters, resid, status);
  if (stats) stats[0] = R.outer;
  Rarray[R_TIME_SOL] = now() - t0;
  return 0;
}

int monolis_hip_matvec_multi_R(monolis_hip_matrix* H, int nrhs, const double* x, double* y) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_TRY(mh::multi_matvec<double>(*H->M, nrhs, x, y, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_matvec_multi_S(monolis_hip_matrix* H, int nrhs, const float* x, float* y) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_TRY(mh::multi_matvec<float>(*H->M, nrhs, x, y, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

static int scalar_op(monolis_hip_matrix* H, int i, int j, int sub_i, int sub_j, double v, int mode, double* out,
                     int* is_find) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  int found = 0;
  double val = 0.0;
  MH_TRY(mh::matrix_scalar(H->M, i, j, sub_i, sub_j, v, mode, &val, &found, c->stream));
  if (out) *out = found ? val : 0.0;
  if (is_find) *is_find = found;
  if (!found && mode != 0) {
    // monolis stops with monolis_stop_by_matrix_assemble (spmat_handler_util.f90:72)
    mh::set_error("block (%d, %d) is not in the sparsity pattern", i, j);
    return ST_ARGS;
  }
  return 0;
}

int monolis_hip_matrix_set_scalar_R(monolis_hip_matrix* H, int i, int j, int sub_i, int sub_j, double val) {
  return scalar_op(H, i, j, sub_i, sub_j, val, 1, nullptr, nullptr);
}

int monolis_hip_matrix_add_scalar_R(monolis_hip_matrix* H, int i, int j, int sub_i, int sub_j, double val) {
  return scalar_op(H, i, j, sub_i, sub_j, val, 2, nullptr, nullptr);
}

int monolis_hip_matrix_get_scalar_R(monolis_hip_matrix* H, int i, int j, int sub_i, int sub_j, double* val,
                                    int* is_find) {
  return scalar_op(H, i, j, sub_i, sub_j, 0.0, 0, val, is_find);
}

static int square_check(const Matrix& M) {
  if (M.n != M.np) {
    mh::set_error("adjoint operations need a square local matrix (n == np), got n=%d np=%d", M.n, M.np);
    return ST_ARGS;
  }
  return 0;
}

int monolis_hip_solve_transpose_R(monolis_hip_matrix* H, const double* b, double* x, int* Iarray, double* Rarray) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_TRY(square_check(*H->M));
  const double t0 = now();
  MH_TRY(mh::matrix_transpose(H->M, c->stream));
  const int st = solve_device(*H->M->transpose, H->W, b, x, Iarray, Rarray, nullptr);
  Iarray[I_IERR] = st;
  Rarray[R_TIME_SOL] = now() - t0;
  return st < 0 ? st : 0;
}

int monolis_hip_outer_grad_R(monolis_hip_matrix* H, const double* lambda, const double* x, double* dA) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_TRY(square_check(*H->M));
  MH_TRY(mh::outer_grad(*H->M, lambda, x, dA, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_adjoint_R(monolis_hip_matrix* H, const double* x, const double* g, double* lambda, double* dA,
                          int symmetric, int* Iarray, double* Rarray) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_TRY(square_check(*H->M));
  const double t0 = now();
  Matrix* A = H->M;
  if (!symmetric) {
    MH_TRY(mh::matrix_transpose(H->M, c->stream));
    A = H->M->transpose;
  }
  const int st = solve_device(*A, H->W, g, lambda, Iarray, Rarray, nullptr);
  Iarray[I_IERR] = st;
  if (st < 0) return st;
  if (dA) {
    MH_TRY(mh::outer_grad(*H->M, lambda, x, dA, c->stream));
    MH_HIP(hipStreamSynchronize(c->stream));
  }
  Rarray[R_TIME_SOL] = now() - t0;
  return 0;
}

int monolis_hip_multi_num_configs(void) { return mh::multi_num_configs(); }

int monolis_hip_multi_config(int cfg, int* steps_per_batch, int* wg_per_cu, int* chunk, int* threads, int* lds) {
  if (cfg < 0 || cfg >= mh::multi_num_configs()) {
    mh::set_error("multi-RHS config %d out of range", cfg);
    return ST_ARGS;
  }
  mh::multi_config_info(cfg, steps_per_batch, wg_per_cu, chunk, threads, lds);
  return 0;
}

int monolis_hip_multi_set_config(monolis_hip_matrix* H, int fp32, int cfg) {
  if (cfg < 0 || cfg >= mh::multi_num_configs()) {
    mh::set_error("multi-RHS config %d out of range", cfg);
    return ST_ARGS;
  }
  if (fp32) H->M->mrhs_cfg32 = cfg;
  else H->M->mrhs_cfg64 = cfg;
  return 0;
}

int monolis_hip_multi_autotune(monolis_hip_matrix* H, int nrhs, int fp32, int reps, int* best, double* ms) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  int b = 0;
  MH_TRY(mh::multi_autotune(*H->M, nrhs, fp32, reps, &b, ms, c->stream));
  if (best) *best = b;
  return 0;
}

__global__ void __launch_bounds__(256) k_stream_copy(long long n, const double* __restrict__ src,
                                                     double* __restrict__ dst) {
  const long long stride = (long long)gridDim.x * blockDim.x * 2;
  for (long long i = ((long long)blockIdx.x * blockDim.x + threadIdx.x) * 2; i < n; i += stride) {
    if (i + 1 < n) {
      const double2 v = *reinterpret_cast<const double2*>(src + i);
      *reinterpret_cast<double2*>(dst + i) = v;
    } else {
      dst[i] = src[i];
    }
  }
}

int monolis_hip_stream_copy(double* dst, const double* src, long long n, int reps, double* ms) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (n <= 0) {
    *ms = 0.0;
    return 0;
  }
  const int grid = 256 * 8 * 4;  // 8 waves per CU worth of workgroups on 256 CUs, grid-stride
  k_stream_copy<<<grid, 256, 0, c->stream>>>(n, src, dst);
  MH_HIP(hipGetLastError());
  hipEvent_t e0, e1;
  MH_HIP(hipEventCreate(&e0));
  MH_HIP(hipEventCreate(&e1));
  MH_HIP(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps; ++r) k_stream_copy<<<grid, 256, 0, c->stream>>>(n, src, dst);
  MH_HIP(hipEventRecord(e1, c->stream));
  MH_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  MH_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = reps > 0 ? (double)t / reps : 0.0;
  return 0;
}

int monolis_hip_stream_read(const double* src, long long n, int reps, double* ms) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  *ms = 0.0;
  if (n < 2 || reps <= 0) return 0;
  // shapes tried: 8-byte loads x 8 in flight and 16-byte loads x 4 in flight,
  // each at 2, 4 and 8 workgroups of 256 threads per CU; the fastest is reported
  const int cus = mh::cu_count();
  mh::DevArray<double> out;
  MH_TRY(out.alloc((size_t)8 * cus * (256 / mh::kWave)));
  hipEvent_t e0, e1;
  MH_HIP(hipEventCreate(&e0));
  MH_HIP(hipEventCreate(&e1));
  double best = 0.0;
  int rc = 0;
  for (int shape = 0; shape < 6 && !rc; ++shape) {
    const int grid = (2 << (shape % 3)) * cus;
    const bool wide = shape >= 3;
    auto launch = [&]() {
      if (wide) k_stream_read<mh::d2v, 4><<<grid, 256, 0, c->stream>>>(n / 2, (const mh::d2v*)src, out.p);
      else k_stream_read<double, 8><<<grid, 256, 0, c->stream>>>(n, src, out.p);
    };
    launch();  // warm-up
    if (hipGetLastError() != hipSuccess || hipEventRecord(e0, c->stream) != hipSuccess) rc = ST_HIP;
    for (int r = 0; r < reps && !rc; ++r) launch();
    float t = 0.f;
    if (!rc && (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&t, e0, e1) != hipSuccess))
      rc = ST_HIP;
    const double per = (double)t / reps;
    if (!rc && (best == 0.0 || per < best)) best = per;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    mh::set_error("stream_read probe failed");
    return rc;
  }
  *ms = best;
  return 0;
}

// ---------------------------------------------------------------------------
// monolis flat ABI
// ---------------------------------------------------------------------------
void monolis_solve_R_c_main(int n, int np, int nz, int n_dof, int NPNDOF, int NZNDOF2, int* n_dof_list,
                            int* n_dof_index, int* n_dof_index2, double* A, double* x, double* b, int* index,
                            int* item, int my_rank, int comm, int comm_size, int recv_n_neib, int recv_nitem,
                            int* recv_neib_pe, int* recv_index, int* recv_item, int send_n_neib, int send_nitem,
                            int* send_neib_pe, int* send_index, int* send_item, int* Iarray, double* Rarray) {
  (void)NZNDOF2; (void)comm;
  (void)recv_n_neib; (void)recv_nitem; (void)recv_neib_pe; (void)recv_index; (void)recv_item;
  (void)send_n_neib; (void)send_nitem; (void)send_neib_pe; (void)send_index; (void)send_item;
  const double t0 = now();
  // monolis_timer_initialize (def_solver_prm.f90:237-251)
  for (int k = R_TIME_SOL; k <= R_TIME_COMM_SPMV; ++k) Rarray[k] = 0.0;
  const int abort_on_error = Iarray[I_IS_ERROR_ABORT] != 0;
  auto fail = [&](int code) {
    Iarray[I_IERR] = code;
    fprintf(stderr, "** monolis error: %s\n", mh::last_error());
    fflush(stderr);
    if (abort_on_error) error_stop("monolis_solve_R (GPU path) failed");
  };
  if (comm_size > 1) {
    mh::set_error("comm_size %d > 1: the GPU path runs on one rank", comm_size);
    fail(ST_MULTI_RANK);
    return;
  }
  int rc = check_method(Iarray[I_METHOD], Iarray[I_PRECOND]);
  if (rc) {
    fail(rc);
    return;
  }
  print_banner(Iarray, my_rank);
  Context* c = nullptr;
  if (mh::context(&c)) {
    fail(ST_HIP);
    return;
  }
  // staging: matrix (SELL build) and vectors, then the preconditioner
  const double tp0 = now();
  monolis_hip_matrix* H = nullptr;
  if (n_dof == -1)
    rc = monolis_hip_matrix_create_V(&H, n, np, nz, n_dof_list, n_dof_index, n_dof_index2, index, item, A);
  else
    rc = monolis_hip_matrix_create(&H, n, np, nz, n_dof, index, item, A);
  if (rc) {
    fail(rc == -2 ? ST_ARGS : ST_HIP);
    return;
  }
  std::unique_ptr<monolis_hip_matrix> guard(H);
  const size_t vec = (size_t)NPNDOF;
  double* dx = x;
  const double* db = b;
  mh::DevArray<double> bx, bb;
  const int x_dev = mh::is_device_pointer(x), b_dev = mh::is_device_pointer(b);
  if (!x_dev) {
    if (bx.alloc(vec) || mh::upload(bx.p, x, sizeof(double) * vec, c->stream)) {
      fail(ST_HIP);
      return;
    }
    dx = bx.p;
  }
  if (!b_dev) {
    if (bb.alloc(vec) || mh::upload(bb.p, b, sizeof(double) * vec, c->stream)) {
      fail(ST_HIP);
      return;
    }
    db = bb.p;
  }
  if (Iarray[I_PRECOND] == 1 && mh::diag_setup<double>(*H->M, c->stream)) {
    fail(ST_HIP);
    return;
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    mh::set_error("staging failed");
    fail(ST_HIP);
    return;
  }
  Rarray[R_TIME_PREP] = now() - tp0;
  std::vector<double> hist;
  const int st = solve_device(*H->M, H->W, db, dx, Iarray, Rarray, &hist);
  if (st < 0) {
    fail(st);
    return;
  }
  if (!x_dev) {
    if (hipMemcpyAsync(x, dx, sizeof(double) * vec, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      mh::set_error("copy of the solution failed");
      fail(ST_HIP);
      return;
    }
  }
  Iarray[I_IERR] = st;
  Rarray[R_TIME_SOL] = now() - t0;
  print_after(Iarray, Rarray, hist, my_rank);
  // mat_converge.f90:134-140; PipeCR has no such stop (PipeCR.f90:81-123 ends the loop silently)
  if (st == ST_NOT_CONVERGED && !Iarray[I_IS_MEASUREMENT] && abort_on_error && Iarray[I_METHOD] != 4)
    error_stop("reached the maximum number of iterations");
}

void monolis_matvec_product_R_c_main(int n, int np, int nz, int n_dof, int NPNDOF, int NZNDOF2, int* n_dof_list,
                                     int* n_dof_index, int* n_dof_index2, double* A, double* x, double* y,
                                     int* index, int* item, int my_rank, int comm, int comm_size,
                                     int recv_n_neib, int recv_nitem, int* recv_neib_pe, int* recv_index,
                                     int* recv_item, int send_n_neib, int send_nitem, int* send_neib_pe,
                                     int* send_index, int* send_item) {
  (void)NZNDOF2; (void)my_rank; (void)comm;
  (void)recv_n_neib; (void)recv_nitem; (void)recv_neib_pe; (void)recv_index; (void)recv_item;
  (void)send_n_neib; (void)send_nitem; (void)send_neib_pe; (void)send_index; (void)send_item;
  if (comm_size > 1) error_stop("monolis_matvec_product_R (GPU path): comm_size > 1 is not supported");
  Context* c = nullptr;
  if (mh::context(&c)) error_stop(mh::last_error());
  monolis_hip_matrix* H = nullptr;
  const int rc = n_dof == -1 ? monolis_hip_matrix_create_V(&H, n, np, nz, n_dof_list, n_dof_index, n_dof_index2,
                                                           index, item, A)
                             : monolis_hip_matrix_create(&H, n, np, nz, n_dof, index, item, A);
  if (rc) error_stop(mh::last_error());
  std::unique_ptr<monolis_hip_matrix> guard(H);
  const size_t vec = (size_t)NPNDOF;
  // internal entries: N * n_dof, or n_dof_index(N) for variable blocks (the scalar rows)
  const size_t m = (size_t)H->M->n * (H->M->vb ? 1 : H->M->b);
  const double* dx = x;
  double* dy = y;
  mh::DevArray<double> bx, by;
  const int x_dev = mh::is_device_pointer(x), y_dev = mh::is_device_pointer(y);
  if (!x_dev) {
    if (bx.alloc(vec) || mh::upload(bx.p, x, sizeof(double) * vec, c->stream)) error_stop(mh::last_error());
    dx = bx.p;
  }
  if (!y_dev) {
    if (by.alloc(vec)) error_stop(mh::last_error());
    dy = by.p;
  }
  if (mh::spmv<double>(*H->M, dx, dy, c->stream)) error_stop(mh::last_error());
  if (!y_dev && m > 0) {
    if (hipMemcpyAsync(y, dy, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
      error_stop("copy of y failed");
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) error_stop("monolis_matvec_product_R (GPU path) failed");
}

void monolis_inner_product_R_c_main(int m, double* x, double* y, double* sum, int comm) {
  (void)comm;
  if (monolis_hip_inner_product_R(m, x, y, sum)) error_stop(mh::last_error());
}

// ---------------------------------------------------------------------------
// Device assembly
// ---------------------------------------------------------------------------
int monolis_hip_matrix_create_by_simple_mesh(monolis_hip_matrix** out, int n_node, int n_base, int n_dof,
                                             int n_elem, const int* elem) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (!out) {
    mh::set_error("null output handle");
    return ST_ARGS;
  }
  *out = nullptr;
  if (n_dof < 1 || n_dof > mh::kMaxB) {
    mh::set_error("block size %d outside 1..%d", n_dof, mh::kMaxB);
    return ST_ARGS;
  }
  mh::DevArray<int> index, item;
  int nz = 0;
  MH_TRY(mh::pattern_by_simple_mesh(n_node, n_base, n_elem, elem, index, item, &nz, c->stream));
  // values start at zero, as monolis_alloc_nonzero_pattern_mat_val_R leaves them
  mh::DevArray<double> zeros;
  const size_t nv = (size_t)nz * n_dof * n_dof;
  MH_TRY(zeros.alloc(nv));
  MH_HIP(hipMemsetAsync(zeros.p, 0, sizeof(double) * (nv ? nv : 1), c->stream));
  Matrix* M = nullptr;
  MH_TRY(mh::matrix_create(&M, n_node, n_node, nz, n_dof, index.p, item.p, zeros.p, c->stream));
  auto* H = new monolis_hip_matrix();
  H->M = M;
  *out = H;
  return 0;
}

int monolis_hip_matrix_create_by_simple_mesh_V(monolis_hip_matrix** out, int n_node, int n_base,
                                               const int* n_dof_list, int n_elem, const int* elem) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (!out) {
    mh::set_error("null output handle");
    return ST_ARGS;
  }
  *out = nullptr;
  Matrix* M = nullptr;
  MH_TRY(create_by_simple_mesh_V<double>(&M, n_node, n_base, n_dof_list, n_elem, elem, c->stream));
  auto* H = new monolis_hip_matrix();
  H->M = M;
  *out = H;
  return 0;
}

int monolis_hip_matrix_create_by_simple_mesh_V_C(monolis_hip_matrix_C** out, int n_node, int n_base,
                                                 const int* n_dof_list, int n_elem, const int* elem) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (!out) {
    mh::set_error("null output handle");
    return ST_ARGS;
  }
  *out = nullptr;
  Matrix* M = nullptr;
  MH_TRY(create_by_simple_mesh_V<mh::zd>(&M, n_node, n_base, n_dof_list, n_elem, elem, c->stream));
  auto* H = new monolis_hip_matrix_C();
  H->M = M;
  *out = H;
  return 0;
}

int monolis_hip_matrix_get_pattern(const monolis_hip_matrix* H, int* index, int* item) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  return mh::matrix_get_pattern(*H->M, index, item, c->stream);
}

int monolis_hip_matrix_get_values_R(const monolis_hip_matrix* H, double* A) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  return mh::matrix_get_values(*H->M, A, c->stream);
}

int monolis_hip_matrix_add_elements_R(monolis_hip_matrix* H, int n_elem, int n_base, const int* elem,
                                      const double* Ke) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  return mh::matrix_add_elements(H->M, n_elem, n_base, elem, Ke, c->stream);
}

int monolis_hip_set_Dirichlet_bc_R(monolis_hip_matrix* H, double* b, int n_bc, const int* node_id,
                                   const int* n_dof_bc, const double* val) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  return mh::matrix_dirichlet(H->M, b, n_bc, node_id, n_dof_bc, val, c->stream);
}

// ---------------------------------------------------------------------------
// Complex path
// ---------------------------------------------------------------------------
int monolis_hip_matrix_create_C(monolis_hip_matrix_C** out, int n, int np, int nz, int n_dof, const int* index,
                                const int* item, const double _Complex* A) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (!out) {
    mh::set_error("null output handle");
    return ST_ARGS;
  }
  *out = nullptr;
  if (n_dof == -1) {
    mh::set_error("variable block sizes (n_dof = -1) need monolis_hip_matrix_create_V_C");
    return ST_ARGS;
  }
  if (n_dof > mh::kMaxB) {
    UniformBlocks u;
    MH_TRY(u.make(np, nz, n_dof));
    return monolis_hip_matrix_create_V_C(out, n, np, nz, u.list.data(), u.dofi.data(), u.dofi2.data(), index, item,
                                         A);
  }
  Matrix* M = nullptr;
  MH_TRY(mh::matrix_create_complex(&M, n, np, nz, n_dof, index, item, reinterpret_cast<const double*>(A),
                                   c->stream));
  auto* H = new monolis_hip_matrix_C();
  H->M = M;
  *out = H;
  return 0;
}

int monolis_hip_matrix_create_V_C(monolis_hip_matrix_C** out, int n, int np, int nz, const int* n_dof_list,
                                  const int* n_dof_index, const int* n_dof_index2, const int* index, const int* item,
                                  const double _Complex* A) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (!out) {
    mh::set_error("null output handle");
    return ST_ARGS;
  }
  *out = nullptr;
  Matrix* M = nullptr;
  MH_TRY(mh::matrix_create_vblock_complex(&M, n, np, nz, n_dof_list, n_dof_index, n_dof_index2, index, item,
                                          reinterpret_cast<const mh::zd*>(A), c->stream));
  auto* H = new monolis_hip_matrix_C();
  H->M = M;
  *out = H;
  return 0;
}

int monolis_hip_matrix_destroy_C(monolis_hip_matrix_C* H) {
  if (!H) return 0;
  Context* c = nullptr;
  if (mh::context(&c) == 0) (void)hipStreamSynchronize(c->stream);
  delete H;
  return 0;
}

int monolis_hip_matrix_set_values_C(monolis_hip_matrix_C* H, const double _Complex* A) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (H->M->vb)
    MH_TRY(mh::matrix_set_values_vblock_complex(H->M, reinterpret_cast<const mh::zd*>(A), c->stream));
  else
    MH_TRY(mh::matrix_set_values(H->M, reinterpret_cast<const double*>(A), c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_matrix_create_by_simple_mesh_C(monolis_hip_matrix_C** out, int n_node, int n_base, int n_dof,
                                               int n_elem, const int* elem) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (!out) {
    mh::set_error("null output handle");
    return ST_ARGS;
  }
  *out = nullptr;
  if (n_dof < 1 || n_dof > mh::kMaxB) {
    mh::set_error("block size %d outside 1..%d", n_dof, mh::kMaxB);
    return ST_ARGS;
  }
  mh::DevArray<int> index, item;
  int nz = 0;
  MH_TRY(mh::pattern_by_simple_mesh(n_node, n_base, n_elem, elem, index, item, &nz, c->stream));
  // values start at zero (monolis_alloc_nonzero_pattern_mat_val_C)
  mh::DevArray<double> zeros;
  const size_t nd = (size_t)nz * n_dof * n_dof * 2;
  MH_TRY(zeros.alloc(nd ? nd : 1));
  MH_HIP(hipMemsetAsync(zeros.p, 0, sizeof(double) * (nd ? nd : 1), c->stream));
  Matrix* M = nullptr;
  MH_TRY(mh::matrix_create_complex(&M, n_node, n_node, nz, n_dof, index.p, item.p, zeros.p, c->stream));
  auto* H = new monolis_hip_matrix_C();
  H->M = M;
  *out = H;
  return 0;
}

int monolis_hip_matrix_get_pattern_C(const monolis_hip_matrix_C* H, int* index, int* item) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  return mh::matrix_get_pattern(*H->M, index, item, c->stream);
}

int monolis_hip_matrix_get_values_C(const monolis_hip_matrix_C* H, double _Complex* A) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  return mh::matrix_get_values_C(*H->M, reinterpret_cast<mh::zd*>(A), c->stream);
}

int monolis_hip_matrix_add_elements_C(monolis_hip_matrix_C* H, int n_elem, int n_base, const int* elem,
                                      const double _Complex* Ke) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  return mh::matrix_add_elements_C(H->M, n_elem, n_base, elem, reinterpret_cast<const mh::zd*>(Ke), c->stream);
}

int monolis_hip_set_Dirichlet_bc_C(monolis_hip_matrix_C* H, double _Complex* b, int n_bc, const int* node_id,
                                   const int* n_dof_bc, const double _Complex* val) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  return mh::matrix_dirichlet_C(H->M, reinterpret_cast<mh::zd*>(b), n_bc, node_id, n_dof_bc,
                                reinterpret_cast<const mh::zd*>(val), c->stream);
}

int monolis_hip_matrix_info_C(const monolis_hip_matrix_C* H, long long* info) {
  const Matrix* M = H->M;
  info[0] = M->n;
  info[1] = M->np;
  info[2] = M->nz;
  info[3] = M->b;
  info[4] = M->nslices;
  info[5] = (long long)M->nslots;
  info[6] = (long long)(M->nslots * mh::kWave * (size_t)M->b * M->b * 16 + M->nslots * mh::kWave * 4);
  info[7] = 0;
  return 0;
}

int monolis_hip_matvec_C(monolis_hip_matrix_C* H, const double _Complex* x, double _Complex* y) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  MH_TRY(mh::zspmv(*H->M, reinterpret_cast<const mh::zd*>(x), reinterpret_cast<mh::zd*>(y), c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_time_matvec_C(monolis_hip_matrix_C* H, const double _Complex* x, double _Complex* y, int warmup,
                              int reps, double* ms) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  const auto* dx = reinterpret_cast<const mh::zd*>(x);
  auto* dy = reinterpret_cast<mh::zd*>(y);
  for (int i = 0; i < warmup; ++i) MH_TRY(mh::zspmv(*H->M, dx, dy, c->stream));
  hipEvent_t e0, e1;
  MH_HIP(hipEventCreate(&e0));
  MH_HIP(hipEventCreate(&e1));
  MH_HIP(hipEventRecord(e0, c->stream));
  int rc = 0;
  for (int i = 0; i < reps && !rc; ++i) rc = mh::zspmv(*H->M, dx, dy, c->stream);
  MH_HIP(hipEventRecord(e1, c->stream));
  MH_HIP(hipEventSynchronize(e1));
  float t = 0.f;
  MH_HIP(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  *ms = reps > 0 ? (double)t / reps : 0.0;
  return 0;
}

int monolis_hip_inner_product_C(int m, const double _Complex* x, const double _Complex* y, double _Complex* sum) {
  Context* c = nullptr;
  MH_TRY(mh::context(&c));
  if (m < 0) {
    mh::set_error("negative length");
    return ST_ARGS;
  }
  const size_t np = mh::partials_needed(m) * 2;
  mh::DevArray<double> scratch;
  MH_TRY(scratch.alloc(np + 16));
  const mh::zd* dx = reinterpret_cast<const mh::zd*>(x);
  const mh::zd* dy = reinterpret_cast<const mh::zd*>(y);
  mh::DevArray<mh::zd> hx, hy;
  if (!mh::is_device_pointer(x)) {
    MH_TRY(hx.alloc((size_t)m));
    MH_TRY(mh::upload(hx.p, x, sizeof(mh::zd) * (size_t)m, c->stream));
    dx = hx.p;
  }
  if (!mh::is_device_pointer(y)) {
    MH_TRY(hy.alloc((size_t)m));
    MH_TRY(mh::upload(hy.p, y, sizeof(mh::zd) * (size_t)m, c->stream));
    dy = hy.p;
  }
  MH_TRY(mh::zdot(m, dx, dy, scratch.p + np, scratch.p, c->stream));
  MH_HIP(hipMemcpyAsync(sum, scratch.p + np, sizeof(mh::zd), hipMemcpyDefault, c->stream));
  MH_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int monolis_hip_solve_C(monolis_hip_matrix_C* H, const double _Complex* b, double _Complex* x, int* Iarray,
                        double* Rarray, double* hist, int hist_len) {
  const double t0 = now();
  std::vector<double> h;
  const int st = solve_device_C(*H->M, H->W, reinterpret_cast<const mh::zd*>(b), reinterpret_cast<mh::zd*>(x),
                                Iarray, Rarray, &h);
  Iarray[I_IERR] = st;
  Rarray[R_TIME_SOL] = now() - t0;
  if (hist && hist_len > 0) {
    for (int k = 0; k < hist_len; ++k) hist[k] = (k < (int)h.size()) ? h[k] : 0.0;
  }
  return st < 0 ? st : 0;
}

void monolis_solve_C_c_main(int n, int np, int nz, int n_dof, int NPNDOF, int NZNDOF2, int* n_dof_list,
                            int* n_dof_index, int* n_dof_index2, double _Complex* A, double _Complex* x,
                            double _Complex* b, int* index, int* item, int my_rank, int comm, int comm_size,
                            int recv_n_neib, int recv_nitem, int* recv_neib_pe, int* recv_index, int* recv_item,
                            int send_n_neib, int send_nitem, int* send_neib_pe, int* send_index, int* send_item,
                            int* Iarray, double* Rarray) {
  (void)NZNDOF2; (void)comm;
  (void)recv_n_neib; (void)recv_nitem; (void)recv_neib_pe; (void)recv_index; (void)recv_item;
  (void)send_n_neib; (void)send_nitem; (void)send_neib_pe; (void)send_index; (void)send_item;
  const double t0 = now();
  for (int k = R_TIME_SOL; k <= R_TIME_COMM_SPMV; ++k) Rarray[k] = 0.0;
  const int abort_on_error = Iarray[I_IS_ERROR_ABORT] != 0;
  auto fail = [&](int code) {
    Iarray[I_IERR] = code;
    fprintf(stderr, "** monolis error: %s\n", mh::last_error());
    fflush(stderr);
    if (abort_on_error) error_stop("monolis_solve_C (GPU path) failed");
  };
  if (comm_size > 1) {
    mh::set_error("comm_size %d > 1: the GPU path runs on one rank", comm_size);
    fail(ST_MULTI_RANK);
    return;
  }
  if (Iarray[I_METHOD] != 13) {
    // src/solver/solver.f90 monolis_solver_select_C: COCG only
    mh::set_error("monolis_solver_select_C: please select a linear solver for complex numbers (COCG = 13, got %d)",
                  Iarray[I_METHOD]);
    fail(ST_METHOD);
    return;
  }
  if (Iarray[I_PRECOND] != 0 && Iarray[I_PRECOND] != 1) {
    mh::set_error("preconditioner %d (%s) is not implemented on the complex GPU path (NONE = 0, DIAG = 1)",
                  Iarray[I_PRECOND], str_prec(Iarray[I_PRECOND]));
    fail(ST_PRECOND);
    return;
  }
  print_banner(Iarray, my_rank);
  Context* c = nullptr;
  if (mh::context(&c)) {
    fail(ST_HIP);
    return;
  }
  const double tp0 = now();
  monolis_hip_matrix_C* H = nullptr;
  int rc = n_dof == -1 ? monolis_hip_matrix_create_V_C(&H, n, np, nz, n_dof_list, n_dof_index, n_dof_index2, index,
                                                       item, A)
                       : monolis_hip_matrix_create_C(&H, n, np, nz, n_dof, index, item, A);
  if (rc) {
    fail(rc == -2 ? ST_ARGS : ST_HIP);
    return;
  }
  std::unique_ptr<monolis_hip_matrix_C> guard(H);
  const size_t vec = (size_t)NPNDOF;
  mh::zd* dx = reinterpret_cast<mh::zd*>(x);
  const mh::zd* db = reinterpret_cast<const mh::zd*>(b);
  mh::DevArray<mh::zd> bx, bb;
  const int x_dev = mh::is_device_pointer(x), b_dev = mh::is_device_pointer(b);
  if (!x_dev) {
    if (bx.alloc(vec) || mh::upload(bx.p, x, sizeof(mh::zd) * vec, c->stream)) {
      fail(ST_HIP);
      return;
    }
    dx = bx.p;
  }
  if (!b_dev) {
    if (bb.alloc(vec) || mh::upload(bb.p, b, sizeof(mh::zd) * vec, c->stream)) {
      fail(ST_HIP);
      return;
    }
    db = bb.p;
  }
  if (Iarray[I_PRECOND] == 1) {
    rc = mh::zdiag_setup(*H->M, c->stream);
    if (rc) {
      fail(rc == -6 ? ST_ZERO_PIVOT : ST_HIP);
      return;
    }
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    mh::set_error("staging failed");
    fail(ST_HIP);
    return;
  }
  Rarray[R_TIME_PREP] = now() - tp0;
  std::vector<double> hist;
  const int st = solve_device_C(*H->M, H->W, db, dx, Iarray, Rarray, &hist);
  if (st < 0) {
    fail(st);
    return;
  }
  if (!x_dev) {
    if (hipMemcpyAsync(x, dx, sizeof(mh::zd) * vec, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      mh::set_error("copy of the solution failed");
      fail(ST_HIP);
      return;
    }
  }
  Iarray[I_IERR] = st;
  Rarray[R_TIME_SOL] = now() - t0;
  print_after(Iarray, Rarray, hist, my_rank);
  if (st == ST_NOT_CONVERGED && !Iarray[I_IS_MEASUREMENT] && abort_on_error)
    error_stop("reached the maximum number of iterations");  // mat_converge.f90:189-194
}

void monolis_matvec_product_C_c_main(int n, int np, int nz, int n_dof, int NPNDOF, int NZNDOF2, int* n_dof_list,
                                     int* n_dof_index, int* n_dof_index2, double _Complex* A, double _Complex* x,
                                     double _Complex* y, int* index, int* item, int my_rank, int comm,
                                     int comm_size, int recv_n_neib, int recv_nitem, int* recv_neib_pe,
                                     int* recv_index, int* recv_item, int send_n_neib, int send_nitem,
                                     int* send_neib_pe, int* send_index, int* send_item) {
  (void)NZNDOF2; (void)my_rank; (void)comm;
  (void)recv_n_neib; (void)recv_nitem; (void)recv_neib_pe; (void)recv_index; (void)recv_item;
  (void)send_n_neib; (void)send_nitem; (void)send_neib_pe; (void)send_index; (void)send_item;
  if (comm_size > 1) error_stop("monolis_matvec_product_C (GPU path): comm_size > 1 is not supported");
  Context* c = nullptr;
  if (mh::context(&c)) error_stop(mh::last_error());
  monolis_hip_matrix_C* H = nullptr;
  const int rc = n_dof == -1 ? monolis_hip_matrix_create_V_C(&H, n, np, nz, n_dof_list, n_dof_index, n_dof_index2,
                                                             index, item, A)
                             : monolis_hip_matrix_create_C(&H, n, np, nz, n_dof, index, item, A);
  if (rc) error_stop(mh::last_error());
  std::unique_ptr<monolis_hip_matrix_C> guard(H);
  const size_t vec = (size_t)NPNDOF;
  const size_t m = (size_t)H->M->n * H->M->b;
  const mh::zd* dx = reinterpret_cast<const mh::zd*>(x);
  mh::zd* dy = reinterpret_cast<mh::zd*>(y);
  mh::DevArray<mh::zd> bx, by;
  const int x_dev = mh::is_device_pointer(x), y_dev = mh::is_device_pointer(y);
  if (!x_dev) {
    if (bx.alloc(vec) || mh::upload(bx.p, x, sizeof(mh::zd) * vec, c->stream)) error_stop(mh::last_error());
    dx = bx.p;
  }
  if (!y_dev) {
    if (by.alloc(vec)) error_stop(mh::last_error());
    dy = by.p;
  }
  if (mh::zspmv(*H->M, dx, dy, c->stream)) error_stop(mh::last_error());
  if (!y_dev && m > 0) {
    if (hipMemcpyAsync(y, dy, sizeof(mh::zd) * m, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
      error_stop("copy of y failed");
  }
  if (hipStreamSynchronize(c->stream) != hipSuccess) error_stop("monolis_matvec_product_C (GPU path) failed");
}

void monolis_inner_product_C_c_main(int m, double _Complex* x, double _Complex* y, double _Complex* sum, int comm) {
  (void)comm;
  if (monolis_hip_inner_product_C(m, x, y, sum)) error_stop(mh::last_error());
}

}  // extern "C"
