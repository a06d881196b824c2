// Multi-right-hand-side CG (SpMM) on one scalar matrix.
#pragma once

#include "mh_internal.hpp"

namespace mh {

// Largest number of right-hand sides one multi-RHS solve takes (one per lane).
constexpr int kMaxRhs = 64;

// SpMM launch configurations: steps (row batches) whose gathers are issued
// together x workgroups per CU of the persistent grid.
int multi_num_configs();
void multi_config_info(int cfg, int* steps_per_batch, int* wg_per_cu, int* chunk = nullptr,
                       int* threads = nullptr, int* lds = nullptr);
// Times `reps` SpMM launches (Q = A P with the (P, Q) partials, as inside CG)
// per configuration for k right-hand sides and keeps the fastest on M for that
// precision.  ms (optional) receives multi_num_configs() averages.
int multi_autotune(Matrix& M, int k, int fp32, int reps, int* best, double* ms, hipStream_t s);

// Y = A X for k columns (np x k row-major device arrays), asynchronous on s.
template <typename T>
int multi_matvec(Matrix& M, int k, const T* x, T* y, hipStream_t s);

struct MultiParams {
  int max_iter = 1000;
  int precond = 1;   // 0 NONE, 1 DIAG
  int init_x = 1;
  double tol = 1e-8;
  int chunk = 8;     // iterations enqueued between convergence polls
};

struct MultiResult {
  std::vector<int> iters;     // per right-hand side
  std::vector<double> resid;  // relative residual sqrt((r,r)/(r0,r0))
  std::vector<int> status;    // 0 converged, 1 not converged, 2 zero right-hand side (x = 0)
  int loop_iters = 0;         // iterations the shared loop ran (max over right-hand sides)
};

// X (np x k, right-hand sides innermost) solves A X = B column by column with
// monolis's CG; all device pointers.  Buffers are carved from W.
template <typename T>
int multi_cg(Matrix& M, const MultiParams& P, int k, const T* b, T* x, Workspace& W, MultiResult* out,
             hipStream_t s);

struct MultiRefineResult {
  int outer = 0;
  std::vector<int> inner_iters;  // per right-hand side, summed over corrections
  std::vector<double> resid;     // final FP64 relative residual per right-hand side
  std::vector<int> status;       // 0 converged, 1 not converged, 2 zero right-hand side
};

// FP32-inner / FP64-refined multi-RHS CG: FP64 residual B - A X and update,
// FP32 multi-RHS CG corrections to a relative inner_tol.
int multi_refine(Matrix& M, const MultiParams& P, double inner_tol, int max_outer, int k, const double* b, double* x,
                 Workspace& Wi, Workspace& Wo, MultiRefineResult* out, hipStream_t s);

}  // namespace mh
