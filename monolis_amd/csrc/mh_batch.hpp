// Batched solver of many independent systems (block-diagonal scalar CSR).
#pragma once

#include "mh_internal.hpp"

namespace mh {

int cu_count();
inline int cu_count_batch() { return cu_count(); }

struct BatchParams {
  int max_iter = 1000;
  double tol = 1e-8;
  int init_x = 1;
  int precond = 1;   // 0 NONE, 1 DIAG (1/a_ii)
};

struct Batch {
  int nsys = 0, n = 0, nz = 0, nslices = 0;
  size_t nslots = 0;
  DevArray<int> sys_slice, sys_row, order, slice_ptr, col, rowlen;
  DevArray<double> val, dinv, work;
  DevArray<int> iters, status;
  DevArray<double> resid;
  int cfg = 0;  // tuned launch configuration
  // captured launch (replayed while arguments are unchanged)
  hipGraphExec_t graph_exec = nullptr;
  const double* g_b = nullptr;
  double* g_x = nullptr;
  int g_cfg = -1, g_max_iter = -1, g_init_x = -1, g_precond = -1;
  double g_tol = -1.0;
  ~Batch();
};

int batch_num_configs();
void batch_config_info(int cfg, int* threads, int* unroll);
int batch_create(Batch** out, int nsys, const int* sys_row_host, int n, int nz, const int* index, const int* item,
                 const double* A, hipStream_t s);
// Enqueues one batched solve (graph replay when use_graph); results in B.iters/resid/status.
int batch_solve(Batch& B, const BatchParams& P, const double* b, double* x, int cfg, int use_graph, hipStream_t s);

}  // namespace mh
