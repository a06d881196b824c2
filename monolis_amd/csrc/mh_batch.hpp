// Batched solver of many independent systems (block-diagonal scalar CSR).
#pragma once

#include <memory>
#include <vector>

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

// Team schedule of one launch configuration: an ordered list of slots (system,
// rank in its team, team size, team index); the largest systems are split over
// teams of workgroups.
struct BatchSched {
  int nslot = 0, nteam = 0, grid = 0;
  double ticks_per_s = 0;   // wall-clock counter rate
  DevArray<int> slot;       // 4 ints per slot
  DevArray<int> ctl;        // [0] next slot, [1] abort flag, [2 ..] team arrival counters
  DevArray<double> part;    // per team: 2 x 32 x 2 partial sums
  DevArray<int> sticky;     // a team wait gave up in some launch since the host last looked
};

struct Batch {
  int nsys = 0, n = 0, nz = 0, nslices = 0;
  size_t nslots = 0;
  DevArray<int> sys_slice, sys_row, order, slice_ptr, col, rowlen;
  DevArray<double> val, dinv, work;
  DevArray<int> iters, status;
  DevArray<double> resid;
  std::vector<int> h_rows, h_nnz, h_slices;  // per system (schedule building)
  std::vector<std::unique_ptr<BatchSched>> sched;  // per configuration, built on first use
  int cfg = 0;  // tuned launch configuration
  double wait_s = 5.0;  // longest team wait before a launch gives up
  // captured launch (replayed while arguments are unchanged)
  hipGraphExec_t graph_exec = nullptr;
  const double* g_b = nullptr;
  double* g_x = nullptr;
  int g_cfg = -1, g_max_iter = -1, g_init_x = -1, g_precond = -1;
  double g_tol = -1.0, g_wait = -1.0;
  ~Batch();
};

int batch_num_configs();
void batch_config_info(int cfg, int* threads, int* unroll);
// Team split of a configuration: 0 = one workgroup per system, else systems whose
// estimated work exceeds split/4 of a workgroup's balanced share get teams.
int batch_config_split(int cfg);
int batch_create(Batch** out, int nsys, const int* sys_row_host, int n, int nz, const int* index, const int* item,
                 const double* A, hipStream_t s);
// Enqueues one batched solve (graph replay when use_graph); results in B.iters/resid/status.
int batch_solve(Batch& B, const BatchParams& P, const double* b, double* x, int cfg, int use_graph, hipStream_t s);
// After launches of configuration cfg have completed: 1 when a team wait gave up
// in any of them since the last call (the affected systems carry status 3), else 0.
int batch_aborted(Batch& B, int cfg, int* aborted, hipStream_t s);
// Team split of cfg and the slot / team counts of its schedule for B (built on
// first use; one slot per system and no teams for split 0).
int batch_schedule_info(Batch& B, int cfg, int* split, int* nslot, int* nteam, hipStream_t s);

}  // namespace mh
