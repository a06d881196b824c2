// Batched solves of many independent sparse systems in one persistent launch.
//
// A batch is a block-diagonal scalar CSR matrix (system s owns rows
// sys_row[s] .. sys_row[s+1]-1 and only couples to its own rows), as produced by
// assembling every system into one monolis matrix.  Each system is solved by
// monolis's CG + diagonal preconditioner (src/iterative/CG.f90:18-119) with its
// own scalars, convergence test and iteration count, exactly as if
// monolis_solve_R were called once per system.
//
// Layout: a system-aligned SELL-64 -- every system starts on a fresh 64-row
// slice, so a workgroup walks only its own slices.  One persistent kernel: each
// workgroup takes systems wg, wg + G, ... (largest first) and runs the whole CG
// loop with workgroup barriers only; there is no cross-workgroup waiting, so
// every workgroup exits after its last system.  The launch (init kernel +
// solver kernel) is captured once in a hipGraph and replayed.
#include <algorithm>
#include <memory>
#include <cmath>
#include <numeric>
#include <vector>

#include "mh_batch.hpp"
#include "mh_kernels.hpp"

namespace mh {

namespace {

constexpr int kIterRR = 200;

// workgroup sum over NT threads (fixed order), result returned to every thread
template <int NT, int NV>
__device__ __forceinline__ void wg_sum(double (&v)[NV], double* red /* NT/64*NV */) {
  constexpr int NW = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[w * NV + k] = v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = red[k];
    for (int i = 1; i < NW; ++i) s += red[i * NV + k];
    v[k] = s;
  }
  __syncthreads();
}

struct BatchView {
  int nsys;
  const int* sys_slice;     // [nsys+1] first slice of each system
  const int* sys_row;       // [nsys+1] first global row of each system
  const int* order;         // systems by decreasing size
  const int* slice_ptr;     // SELL slot offsets per slice
  const int* col;           // global 0-based columns
  const double* val;
  const int* rowlen;        // per SELL row (padding rows: 0)
  const double* dinv;       // diagonal preconditioner per global row (1/a_ii, zero pivot -> 1)
};

// Per-system results written by the solver kernel.
struct BatchOut {
  int* iters;
  double* resid;
  int* status;   // 0 converged, 1 not converged, 2 zero rhs
};

template <int U>
__device__ __forceinline__ double sell_row_scalar(const BatchView& V, int s, int lane, int len, const double* x) {
  SellView<double> M;
  M.nrows = 0x7fffffff;
  M.nslices = 0;
  M.slice_ptr = V.slice_ptr;
  M.rowptr = nullptr;
  M.col = V.col;
  M.val = V.val;
  M.col16 = nullptr;
  double acc[1] = {0.0};
  auto xg = [&](int c, double(&xv)[1]) { xv[0] = x[c]; };
  sell_row<double, double, 1, U, true>(M, s, lane, len, xg, acc);
  return acc[0];
}

// The persistent solver.  NT threads per workgroup; U slots per step.
template <int NT, int U>
__global__ void __launch_bounds__(NT) k_batch_cg(BatchView V, const double* __restrict__ b, double* __restrict__ x,
                                                 double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ p, double* __restrict__ q, BatchOut out,
                                                 int max_iter, double tol, int init_x, int precond) {
  __shared__ double red[NT / kWave * 2];
  constexpr int NW = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  for (int k = blockIdx.x; k < V.nsys; k += gridDim.x) {
    const int sys = V.order[k];
    const int s0 = V.sys_slice[sys], s1 = V.sys_slice[sys + 1];
    const int row0 = V.sys_row[sys], nrow = V.sys_row[sys + 1] - row0;
    // rows of this system: local row l = (slice - s0) * 64 + lane; global row0 + l
    auto for_rows = [&](auto&& f) {
      for (int s = s0 + w; s < s1; s += NW) {
        const int l = (s - s0) * kWave + lane;
        f(s, l, l < nrow);
      }
    };
    // r = b - A x ; B2 = (r, r)
    if (init_x) {
      for (int l = threadIdx.x; l < nrow; l += NT) x[row0 + l] = 0.0;
      __syncthreads();
    }
    double acc2[2] = {0.0, 0.0};
    for_rows([&](int s, int l, bool valid) {
      const int len = valid ? V.rowlen[s * kWave + lane] : 0;
      const double ax = sell_row_scalar<U>(V, s, lane, len, x);
      if (valid) {
        const double rv = b[row0 + l] - ax;
        r[row0 + l] = rv;
        acc2[0] += rv * rv;
      }
    });
    wg_sum<NT, 2>(acc2, red);
    const double B2 = acc2[0];
    if (B2 == 0.0) {  // mat_converge.f90:40-48
      for (int l = threadIdx.x; l < nrow; l += NT) x[row0 + l] = 0.0;
      if (threadIdx.x == 0) {
        out.iters[sys] = 0;
        out.resid[sys] = 0.0;
        out.status[sys] = 2;
      }
      __syncthreads();
      continue;
    }
    // z = M^-1 r ; rho = (r, z) ; p = 0
    double rho = 0.0;
    {
      double a[2] = {0.0, 0.0};
      for (int l = threadIdx.x; l < nrow; l += NT) {
        const int g = row0 + l;
        const double zv = precond ? V.dinv[g] * r[g] : r[g];
        z[g] = zv;
        p[g] = 0.0;
        a[0] += r[g] * zv;
      }
      wg_sum<NT, 2>(a, red);
      rho = a[0];
    }
    double beta = 0.0, resid = 1.0;
    int it = 1, conv = 0;
    for (; it <= max_iter; ++it) {
      // p = z + beta p   (CG.f90:80-84)
      for (int l = threadIdx.x; l < nrow; l += NT) p[row0 + l] = fma(beta, p[row0 + l], z[row0 + l]);
      __syncthreads();
      // q = A p ; (p, q)
      double a[2] = {0.0, 0.0};
      for_rows([&](int s, int l, bool valid) {
        const int len = valid ? V.rowlen[s * kWave + lane] : 0;
        const double qv = sell_row_scalar<U>(V, s, lane, len, p);
        if (valid) {
          q[row0 + l] = qv;
          a[0] += p[row0 + l] * qv;
        }
      });
      wg_sum<NT, 2>(a, red);
      const double alpha = rho / a[0];
      // x += alpha p ; r -= alpha q (or the true residual every 200 iterations);
      // ||r||, z = M^-1 r, (r, z).  Outside the replacement iterations both are
      // one pass over the rows (the same operations per row, the same per-thread
      // order of the sums).
      const bool rr = (it % kIterRR) == 0;
      double c[2] = {0.0, 0.0};
      if (!rr) {
        for (int l = threadIdx.x; l < nrow; l += NT) {
          const int g = row0 + l;
          x[g] = fma(alpha, p[g], x[g]);
          const double rv = fma(-alpha, q[g], r[g]);
          r[g] = rv;
          const double zv = precond ? V.dinv[g] * rv : rv;
          z[g] = zv;
          c[0] += rv * rv;
          c[1] += rv * zv;
        }
      } else {
        for (int l = threadIdx.x; l < nrow; l += NT) {
          const int g = row0 + l;
          x[g] = fma(alpha, p[g], x[g]);
        }
        __syncthreads();
        for_rows([&](int s, int l, bool valid) {
          const int len = valid ? V.rowlen[s * kWave + lane] : 0;
          const double ax = sell_row_scalar<U>(V, s, lane, len, x);
          if (valid) r[row0 + l] = b[row0 + l] - ax;
        });
        __syncthreads();
        for (int l = threadIdx.x; l < nrow; l += NT) {
          const int g = row0 + l;
          const double rv = r[g];
          const double zv = precond ? V.dinv[g] * rv : rv;
          z[g] = zv;
          c[0] += rv * rv;
          c[1] += rv * zv;
        }
      }
      wg_sum<NT, 2>(c, red);
      resid = sqrt(c[0] / B2);
      if (resid < tol) {
        conv = 1;
        break;
      }
      beta = c[1] / rho;
      rho = c[1];
    }
    if (threadIdx.x == 0) {
      out.iters[sys] = conv ? it : max_iter;
      out.resid[sys] = resid;
      out.status[sys] = conv ? 0 : 1;
    }
    __syncthreads();
  }
}

__global__ void k_batch_reset(int nsys, int* iters, double* resid, int* status) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsys) {
    iters[i] = -1;
    resid[i] = 0.0;
    status[i] = -1;
  }
}

// system-aligned SELL build: one thread per SELL row
__global__ void k_batch_fill(int nsys_rows_total, const int* __restrict__ sell_sys, const int* __restrict__ sys_slice,
                             const int* __restrict__ sys_row, const int* __restrict__ rowptr,
                             const int* __restrict__ item, const double* __restrict__ A,
                             const int* __restrict__ slice_ptr, int* __restrict__ col, double* __restrict__ val,
                             int* __restrict__ rowlen, double* __restrict__ dinv) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsys_rows_total) return;
  const int s = t / kWave, lane = t % kWave;
  const int sys = sell_sys[s];
  const int l = (s - sys_slice[sys]) * kWave + lane;
  const int nrow = sys_row[sys + 1] - sys_row[sys];
  const int g = sys_row[sys] + l;
  const int k0 = slice_ptr[s], w = slice_ptr[s + 1] - k0;
  int j0 = 0, len = 0;
  if (l < nrow) {
    j0 = rowptr[g];
    len = rowptr[g + 1] - j0;
  }
  rowlen[t] = len;
  double d = 0.0;
  for (int k = 0; k < w; ++k) {
    const size_t slot = (size_t)(k0 + k);
    if (k < len) {
      const int c = item[j0 + k] - 1;
      col[slot * kWave + lane] = c;
      val[slot * kWave + lane] = A[j0 + k];
      if (c == g) d = A[j0 + k];
    } else {
      col[slot * kWave + lane] = l < nrow ? g : sys_row[sys];
      val[slot * kWave + lane] = 0.0;
    }
  }
  if (l < nrow) {
    // diag_nn setup with b = 1 (diag_nn.f90:45-80): zero pivot -> 1, store 1/d
    if (d == 0.0) d = 1.0;
    dinv[g] = 1.0 / d;
  }
}

#define MH_BATCH_CFG(cfg, ...)                                                  \
  switch (cfg) {                                                                \
    case 1: { constexpr int NT = 256; constexpr int UU = 2; __VA_ARGS__; } break;  \
    case 2: { constexpr int NT = 512; constexpr int UU = 2; __VA_ARGS__; } break;  \
    case 3: { constexpr int NT = 1024; constexpr int UU = 2; __VA_ARGS__; } break; \
    case 4: { constexpr int NT = 512; constexpr int UU = 4; __VA_ARGS__; } break;  \
    case 5: { constexpr int NT = 1024; constexpr int UU = 4; __VA_ARGS__; } break; \
    case 6: { constexpr int NT = 512; constexpr int UU = 1; __VA_ARGS__; } break;  \
    default: { constexpr int NT = 256; constexpr int UU = 4; __VA_ARGS__; } break; \
  }

}  // namespace

int batch_num_configs() { return 7; }

void batch_config_info(int cfg, int* threads, int* unroll) {
  MH_BATCH_CFG(cfg, *threads = NT; *unroll = UU);
}

Batch::~Batch() {
  if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
}

int batch_create(Batch** out, int nsys, const int* sys_row_h, int n, int nz, const int* index, const int* item,
                 const double* A, hipStream_t s) {
  *out = nullptr;
  if (nsys < 1 || n < 1 || nz < 1 || sys_row_h[0] != 0 || sys_row_h[nsys] != n) {
    set_error("invalid batch: nsys=%d n=%d nz=%d", nsys, n, nz);
    return -2;
  }
  for (int k = 0; k < nsys; ++k)
    if (sys_row_h[k + 1] < sys_row_h[k]) {
      set_error("batch row offsets are not non-decreasing");
      return -2;
    }
  // host: validate structure (block-diagonal, in range) on the caller's arrays
  std::vector<int> hidx((size_t)n + 1), hitem((size_t)nz);
  MH_HIP(hipMemcpy(hidx.data(), index, sizeof(int) * ((size_t)n + 1), hipMemcpyDefault));
  MH_HIP(hipMemcpy(hitem.data(), item, sizeof(int) * (size_t)nz, hipMemcpyDefault));
  if (hidx[0] != 0 || hidx[n] != nz) {
    set_error("invalid batch CSR offsets");
    return -2;
  }
  std::unique_ptr<Batch> B(new Batch());
  B->nsys = nsys;
  B->n = n;
  B->nz = nz;
  std::vector<int> sys_slice((size_t)nsys + 1, 0), width, sell_sys;
  for (int k = 0; k < nsys; ++k) {
    const int r0 = sys_row_h[k], r1 = sys_row_h[k + 1];
    for (int i = r0; i < r1; ++i) {
      if (hidx[i + 1] < hidx[i]) {
        set_error("batch CSR offsets decrease at row %d", i);
        return -2;
      }
      for (int j = hidx[i]; j < hidx[i + 1]; ++j)
        if (hitem[j] - 1 < r0 || hitem[j] - 1 >= r1) {
          set_error("batch matrix couples system %d to a row outside it (row %d)", k, i);
          return -2;
        }
    }
    const int ns = (r1 - r0 + kWave - 1) / kWave;
    sys_slice[k + 1] = sys_slice[k] + ns;
    for (int q = 0; q < ns; ++q) {
      int wmax = 0;
      for (int l = q * kWave; l < std::min((q + 1) * kWave, r1 - r0); ++l)
        wmax = std::max(wmax, hidx[r0 + l + 1] - hidx[r0 + l]);
      width.push_back(wmax);
      sell_sys.push_back(k);
    }
  }
  const int nslices = sys_slice[nsys];
  std::vector<int> slice_ptr((size_t)nslices + 1, 0);
  for (int q = 0; q < nslices; ++q) slice_ptr[q + 1] = slice_ptr[q] + width[q];
  const size_t nslots = (size_t)slice_ptr[nslices];
  std::vector<int> order(nsys);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (sys_row_h[a + 1] - sys_row_h[a]) > (sys_row_h[b + 1] - sys_row_h[b]);
  });
  B->nslices = nslices;
  B->nslots = nslots;
  DevArray<int> d_idx, d_item, d_sell_sys;
  DevArray<double> d_A;
  MH_TRY(d_idx.alloc((size_t)n + 1));
  MH_TRY(d_item.alloc((size_t)nz));
  MH_TRY(d_A.alloc((size_t)nz));
  MH_TRY(d_sell_sys.alloc((size_t)nslices));
  MH_TRY(upload(d_idx.p, hidx.data(), sizeof(int) * ((size_t)n + 1), s));
  MH_TRY(upload(d_item.p, hitem.data(), sizeof(int) * (size_t)nz, s));
  MH_TRY(upload(d_A.p, A, sizeof(double) * (size_t)nz, s));
  MH_TRY(upload(d_sell_sys.p, sell_sys.data(), sizeof(int) * (size_t)nslices, s));
  MH_TRY(B->sys_slice.alloc((size_t)nsys + 1));
  MH_TRY(B->sys_row.alloc((size_t)nsys + 1));
  MH_TRY(B->order.alloc((size_t)nsys));
  MH_TRY(B->slice_ptr.alloc((size_t)nslices + 1));
  MH_TRY(B->col.alloc(nslots * kWave));
  MH_TRY(B->val.alloc(nslots * kWave));
  MH_TRY(B->rowlen.alloc((size_t)nslices * kWave));
  MH_TRY(B->dinv.alloc((size_t)n));
  MH_TRY(B->work.alloc((size_t)n * 4));
  MH_TRY(B->iters.alloc((size_t)nsys));
  MH_TRY(B->resid.alloc((size_t)nsys));
  MH_TRY(B->status.alloc((size_t)nsys));
  MH_TRY(upload(B->sys_slice.p, sys_slice.data(), sizeof(int) * ((size_t)nsys + 1), s));
  MH_TRY(upload(B->sys_row.p, sys_row_h, sizeof(int) * ((size_t)nsys + 1), s));
  MH_TRY(upload(B->order.p, order.data(), sizeof(int) * (size_t)nsys, s));
  MH_TRY(upload(B->slice_ptr.p, slice_ptr.data(), sizeof(int) * ((size_t)nslices + 1), s));
  const int total = nslices * kWave;
  k_batch_fill<<<(total + 255) / 256, 256, 0, s>>>(total, d_sell_sys.p, B->sys_slice.p, B->sys_row.p, d_idx.p, d_item.p,
                                                   d_A.p, B->slice_ptr.p, B->col.p, B->val.p, B->rowlen.p, B->dinv.p);
  MH_HIP(hipGetLastError());
  MH_HIP(hipStreamSynchronize(s));
  *out = B.release();
  return 0;
}

int batch_solve(Batch& B, const BatchParams& P, const double* b, double* x, int cfg, int use_graph, hipStream_t s) {
  BatchView V{B.nsys, B.sys_slice.p, B.sys_row.p, B.order.p, B.slice_ptr.p, B.col.p, B.val.p, B.rowlen.p, B.dinv.p};
  BatchOut O{B.iters.p, B.resid.p, B.status.p};
  double* r = B.work.p;
  double* z = r + B.n;
  double* p = z + B.n;
  double* q = p + B.n;
  int threads = 256, unroll = 2;
  batch_config_info(cfg, &threads, &unroll);
  // one workgroup per system while they fit two per CU, else grid-stride over systems
  int grid = std::min(B.nsys, cu_count_batch() * 2);
  if (grid < 1) grid = 1;
  auto enqueue = [&]() -> int {
    k_batch_reset<<<(B.nsys + 255) / 256, 256, 0, s>>>(B.nsys, B.iters.p, B.resid.p, B.status.p);
    MH_BATCH_CFG(cfg, k_batch_cg<NT, UU><<<grid, NT, 0, s>>>(V, b, x, r, z, p, q, O, P.max_iter, P.tol, P.init_x,
                                                             P.precond));
    MH_HIP(hipGetLastError());
    return 0;
  };
  if (!use_graph) return enqueue();
  const bool same = B.graph_exec && B.g_b == b && B.g_x == x && B.g_cfg == cfg && B.g_max_iter == P.max_iter &&
                    B.g_tol == P.tol && B.g_init_x == P.init_x && B.g_precond == P.precond;
  if (!same) {
    if (B.graph_exec) {
      (void)hipGraphExecDestroy(B.graph_exec);
      B.graph_exec = nullptr;
    }
    hipGraph_t g = nullptr;
    MH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue();
    hipError_t e = hipStreamEndCapture(s, &g);
    if (rc || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);  // a partial capture still owns its graph
      if (rc) return rc;
      set_error("graph capture failed: %s", hipGetErrorString(e));
      return -1;
    }
    e = hipGraphInstantiate(&B.graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
      B.graph_exec = nullptr;
      set_error("graph instantiate failed: %s", hipGetErrorString(e));
      return -1;
    }
    B.g_b = b;
    B.g_x = x;
    B.g_cfg = cfg;
    B.g_max_iter = P.max_iter;
    B.g_tol = P.tol;
    B.g_init_x = P.init_x;
    B.g_precond = P.precond;
  }
  MH_HIP(hipGraphLaunch(B.graph_exec, s));
  return 0;
}

}  // namespace mh
