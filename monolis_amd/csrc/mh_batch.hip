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
// slice, so a workgroup walks only its own slices.  Two persistent kernels:
//  * k_batch_cg: each workgroup takes systems wg, wg + G, ... (largest first) and
//    runs the whole CG loop with workgroup barriers only;
//  * k_batch_sched: the largest systems are split over teams of up to 32
//    workgroups (contiguous slice ranges, team barriers in global memory with a
//    bounded wait), so that one system no longer sets the critical path.
// The launch (init kernel + solver kernel) is captured once in a hipGraph and
// replayed.
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
constexpr int kTeamMax = 32;        // workgroups per team at most
constexpr int kSpinMax = 1 << 24;   // polls of one team wait at most (a bound that needs no clock)

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
  int* status;   // 0 converged, 1 not converged, 2 zero rhs, 3 a team wait gave up
};

// One SELL row times a vector whose entry c is gather(c).
template <int U, class G>
__device__ __forceinline__ double sell_row_g(const BatchView& V, int s, int lane, int len, G&& gather) {
  SellView<double> M;
  M.nrows = 0x7fffffff;
  M.nslices = 0;
  M.slice_ptr = V.slice_ptr;
  M.rowptr = nullptr;
  M.col = V.col;
  M.val = V.val;
  M.col16 = nullptr;
  double acc[1] = {0.0};
  auto xg = [&](int c, double(&xv)[1]) { xv[0] = gather(c); };
  sell_row<double, double, 1, U, true>(M, s, lane, len, xg, acc);
  return acc[0];
}

template <int U>
__device__ __forceinline__ double sell_row_scalar(const BatchView& V, int s, int lane, int len, const double* x) {
  return sell_row_g<U>(V, s, lane, len, [&](int c) { return x[c]; });
}


// Vector loads and stores of the elementwise passes: plain, or non-temporal
// when SV (a batch whose vectors exceed what L2 / MALL keep between passes).
template <bool SV>
__device__ __forceinline__ double ldv(const double* p) {
  if constexpr (SV) return __builtin_nontemporal_load(p);
  else return *p;
}
template <bool SV>
__device__ __forceinline__ void stv(double* p, double v) {
  if constexpr (SV) __builtin_nontemporal_store(v, p);
  else *p = v;
}

// One system solved by one workgroup (NT threads, U slots per step).
template <int NT, int U, bool SV>
__device__ __forceinline__ void cg_solo(const BatchView& V, int sys, const double* __restrict__ b,
                                        double* __restrict__ x, double* __restrict__ r, double* __restrict__ z,
                                        double* __restrict__ p, double* __restrict__ q, const BatchOut& out,
                                        int max_iter, double tol, int init_x, int precond, double* red) {
  constexpr int NW = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  {
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
      return;
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
      for (int l = threadIdx.x; l < nrow; l += NT)
        stv<false>(p + row0 + l, fma(beta, ldv<SV>(p + row0 + l), ldv<SV>(z + row0 + l)));
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
          stv<SV>(x + g, fma(alpha, ldv<SV>(p + g), ldv<SV>(x + g)));
          const double rv = fma(-alpha, ldv<SV>(q + g), ldv<SV>(r + g));
          stv<SV>(r + g, rv);
          const double zv = precond ? ldv<SV>(V.dinv + g) * rv : rv;
          stv<SV>(z + g, zv);
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

// The persistent solver, one workgroup per system.
template <int NT, int U, bool SV>
__global__ void __launch_bounds__(NT) k_batch_cg(BatchView V, const double* __restrict__ b, double* __restrict__ x,
                                                 double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ p, double* __restrict__ q, BatchOut out,
                                                 int max_iter, double tol, int init_x, int precond) {
  __shared__ double red[NT / kWave * 2];
  for (int k = blockIdx.x; k < V.nsys; k += gridDim.x)
    cg_solo<NT, U, SV>(V, V.order[k], b, x, r, z, p, q, out, max_iter, tol, init_x, precond, red);
}

// Team schedule.  Workgroups take slots in order from one counter; the slots of
// a team are consecutive, so at most one team is incomplete at any time and a
// member waiting for a teammate that has not started waits only for running
// workgroups to finish their own complete teams -- nothing assumes that the
// whole grid is resident.  Every wait is bounded (wall clock and poll count):
// one that gives up raises the abort flag, after which every team wait returns
// at once and the systems concerned report status 3.
struct TeamView {
  int nslot;
  const int* slot;     // 4 ints per slot: system, rank in its team, team size, team index
  int* ctl;            // [0] next slot, [1] abort flag, [2 + t] arrivals at team t's barriers
  double* part;        // per team: 2 alternating sets x kTeamMax members x 2 partial sums
  long long limit;     // wall-clock ticks one wait may last
  int* sticky;         // set with the abort flag, cleared only by the host (survives later launches)
};

// Team barrier: arrival `target / T` of a team of T workgroups.  After the
// workgroup barrier one thread releases the workgroup's stores at device scope
// (they precede it through the barrier), counts the arrival, polls the counter
// and acquires, then the workgroup barrier again -- one cache write-back and one
// invalidation per member and barrier, so loads after it see what the other
// members stored before it, whichever XCD they run on.
__device__ __forceinline__ bool team_wait(int* cnt, int target, int* abort_flag, int* sticky, long long limit,
                                          int* flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int ok = 1;
    const long long t0 = (long long)wall_clock64();
    for (int k = 0;; ++k) {
      if (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target) break;
      if (__hip_atomic_load(abort_flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || k >= kSpinMax ||
          (long long)wall_clock64() - t0 > limit) {
        __hip_atomic_store(abort_flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_store(sticky, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    *flag = ok;
  }
  __syncthreads();
  return *flag != 0;
}

// Member `rank` of a team of T workgroups solving one system: it owns slices
// [s0 + ns*rank/T, s0 + ns*(rank+1)/T) and their rows; the SpMV reads the
// other members' p (and x) after a team barrier; each dot product is a
// workgroup sum per member, then the T member sums added in rank order (the
// same order in every member, so all members take the same branches).
template <int NT, int U, bool SV>
__device__ __forceinline__ void cg_team(const BatchView& V, const TeamView& S, int sys, int rank, int T, int team,
                                        const double* __restrict__ b, double* __restrict__ x,
                                        double* __restrict__ r, double* __restrict__ z, double* __restrict__ p,
                                        double* __restrict__ p2, double* __restrict__ q, const BatchOut& out,
                                        int max_iter, double tol, int init_x, int precond, double* red, int* flag) {
  constexpr int NW = NT / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int s0 = V.sys_slice[sys], ns = V.sys_slice[sys + 1] - s0;
  const int a0 = s0 + (int)((long long)ns * rank / T), a1 = s0 + (int)((long long)ns * (rank + 1) / T);
  const int row0 = V.sys_row[sys], nrow = V.sys_row[sys + 1] - row0;
  const int m0 = (a0 - s0) * kWave, m1 = min((a1 - s0) * kWave, nrow);  // this member's local rows
  int* cnt = S.ctl + 2 + team;
  double* part = S.part + (size_t)team * (2 * kTeamMax * 2);
  int nbar = 0, nred = 0;
  auto wait = [&]() -> bool {
    ++nbar;
    return team_wait(cnt, nbar * T, S.ctl + 1, S.sticky, S.limit, flag);
  };
  // team-wide sums; consecutive reductions alternate between two partial sets,
  // and a set is rewritten only after the next reduction's barrier, which every
  // member passes after reading it
  auto tsum = [&](double(&v)[2]) -> bool {
    wg_sum<NT, 2>(v, red);
    double* ps = part + (nred & 1) * (kTeamMax * 2);
    ++nred;
    if (threadIdx.x == 0) {
      ps[2 * rank] = v[0];
      ps[2 * rank + 1] = v[1];
    }
    if (!wait()) return false;
    double t0 = 0.0, t1 = 0.0;
    for (int m = 0; m < T; ++m) {
      t0 += ps[2 * m];
      t1 += ps[2 * m + 1];
    }
    v[0] = t0;
    v[1] = t1;
    return true;
  };
  auto fail = [&]() {
    if (threadIdx.x == 0) __hip_atomic_fetch_max(out.status + sys, 3, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  auto finish = [&](int iters, double resid, int st) {
    if (rank == 0 && threadIdx.x == 0) {
      out.iters[sys] = iters;
      out.resid[sys] = resid;
    }
    if (threadIdx.x == 0) __hip_atomic_fetch_max(out.status + sys, st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  auto for_rows = [&](auto&& f) {
    for (int s = a0 + w; s < a1; s += NW) {
      const int l = (s - s0) * kWave + lane;
      f(s, l, l < nrow);
    }
  };
  if (init_x) {
    for (int l = m0 + threadIdx.x; l < m1; l += NT) x[row0 + l] = 0.0;
    if (!wait()) return fail();
  }
  // r = b - A x ; B2 = (r, r)
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
  if (!tsum(acc2)) return fail();
  const double B2 = acc2[0];
  if (B2 == 0.0) {  // mat_converge.f90:40-48
    for (int l = m0 + threadIdx.x; l < m1; l += NT) x[row0 + l] = 0.0;
    return finish(0, 0.0, 2);
  }
  // z = M^-1 r ; rho = (r, z) ; p = 0
  double rho = 0.0;
  {
    double a[2] = {0.0, 0.0};
    for (int l = m0 + threadIdx.x; l < m1; l += NT) {
      const int g = row0 + l;
      const double zv = precond ? V.dinv[g] * r[g] : r[g];
      z[g] = zv;
      p[g] = 0.0;
      a[0] += r[g] * zv;
    }
    if (!tsum(a)) return fail();
    rho = a[0];
  }
  double beta = 0.0, resid = 1.0;
  int it = 1, conv = 0;
  // p alternates between two buffers: iteration it writes pn and reads po
  // (zero at it = 1).  The other members' rows of the new p are not waited for:
  // the SpMV forms them from their z and previous p, complete since the last
  // barrier, exactly as their owners compute them.
  const int g0 = row0 + m0, g1 = row0 + m1;
  for (; it <= max_iter; ++it) {
    const double* po = (it & 1) ? p : p2;
    double* pn = (it & 1) ? p2 : p;
    // p = z + beta p   (CG.f90:80-84), this member's rows
    for (int l = m0 + threadIdx.x; l < m1; l += NT)
      stv<false>(pn + row0 + l, fma(beta, ldv<SV>(po + row0 + l), ldv<SV>(z + row0 + l)));
    __syncthreads();
    // q = A p ; (p, q)
    double a[2] = {0.0, 0.0};
    for_rows([&](int s, int l, bool valid) {
      const int len = valid ? V.rowlen[s * kWave + lane] : 0;
      const double qv = sell_row_g<U>(V, s, lane, len, [&](int c) {
        double v;
        if (c >= g0 && c < g1)
          v = pn[c];
        else
          v = fma(beta, po[c], z[c]);
        return v;
      });
      if (valid) {
        q[row0 + l] = qv;
        a[0] += pn[row0 + l] * qv;
      }
    });
    if (!tsum(a)) return fail();
    const double alpha = rho / a[0];
    const bool rr = (it % kIterRR) == 0;
    double c[2] = {0.0, 0.0};
    if (!rr) {
      for (int l = m0 + threadIdx.x; l < m1; l += NT) {
        const int g = row0 + l;
        stv<SV>(x + g, fma(alpha, ldv<SV>(pn + g), ldv<SV>(x + g)));
        const double rv = fma(-alpha, ldv<SV>(q + g), ldv<SV>(r + g));
        stv<SV>(r + g, rv);
        const double zv = precond ? ldv<SV>(V.dinv + g) * rv : rv;
        stv<SV>(z + g, zv);
        c[0] += rv * rv;
        c[1] += rv * zv;
      }
    } else {
      for (int l = m0 + threadIdx.x; l < m1; l += NT) {
        const int g = row0 + l;
        x[g] = fma(alpha, pn[g], x[g]);
      }
      if (!wait()) return fail();
      for_rows([&](int s, int l, bool valid) {
        const int len = valid ? V.rowlen[s * kWave + lane] : 0;
        const double ax = sell_row_scalar<U>(V, s, lane, len, x);
        if (valid) r[row0 + l] = b[row0 + l] - ax;
      });
      __syncthreads();
      for (int l = m0 + threadIdx.x; l < m1; l += NT) {
        const int g = row0 + l;
        const double rv = r[g];
        const double zv = precond ? V.dinv[g] * rv : rv;
        z[g] = zv;
        c[0] += rv * rv;
        c[1] += rv * zv;
      }
    }
    if (!tsum(c)) return fail();
    resid = sqrt(c[0] / B2);
    if (resid < tol) {
      conv = 1;
      break;
    }
    beta = c[1] / rho;
    rho = c[1];
  }
  finish(conv ? it : max_iter, resid, conv ? 0 : 1);
}

template <int NT, int U, bool SV>
__global__ void __launch_bounds__(NT) k_batch_sched(BatchView V, TeamView S, const double* __restrict__ b,
                                                    double* __restrict__ x, double* __restrict__ r,
                                                    double* __restrict__ z, double* __restrict__ p,
                                                    double* __restrict__ p2, double* __restrict__ q, BatchOut out,
                                                    int max_iter, double tol, int init_x, int precond) {
  __shared__ double red[NT / kWave * 2];
  __shared__ int sh[2];
  for (;;) {
    if (threadIdx.x == 0) sh[0] = __hip_atomic_fetch_add(S.ctl, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int k = sh[0];
    if (k >= S.nslot) break;
    const int sys = S.slot[4 * k], rank = S.slot[4 * k + 1], T = S.slot[4 * k + 2], team = S.slot[4 * k + 3];
    if (T == 1)
      cg_solo<NT, U, SV>(V, sys, b, x, r, z, p, q, out, max_iter, tol, init_x, precond, red);
    else
      cg_team<NT, U, SV>(V, S, sys, rank, T, team, b, x, r, z, p, p2, q, out, max_iter, tol, init_x, precond, red,
                     sh + 1);
    __syncthreads();
  }
}

// Per-launch reset (a kernel, so that the captured graph holds no memset
// nodes): results to "not run", the slot counter, abort flag and team counters
// to zero.
__global__ void k_batch_reset(int nsys, int* iters, double* resid, int* status, int* ctl, int nctl) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nsys) {
    iters[i] = -1;
    resid[i] = 0.0;
    status[i] = -1;
  }
  if (i < nctl) ctl[i] = 0;
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

// Launch configurations: threads per workgroup, slots per SpMV step, team split
// (0: k_batch_cg; s > 0: k_batch_sched with teams for systems above s/4 of a
// workgroup's balanced share of the estimated work), non-temporal vector
// streams (measured: 10 % faster per iteration when 256 40^3 systems fill the
// GPU, slower when the active vectors fit in L2 / MALL).
#define MH_BATCH_CASE(c, nt, u, ts, sv, ...) \
  case c: {                                  \
    constexpr int NT = nt;                   \
    constexpr int UU = u;                    \
    constexpr int TS = ts;                   \
    constexpr bool SV = sv;                  \
    __VA_ARGS__;                             \
  } break;
#define MH_BATCH_CFG(cfg, ...)                       \
  switch (cfg) {                                     \
    MH_BATCH_CASE(1, 256, 2, 0, false, __VA_ARGS__)  \
    MH_BATCH_CASE(2, 512, 2, 0, false, __VA_ARGS__)  \
    MH_BATCH_CASE(3, 1024, 2, 0, false, __VA_ARGS__) \
    MH_BATCH_CASE(4, 512, 4, 0, false, __VA_ARGS__)  \
    MH_BATCH_CASE(5, 1024, 4, 0, false, __VA_ARGS__) \
    MH_BATCH_CASE(6, 512, 1, 0, false, __VA_ARGS__)  \
    MH_BATCH_CASE(7, 1024, 4, 4, false, __VA_ARGS__) \
    MH_BATCH_CASE(8, 1024, 4, 2, false, __VA_ARGS__) \
    MH_BATCH_CASE(9, 1024, 2, 2, false, __VA_ARGS__) \
    MH_BATCH_CASE(10, 512, 4, 2, false, __VA_ARGS__) \
    MH_BATCH_CASE(11, 1024, 2, 4, false, __VA_ARGS__) \
    MH_BATCH_CASE(12, 1024, 2, 3, false, __VA_ARGS__) \
    MH_BATCH_CASE(13, 1024, 2, 0, true, __VA_ARGS__) \
    MH_BATCH_CASE(14, 1024, 2, 4, true, __VA_ARGS__) \
    default: {                                       \
      constexpr int NT = 256;                        \
      constexpr int UU = 4;                          \
      constexpr int TS = 0;                          \
      constexpr bool SV = false;                     \
      __VA_ARGS__;                                   \
    } break;                                         \
  }

// Team schedule of a configuration (NT threads per workgroup, `resident`
// workgroups at once, split factor ts): the estimated work of a system is
// (entries + 6 rows) x rows^(1/3) (the SpMV and vector bytes times an iteration
// count that grows like a 3-D mesh's edge); systems above ts/4 of the balanced
// share get ceil(work / that) members, at most 32, at most the system's share of
// the resident workgroups and at most one slice per wave and member; tasks are ordered by work per member, largest first (teams and
// single workgroups interleaved: putting every team first left the largest
// single-workgroup systems for the end).
int build_sched(const Batch& B, int NT, int resident, int ts, BatchSched& S, hipStream_t s) {
  const int nsys = B.nsys;
  std::vector<double> w((size_t)nsys);
  double W = 0.0;
  for (int k = 0; k < nsys; ++k) {
    const double rows = (double)B.h_rows[k];
    w[k] = ((double)B.h_nnz[k] + 6.0 * rows) * std::cbrt(std::max(rows, 1.0));
    W += w[k];
  }
  const double share = W / std::max(resident, 1) * ts / 4.0;
  struct Task {
    int sys, T;
    double d;
  };
  std::vector<Task> tasks((size_t)nsys);
  for (int k = 0; k < nsys; ++k) {
    // at most its proportional part of the resident workgroups (rounded up):
    // teams that together ask for more than are resident would wait for one another
    const int tfair = (int)std::ceil((double)resident * w[k] / W);
    const int tmax = std::min(std::min(kTeamMax, tfair), B.h_slices[k] / (NT / kWave));
    int T = 1;
    if (tmax > 1 && share > 0.0 && w[k] > share) T = std::min(tmax, (int)std::ceil(w[k] / share));
    tasks[k] = Task{k, T, w[k] / T};
  }
  std::stable_sort(tasks.begin(), tasks.end(), [](const Task& a, const Task& b) { return a.d > b.d; });
  std::vector<int> slot;
  int nteam = 0;
  for (const Task& t : tasks) {
    const int team = t.T > 1 ? nteam++ : -1;
    for (int m = 0; m < t.T; ++m) {
      slot.push_back(t.sys);
      slot.push_back(m);
      slot.push_back(t.T);
      slot.push_back(team);
    }
  }
  S.nslot = (int)(slot.size() / 4);
  S.nteam = nteam;
  S.grid = std::max(1, std::min(S.nslot, resident));
  MH_TRY(S.slot.alloc(slot.size()));
  MH_TRY(S.ctl.alloc((size_t)2 + nteam));
  MH_TRY(S.part.alloc((size_t)std::max(nteam, 1) * 2 * kTeamMax * 2));
  MH_TRY(S.sticky.alloc(1));
  MH_TRY(upload(S.slot.p, slot.data(), sizeof(int) * slot.size(), s));
  MH_HIP(hipMemsetAsync(S.sticky.p, 0, sizeof(int), s));
  MH_HIP(hipStreamSynchronize(s));  // `slot` is freed on return
  int dev = 0, khz = 0;
  MH_HIP(hipGetDevice(&dev));
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
  S.ticks_per_s = (double)khz * 1000.0;
  return 0;
}

}  // namespace

int batch_num_configs() { return 15; }

void batch_config_info(int cfg, int* threads, int* unroll) {
  MH_BATCH_CFG(cfg, *threads = NT; *unroll = UU);
}

int batch_config_split(int cfg) {
  int ts = 0;
  MH_BATCH_CFG(cfg, ts = TS);
  return ts;
}

// The team schedule of configuration cfg, built on first use (outside any
// stream capture: it uploads synchronously).  Its control words are zeroed by
// the reset kernel of every launch.
static int ensure_sched(Batch& B, int cfg, BatchSched** out, hipStream_t s) {
  *out = nullptr;
  if ((int)B.sched.size() < batch_num_configs()) B.sched.resize(batch_num_configs());
  if (!B.sched[cfg]) {
    std::unique_ptr<BatchSched> S(new BatchSched());
    int rc = 0;
    MH_BATCH_CFG(cfg, {
      if constexpr (TS > 0) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_batch_sched<NT, UU, SV>, NT, 0) != hipSuccess ||
            per_cu < 1)
          per_cu = 1;
        rc = build_sched(B, NT, per_cu * cu_count_batch(), TS, *S, s);
      }
    });
    if (rc) return rc;
    B.sched[cfg] = std::move(S);
  }
  *out = B.sched[cfg].get();
  return 0;
}

int batch_schedule_info(Batch& B, int cfg, int* split, int* nslot, int* nteam, hipStream_t s) {
  if (cfg < 0 || cfg >= batch_num_configs()) {
    set_error("batch config %d out of range", cfg);
    return -2;
  }
  *split = batch_config_split(cfg);
  *nslot = B.nsys;
  *nteam = 0;
  if (*split == 0) return 0;
  BatchSched* S = nullptr;
  MH_TRY(ensure_sched(B, cfg, &S, s));
  *nslot = S->nslot;
  *nteam = S->nteam;
  return 0;
}

int batch_aborted(Batch& B, int cfg, int* aborted, hipStream_t s) {
  *aborted = 0;
  if (cfg < 0 || cfg >= (int)B.sched.size() || !B.sched[cfg] || batch_config_split(cfg) == 0) return 0;
  // the sticky flag: set by any launch since the last check, then cleared
  int flag = 0;
  BatchSched& S = *B.sched[cfg];
  MH_HIP(hipMemcpyAsync(&flag, S.sticky.p, sizeof(int), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  if (flag) {
    MH_HIP(hipMemsetAsync(S.sticky.p, 0, sizeof(int), s));
    MH_HIP(hipStreamSynchronize(s));
  }
  *aborted = flag != 0;
  return 0;
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
  B->h_rows.resize((size_t)nsys);
  B->h_nnz.resize((size_t)nsys);
  B->h_slices.resize((size_t)nsys);
  for (int k = 0; k < nsys; ++k) {
    const int r0 = sys_row_h[k], r1 = sys_row_h[k + 1];
    B->h_rows[k] = r1 - r0;
    B->h_nnz[k] = hidx[r1] - hidx[r0];
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
    B->h_slices[k] = ns;
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
  MH_TRY(B->work.alloc((size_t)n * 5));  // r, z, p, q and the second p buffer of the team solver
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
  double* p2 = q + B.n;
  if (cfg < 0 || cfg >= batch_num_configs()) {
    set_error("batch config %d out of range", cfg);
    return -2;
  }
  BatchSched* S = nullptr;
  if (batch_config_split(cfg) > 0) MH_TRY(ensure_sched(B, cfg, &S, s));
  // k_batch_cg: one workgroup per system while they fit two per CU, else
  // grid-stride over systems
  int grid = std::min(B.nsys, cu_count_batch() * 2);
  if (grid < 1) grid = 1;
  auto enqueue = [&]() -> int {
    const int nctl = S ? 2 + S->nteam : 0;
    k_batch_reset<<<(std::max(B.nsys, nctl) + 255) / 256, 256, 0, s>>>(B.nsys, B.iters.p, B.resid.p, B.status.p,
                                                                         S ? S->ctl.p : nullptr, nctl);
    MH_BATCH_CFG(cfg, {
      if constexpr (TS > 0) {
        const TeamView TV{S->nslot, S->slot.p, S->ctl.p, S->part.p, (long long)(S->ticks_per_s * B.wait_s),
                          S->sticky.p};
        k_batch_sched<NT, UU, SV><<<S->grid, NT, 0, s>>>(V, TV, b, x, r, z, p, p2, q, O, P.max_iter, P.tol,
                                                     P.init_x, P.precond);
      } else {
        k_batch_cg<NT, UU, SV><<<grid, NT, 0, s>>>(V, b, x, r, z, p, q, O, P.max_iter, P.tol, P.init_x, P.precond);
      }
    });
    MH_HIP(hipGetLastError());
    return 0;
  };
  if (!use_graph) return enqueue();
  const bool same = B.graph_exec && B.g_b == b && B.g_x == x && B.g_cfg == cfg && B.g_max_iter == P.max_iter &&
                    B.g_tol == P.tol && B.g_init_x == P.init_x && B.g_precond == P.precond && B.g_wait == B.wait_s;
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
    B.g_wait = B.wait_s;
  }
  MH_HIP(hipGraphLaunch(B.graph_exec, s));
  return 0;
}

}  // namespace mh
