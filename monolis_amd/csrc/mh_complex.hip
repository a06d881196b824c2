// Complex FP64 Krylov path: SpMV (monolis src/linalg/matvec.f90:192-227,
// matvec_*_C), the unconjugated dot (src/linalg/inner_product.f90:457-499),
// complex block-Jacobi (src/prec/diag/diag_nn.f90:90-159, :235-284) and COCG
// (src/iterative/COCG.f90:18-105) with the complex convergence test
// (src/linalg/mat_converge.f90:53-86, :145-196).
//
// Values are 16-byte complex numbers (zd) in the real path's SELL-64 layout:
// one slot of a wavefront reads 64 x 16 B = 1 KiB contiguous per block entry.
// COCG runs like the device CG: three fused kernels per iteration (p update;
// SpMV with (p, q) and alpha; x / r update with (r, r), the test, z = M^-1 r and
// (r, z), beta), complex scalars kept in a device ZState, `done` polled one
// chunk behind.
#include <algorithm>
#include <cstring>

#include "mh_complex.hpp"
#include "mh_launch.hpp"

namespace mh {

namespace {

constexpr int kIterRR = 200;  // COCG.f90: iter_RR = 200
enum : int { F_UPD = 1, F_CHECK = 2, F_FIRST = 4, F_XONLY = 8, F_EXTPREC = 16 };

// part[k] + i part[k+1] += a * b
template <int NV>
__device__ __forceinline__ void zacc(double (&part)[NV], int k, zd a, zd b) {
  part[k] += a.re * b.re - a.im * b.im;
  part[k + 1] += a.re * b.im + a.im * b.re;
}

// Principal square root (the Fortran intrinsic sqrt of a complex argument).
__device__ __forceinline__ zd zsqrt(zd q) {
  const double m = hypot(q.re, q.im);
  if (m == 0.0) return zd(0.0, q.im);
  if (q.re >= 0.0) {
    const double t = sqrt(0.5 * (m + q.re));
    return zd(t, q.im / (2.0 * t));
  }
  const double t = sqrt(0.5 * (m - q.re));
  return zd(fabs(q.im) / (2.0 * t), copysign(t, q.im));
}

// Block-Jacobi apply of one block row in the diag_nn_apply_C order; D is the
// structure-of-arrays factor (entry e of row i at D[e * nrows + i]).
template <int B>
__device__ __forceinline__ void zdiag_apply(const zd* __restrict__ D, int nrows, int i, zd (&x)[B]) {
  const size_t n = (size_t)nrows;
#pragma unroll
  for (int j = 1; j < B; ++j)
#pragma unroll
    for (int k = 0; k < j; ++k) x[j] = x[j] - D[(size_t)(B * j + k) * n + i] * x[k];
#pragma unroll
  for (int j = B - 1; j >= 0; --j) {
#pragma unroll
    for (int k = B - 1; k > j; --k) x[j] = x[j] - D[(size_t)(B * j + k) * n + i] * x[k];
    x[j] = D[(size_t)((B + 1) * j) * n + i] * x[j];
  }
}

// Solver scalars in device memory.
struct ZState {
  zd rho, alpha, beta, B2;
  double resid_abs;  // resid_abs, iter, done: the PollView the host polls
  int iter, done;
  double resid_re, tol;
  int max_iter, converged;
};
static_assert(offsetof(ZState, iter) == offsetof(ZState, resid_abs) + 8 &&
                  offsetof(ZState, done) == offsetof(ZState, resid_abs) + 12,
              "ZState::resid_abs, iter, done must form a PollView");

__global__ void k_zstate_init(ZState* st, double tol, int max_iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    ZState z;
    memset(&z, 0, sizeof(z));
    z.tol = tol;
    z.max_iter = max_iter;
    *st = z;
  }
}

// monolis_check_converge_C: resid = sqrt(R2 / B2), cur_resid = its real part,
// converged when |resid| < tol.
__device__ __forceinline__ bool zfinalize_check(ZState* st, double* hist, int it, zd R2) {
  const zd rs = zsqrt(zdiv(R2, st->B2));
  const double a = hypot(rs.re, rs.im);
  st->resid_abs = a;
  st->resid_re = rs.re;
  st->iter = it;
  if (hist) hist[it] = a;
  if (a < st->tol) {
    st->converged = 1;
    st->done = 1;
    return true;
  }
  if (it >= st->max_iter) {
    st->done = 1;
    return true;
  }
  return false;
}

// SELL launch for complex values: a slot of a B x B complex block holds 4*B*B
// VGPRs, so one slot per step (twice that for B = 1, through SellCfg) with the
// geometry of real-path variant 0: contiguous slice ranges at 8 waves per CU.
// Measured on the 100^3 BCSR3 pattern (26.5 M complex 3x3 blocks) it is the
// fastest of unroll 1..3, the pipelined loop and wpc 0..24 (664-689 us,
// 73-76 % of HBM peak; unroll 2 705-735 us; the pipelined loop 1.2-2.1 ms;
// loading the next slot's column ids during the current slot 698 vs 676 us).
constexpr int kZSellGeom = 0;
template <int B>
struct ZUnroll {
  static constexpr int U = 1;
};

template <int B, class Op>
int launch_zsell(const Matrix& M, const Op& op, double* partials, hipStream_t s) {
  int nwg = 0, spw = 0;
  if (M.nslices > 0) {
    sell_geometry(M, kZSellGeom, &nwg, &spw);
    k_sell<zd, zd, B, ZUnroll<B>::U, Op><<<padded_grid(nwg), kBlock, 0, s>>>(view<zd>(M), op, partials, nwg, spw);
    MH_HIP(hipGetLastError());
    ktimer_mark(kClsSpmv);
  }
  if constexpr (Op::NV > 0) {
    k_finalize<Op><<<1, kBlock, 0, s>>>(op, partials, nwg);
    MH_HIP(hipGetLastError());
    ktimer_mark(kClsDot);
  }
  return 0;
}

template <int B>
__device__ __forceinline__ void zgather(const zd* __restrict__ v, int c, zd (&xv)[B]) {
#pragma unroll
  for (int k = 0; k < B; ++k) xv[k] = v[(size_t)c * B + k];
}

// ---- y = A x ----------------------------------------------------------------
template <int B>
struct OpZSpmv {
  static constexpr int NV = 0;
  const zd* __restrict__ x;
  zd* __restrict__ y;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ void gather(int c, zd (&xv)[B]) const { zgather<B>(x, c, xv); }
  __device__ void epilogue(int, bool, const zd (&acc)[B], double (&)[1], const WaveRows<zd, B>& io) const {
    io.store(y, acc);
  }
  __device__ void finalize(const double (&)[1]) const {}
};

// ---- r = b - A x; (r, r) (matvec.f90:907-933) ----------------------------------
template <int B>
struct OpZResid {
  static constexpr int NV = 2;
  const zd* __restrict__ x;
  const zd* __restrict__ b;
  zd* __restrict__ r;
  double* out;          // (r, r) as two doubles, may be null
  const ZState* st;     // may be null
  __device__ bool skip() const { return st && st->done != 0; }
  __device__ void begin() {}
  __device__ void gather(int c, zd (&xv)[B]) const { zgather<B>(x, c, xv); }
  __device__ void epilogue(int, bool valid, const zd (&acc)[B], double (&part)[2], const WaveRows<zd, B>& io) const {
    zd v[B];
    io.load(b, v);
#pragma unroll
    for (int k = 0; k < B; ++k) {
      v[k] = v[k] - acc[k];
      if (valid) zacc(part, 0, v[k], v[k]);
    }
    io.store(r, v);
  }
  __device__ void finalize(const double (&tot)[2]) const {
    if (out) {
      out[0] = tot[0];
      out[1] = tot[1];
    }
  }
};

// ---- (x, y), unconjugated ------------------------------------------------------
struct OpZDot {
  using VT = zd;
  static constexpr int BS = 1;
  static constexpr int NV = 2;
  const zd* __restrict__ x;
  const zd* __restrict__ y;
  double* out;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ void row(int i, bool valid, double (&part)[2], const WaveRows<zd, 1>&) const {
    if (valid) zacc(part, 0, x[i], y[i]);
  }
  __device__ void finalize(const double (&tot)[2]) const {
    out[0] = tot[0];
    out[1] = tot[1];
  }
};

// ---- COCG: p = beta p + z (COCG.f90:78-83; first iteration beta = 0: p = z) ----
struct OpZP {
  using VT = zd;
  static constexpr int BS = 1;
  static constexpr int NV = 0;
  const zd* __restrict__ z;
  zd* __restrict__ p;
  ZState* st;
  zd beta;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() { beta = st->beta; }
  __device__ void row(int i, bool valid, double (&)[1], const WaveRows<zd, 1>&) const {
    if (valid) p[i] = beta * p[i] + z[i];
  }
  __device__ void finalize(const double (&)[1]) const {}
};

// ---- COCG: q = A p; omega = (p, q); alpha = rho / omega (COCG.f90:84-86) ------
template <int B>
struct OpZSpmvPQ {
  static constexpr int NV = 2;
  const zd* __restrict__ p;
  zd* __restrict__ q;
  ZState* st;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void gather(int c, zd (&xv)[B]) const { zgather<B>(p, c, xv); }
  __device__ void epilogue(int, bool valid, const zd (&acc)[B], double (&part)[2], const WaveRows<zd, B>& io) const {
    io.store(q, acc);
    zd pv[B];
    io.load(p, pv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) zacc(part, 0, pv[k], acc[k]);
    }
  }
  __device__ void finalize(const double (&tot)[2]) const { st->alpha = zdiv(st->rho, zd(tot[0], tot[1])); }
};

// ---- COCG: x += alpha p; r -= alpha q; test; z = M^-1 r; rho = (r, z); beta ----
// F_EXTPREC: z, rho and beta come from OpZVPrec (variable blocks) instead.
template <int B, bool PREC>
struct OpZUpdate {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = zd;
  static constexpr int BS = B;
  static constexpr int MIN_WAVES = 8;  // two 1024-thread workgroups per CU
  static constexpr int NV = 4;
  zd* __restrict__ x;
  zd* __restrict__ r;
  zd* __restrict__ z;
  const zd* __restrict__ p;
  const zd* __restrict__ q;
  const zd* __restrict__ D;
  int n;
  ZState* st;
  double* hist;
  int it;
  int flags;
  zd alpha;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() { alpha = st->alpha; }
  __device__ void row(int i, bool valid, double (&part)[4], const WaveRows<zd, B>& io) const {
    if (flags & (F_UPD | F_XONLY)) {
      zd xv[B], pv[B];
      io.load(x, xv);
      io.load(p, pv);
#pragma unroll
      for (int k = 0; k < B; ++k) xv[k] = alpha * pv[k] + xv[k];  // X = alpha*P + X  (COCG.f90:87)
      io.store(x, xv);
    }
    if (flags & F_XONLY) return;
    zd v[B];
    io.load(r, v);
    if (flags & F_UPD) {
      zd qv[B];
      io.load(q, qv);
#pragma unroll
      for (int k = 0; k < B; ++k) v[k] = (-alpha) * qv[k] + v[k];  // R = -alpha*Q + R  (COCG.f90:92)
      io.store(r, v);
    }
    if (flags & F_EXTPREC) {
      if (valid) {
#pragma unroll
        for (int k = 0; k < B; ++k) zacc(part, 0, v[k], v[k]);
      }
      return;
    }
    zd w[B];
#pragma unroll
    for (int k = 0; k < B; ++k) w[k] = v[k];
    if (valid) {
      if constexpr (PREC) zdiag_apply<B>(D, n, i, w);
#pragma unroll
      for (int k = 0; k < B; ++k) {
        zacc(part, 0, v[k], v[k]);
        zacc(part, 2, v[k], w[k]);
      }
    }
    io.store(z, w);
  }
  __device__ void finalize(const double (&tot)[4]) const {
    if (flags & F_XONLY) return;
    if (flags & F_CHECK) {
      if (zfinalize_check(st, hist, it, zd(tot[0], tot[1]))) return;
    }
    if (flags & F_EXTPREC) return;
    const zd rho(tot[2], tot[3]);
    st->beta = (flags & F_FIRST) ? zd(0.0) : zdiv(rho, st->rho);
    st->rho = rho;
  }
};

// ---- variable blocks: z = M_V^-1 r per block row (diag_V.f90:208-250), rho = (r, z),
// beta = rho / rho_old (first: 0) ------------------------------------------------
struct OpZVPrec {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = zd;
  static constexpr int BS = 1;
  static constexpr int NV = 2;
  const zd* __restrict__ r;
  zd* __restrict__ z;
  const int* __restrict__ list;
  const int* __restrict__ dof_index;
  const long long* __restrict__ doff;
  const zd* __restrict__ D;
  ZState* st;
  int first;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void row(int i, bool valid, double (&part)[2], const WaveRows<zd, 1>&) const {
    if (!valid) return;
    const int nd = list[i];
    const int o = dof_index[i];
    const zd* L = D + doff[i];
    zd t[kMaxVBlock];
    for (int j = 0; j < nd; ++j) t[j] = r[o + j];
    for (int j = 1; j < nd; ++j)
      for (int k = 0; k < j; ++k) t[j] = t[j] - L[nd * j + k] * t[k];
    for (int j = nd - 1; j >= 0; --j) {
      for (int k = nd - 1; k > j; --k) t[j] = t[j] - L[nd * j + k] * t[k];
      t[j] = L[(nd + 1) * j] * t[j];
    }
    for (int j = 0; j < nd; ++j) {
      z[o + j] = t[j];
      zacc(part, 0, r[o + j], t[j]);
    }
  }
  __device__ void finalize(const double (&tot)[2]) const {
    const zd rho(tot[0], tot[1]);
    st->beta = first ? zd(0.0) : zdiv(rho, st->rho);
    st->rho = rho;
  }
};

// ---- complex block-Jacobi setup (diag_nn.f90:90-159) ---------------------------
// One thread per block row: copy the diagonal block, then the in-place no-pivot
// LU with reciprocal pivots.  Rows without a diagonal block keep zeros.  A zero
// pivot records the smallest such row in *bad (the reference stops).
__global__ void k_zdiag_setup(int n, int b, const int* __restrict__ slice_ptr, const int* __restrict__ rowptr,
                              const int* __restrict__ col, const zd* __restrict__ val, zd* __restrict__ D,
                              int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = i / kWave, lane = i % kWave;
  const int bb = b * b;
  const size_t N = (size_t)n;
  const int k0 = slice_ptr[s];
  const int len = rowptr[i + 1] - rowptr[i];
  int found = 0;
  for (int e = 0; e < bb; ++e) D[(size_t)e * N + i] = zd(0.0);
  for (int k = 0; k < len; ++k) {
    const size_t slot = (size_t)(k0 + k);
    if (col[slot * kWave + lane] == i) {
      found = 1;
      for (int e = 0; e < bb; ++e) D[(size_t)e * N + i] = val[sell_vpos<zd>(slot, e, bb, lane)];
    }
  }
  if (!found) return;
#define LU(r, c) D[(size_t)((r) * b + (c)) * N + i]
  for (int k = 0; k < b; ++k) {
    const zd piv = LU(k, k);
    if (piv.re == 0.0 && piv.im == 0.0) {
      atomicMin(bad, i);
      return;
    }
    LU(k, k) = zdiv(zd(1.0), piv);
    for (int l = k + 1; l < b; ++l) {
      LU(l, k) = LU(l, k) * LU(k, k);
      for (int j = k + 1; j < b; ++j) LU(l, j) = LU(l, j) - LU(l, k) * LU(k, j);
    }
  }
#undef LU
}

size_t cocg_bytes(const Matrix& M, int max_iter) {
  const size_t vec = (size_t)M.np * M.b;
  const size_t parts = partials_needed((long long)M.n) * 4 + partials_needed((long long)M.nslices * kWave) * 4;
  return 4 * (vec * sizeof(zd) + 256) + parts * sizeof(double) + sizeof(ZState) + 256 +
         ((size_t)max_iter + 2) * sizeof(double) + 4096;
}

template <int B, bool PREC>
int cocg_impl(Matrix& M, const SolveParams& P, const zd* b, zd* x, Workspace& W, ZSolveResult* out,
              hipStream_t s) {
  Context* c = nullptr;
  MH_TRY(context(&c));
  const size_t vec = (size_t)M.np * B;
  MH_TRY(W.ensure(cocg_bytes(M, P.max_iter)));
  Carve cv{W.raw.p};
  zd* r = cv.take<zd>(vec);
  zd* z = cv.take<zd>(vec);
  zd* q = cv.take<zd>(vec);
  zd* p = cv.take<zd>(vec);
  const size_t nparts = partials_needed((long long)M.n) * 4 + partials_needed((long long)M.nslices * kWave) * 4;
  double* parts = cv.take<double>(nparts);
  ZState* st = cv.take<ZState>(1);
  double* hist = cv.take<double>((size_t)P.max_iter + 2);
  const zd* D = PREC ? M.Dz.p : nullptr;
  const int n = M.n;
  const size_t m = (size_t)n * B;
  const bool vprec = PREC && B == 1 && M.vb != nullptr;  // variable blocks: diag_V_C per block row

  k_zstate_init<<<1, 64, 0, s>>>(st, P.tol, P.max_iter);
  MH_HIP(hipGetLastError());
  if (P.init_x) MH_HIP(hipMemsetAsync(x, 0, sizeof(zd) * m, s));
  MH_HIP(hipMemsetAsync(p, 0, sizeof(zd) * vec, s));
  if (vec > m) MH_HIP(hipMemsetAsync(z + m, 0, sizeof(zd) * (vec - m), s));
  // R = B - A X; B2 = (R, R)  (COCG.f90:52-53, mat_converge.f90:77)
  {
    OpZResid<B> op{x, b, r, (double*)&st->B2, nullptr};
    MH_TRY((launch_zsell<B>(M, op, parts, s)));
  }
  zd B2(0.0);
  MH_HIP(hipMemcpyAsync(&B2, &st->B2, sizeof(zd), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  out->hist.clear();
  if (B2.re == 0.0 && B2.im == 0.0) {  // mat_converge.f90:79-85: X = 0, converged, no iteration
    MH_HIP(hipMemsetAsync(x, 0, sizeof(zd) * m, s));
    MH_HIP(hipStreamSynchronize(s));
    out->zero_rhs = 1;
    out->converged = 1;
    out->iter = -1;
    return 0;
  }
  auto vprec_apply = [&](int first) -> int {
    const VBlock& V = *M.vb;
    OpZVPrec op{r, z, V.list.p, V.dof_index.p, V.doff.p, V.Dz.p, st, first};
    return launch_rows(V.nb, op, parts, s);
  };
  auto upd = [&](int it, int flags) -> int {
    if (vprec) {
      if (flags & (F_UPD | F_XONLY | F_CHECK)) {
        OpZUpdate<B, false> op{x, r, z, p, q, nullptr, n, st, P.keep_hist ? hist : nullptr, it, flags | F_EXTPREC,
                               zd(0.0)};
        MH_TRY(launch_rows(n, op, parts, s));
      }
      return (flags & F_XONLY) ? 0 : vprec_apply((flags & F_FIRST) != 0);
    }
    OpZUpdate<B, PREC> op{x, r, z, p, q, D, n, st, P.keep_hist ? hist : nullptr, it, flags, zd(0.0)};
    return launch_rows(n, op, parts, s);
  };
  MH_TRY(upd(0, F_FIRST));  // z = M^-1 r, rho of iteration 1
  Poller poll{c, &st->resid_abs, P.tol};
  int done = 0;
  for (int it = 1; it <= P.max_iter && !done;) {
    const int end = std::min(it + poll.next_chunk(it, P.chunk) - 1, P.max_iter);
    for (; it <= end; ++it) {
      OpZP pu{z, p, st, zd(0.0)};
      MH_TRY(launch_rows((int)m, pu, parts, s));
      OpZSpmvPQ<B> sp{p, q, st};
      MH_TRY((launch_zsell<B>(M, sp, parts, s)));
      if (it % kIterRR == 0) {  // COCG.f90:88-90: true residual
        MH_TRY(upd(it, F_XONLY));
        OpZResid<B> rs{x, b, r, nullptr, st};
        MH_TRY((launch_zsell<B>(M, rs, parts, s)));
        MH_TRY(upd(it, F_CHECK));
      } else {
        MH_TRY(upd(it, F_UPD | F_CHECK));
      }
    }
    MH_TRY(poll.issue(s));
    MH_TRY(poll.check(1, &done));
  }
  MH_HIP(hipStreamSynchronize(s));
  ZState hs;
  MH_HIP(hipMemcpy(&hs, st, sizeof(ZState), hipMemcpyDeviceToHost));
  out->iter = hs.iter;
  out->resid_abs = hs.resid_abs;
  out->resid_re = hs.resid_re;
  out->converged = hs.converged;
  if (P.keep_hist && hs.iter > 0) {
    out->hist.resize((size_t)hs.iter + 1);
    MH_HIP(hipMemcpy(out->hist.data(), hist, sizeof(double) * ((size_t)hs.iter + 1), hipMemcpyDeviceToHost));
  }
  return 0;
}

int require_complex(const Matrix& M) {
  if (!M.is_complex) {
    set_error("matrix holds real values; the complex path needs a complex matrix");
    return -2;
  }
  return 0;
}

}  // namespace

int zspmv(const Matrix& M, const zd* x, zd* y, hipStream_t s) {
  MH_TRY(require_complex(M));
  MH_DISPATCH_B(M.b, {
    OpZSpmv<B> op{x, y};
    MH_TRY((launch_zsell<B>(M, op, nullptr, s)));
  });
  return 0;
}

int zdot(long long m, const zd* x, const zd* y, double* out, double* partials, hipStream_t s) {
  if (m > 0x7fffffffLL) {
    set_error("vector too long");
    return -2;
  }
  OpZDot op{x, y, out};
  return launch_rows((int)m, op, partials, s);
}

int zdiag_setup(Matrix& M, hipStream_t s) {
  MH_TRY(require_complex(M));
  if (M.vb) {  // the diag_V_C factors are built with the values
    if (M.vb->zbad >= 0) {
      set_error("zero diag in monolis_precond_diag_V_setup_C: zero pivot in the diagonal block of block row %d "
                "(0-based)", M.vb->zbad);
      return -6;
    }
    M.diagz_ready = true;
    return 0;
  }
  if (M.diagz_ready) return 0;
  const size_t need = (size_t)M.n * M.b * M.b;
  if (M.Dz.n < need) MH_TRY(M.Dz.alloc(need));
  if (M.n > 0) {
    DevArray<int> bad;
    MH_TRY(bad.alloc(1));
    MH_HIP(hipMemsetAsync(bad.p, 0x7f, sizeof(int), s));
    k_zdiag_setup<<<(M.n + kBlock - 1) / kBlock, kBlock, 0, s>>>(M.n, M.b, M.slice_ptr.p, M.rowptr.p, M.col.p,
                                                                  M.valz.p, M.Dz.p, bad.p);
    MH_HIP(hipGetLastError());
    int hbad = 0;
    MH_HIP(hipMemcpyAsync(&hbad, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    MH_HIP(hipStreamSynchronize(s));
    if (hbad != 0x7f7f7f7f) {
      set_error("zero diag in monolis_precond_diag_nn_setup: zero pivot in the diagonal block of block row %d "
                "(0-based)", hbad);
      return -6;
    }
  }
  M.diagz_ready = true;
  return 0;
}

int cocg_solve(Matrix& M, const SolveParams& P, const zd* b, zd* x, Workspace& W, ZSolveResult* out,
               hipStream_t s) {
  MH_TRY(require_complex(M));
  if (P.method != 13) {
    set_error("method %d: the complex path runs COCG (13) only, as monolis_solver_select_C", P.method);
    return -4;
  }
  if (P.precond != 0 && P.precond != 1) {
    set_error("preconditioner %d is not implemented on the complex GPU path (NONE = 0, DIAG = 1)", P.precond);
    return -3;
  }
  if (P.max_iter < 0) {
    set_error("max_iter < 0");
    return -2;
  }
  if (P.precond == 1) MH_TRY(zdiag_setup(M, s));
  int rc = 0;
  if (P.precond == 1) {
    MH_DISPATCH_B(M.b, rc = (cocg_impl<B, true>(M, P, b, x, W, out, s)));
  } else {
    MH_DISPATCH_B(M.b, rc = (cocg_impl<B, false>(M, P, b, x, W, out, s)));
  }
  return rc;
}

}  // namespace mh
