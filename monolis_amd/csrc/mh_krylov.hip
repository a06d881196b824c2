// Device-resident CG and BiCGSTAB (monolis src/iterative/CG.f90:18-119,
// src/iterative/BiCGSTAB.f90:18-192) and FP32-inner / FP64 iterative refinement.
//
// Every iteration is two (CG) or five (BiCGSTAB) fused kernels.  Scalars
// (rho, alpha, beta, omega, residual) never leave the GPU: the last workgroup
// of each reduction derives them and writes them to a KState in device
// memory, and it also runs monolis's convergence test
// (src/linalg/mat_converge.f90:90-141) and raises `done`.  The host enqueues
// iterations in chunks and polls `done` through pinned memory one chunk behind,
// so the GPU never waits for the CPU; kernels launched after `done` exit at
// once.  The true-residual replacement every 200 iterations (CG.f90:91-95,
// BiCGSTAB.f90:163-167) is kept.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>

#include "mh_internal.hpp"
#include "mh_launch.hpp"
#include "mh_vblock.hpp"

namespace mh {

KTimer*& active_ktimer() {
  thread_local KTimer* t = nullptr;
  return t;
}

namespace {

constexpr int kIterRR = 200;

// F_EXTPREC: the preconditioner runs as its own kernel afterwards (diag_V for
// variable block sizes); this step only updates x, r and tests (r, r).
enum : int { F_UPD = 1, F_CHECK = 2, F_FIRST = 4, F_XONLY = 8, F_EXTPREC = 16 };

// ---- shared finalize: monolis_check_converge_R --------------------------------
// B2 = (r0, r0) read back for the zero right-hand-side test (mat_converge.f90:40-48).
// Callers that know r0 != 0 (the inner solves of the refinement, whose right-hand
// side is a normalised nonzero residual) skip the host round trip.
int read_B2(const SolveParams& P, const KState* st, double* B2, hipStream_t s) {
  *B2 = 1.0;
  if (P.rhs_nonzero) return 0;
  MH_HIP(hipMemcpyAsync(B2, &st->B2, sizeof(double), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  return 0;
}

__device__ __forceinline__ bool finalize_check(KState* st, double* hist, int it, double rr) {
  const double resid = sqrt(rr / st->B2);
  st->rr = rr;
  st->resid = resid;
  st->iter = it;
  if (hist) hist[it] = resid;
  if (resid < st->tol) {
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

// ---- residual used inside a solve (honours `done`, optional copy) -----------
template <typename T, int B>
struct OpResidSolve {
  static constexpr int NV = 1;
  const T* __restrict__ x;
  const T* __restrict__ b;
  T* __restrict__ r;
  T* __restrict__ r2;   // optional second copy (BiCGSTAB r~)
  double* out;          // (r, r) destination, may be null
  const KState* st;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void gather(int c, T (&xv)[B]) const {
#pragma unroll
    for (int k = 0; k < B; ++k) xv[k] = x[(size_t)c * B + k];
  }
  __device__ void epilogue(int, bool valid, const T (&acc)[B], double (&part)[1], const WaveRows<T, B>& io) const {
    T v[B];
    io.load(b, v);
#pragma unroll
    for (int k = 0; k < B; ++k) {
      v[k] = v[k] - acc[k];  // R = 1*B + (-1)*(A X)  (matvec.f90:901)
      if (valid) part[0] += (double)v[k] * (double)v[k];
    }
    io.store(r, v);
    if (r2) io.store(r2, v);
  }
  __device__ void finalize(const double (&tot)[1]) const {
    if (out) *out = tot[0];
  }
};

// ---- deferred convergence test (mat_converge.f90:90-141) ---------------------
// Evaluated by every workgroup of the kernel that consumes iteration it's
// (r, r) partials; only the writer records it.  True when the solve is done.
__device__ __forceinline__ bool check_deferred(KState* st, double* hist, int it, double rr, bool writer) {
  const double resid = sqrt(rr / st->B2);
  const bool conv = resid < st->tol;
  const bool done = conv || it >= st->max_iter;
  if (writer) {
    st->rr = rr;
    st->resid = resid;
    st->iter = it;
    if (hist) hist[it] = resid;
    if (conv) st->converged = 1;
    if (done) st->done = 1;
  }
  return done;
}

// ---- CG: test of iteration it-1; rho, beta; p = z + beta p ---------------------
// (CG.f90:71-84, and mat_converge.f90:90-141 for the previous iteration.)
// Elementwise over n*B entries.  The prologue takes the previous update's
// partials (r, r), (r, z) -- with variable blocks (r, z) comes from the
// preconditioner kernel's own partials -- and derives rho and
// beta = rho / rho_1 (first iteration: beta = 0 with p = 0, so p = z).
// (Forming p at every gather of the SpMV instead, from z and the previous p,
// measured 361 us against 334 + 14 us for the SpMV and this pass.)
template <typename T>
struct OpCgP {
  using VT = T;
  static constexpr int BS = 1;
  static constexpr int MIN_WAVES = 8;  // two 1024-thread workgroups per CU
  static constexpr int NV = 0;
  static constexpr int NPRE = 2;
  const T* __restrict__ z;
  T* __restrict__ p;
  KState* st;
  double* hist;
  const double* pre;   // update partials, (r, r) and (r, z) per tile
  int npre;
  const double* pre2;  // optional (r, z) partials, one per tile
  int npre2;
  int it;
  T beta;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ bool prologue(const double (&tot)[2], bool writer) {
    if (it > 1 && check_deferred(st, hist, it - 1, tot[0], writer)) return false;
    const double rho = tot[1];
    const double bt = it > 1 ? rho / st->rhok[(it - 1) & 1] : 0.0;  // CG.f90:80-84
    beta = (T)bt;
    if (writer) {
      st->rhok[it & 1] = rho;
      st->rho = rho;
      st->beta = bt;
    }
    return true;
  }
  __device__ void row(int i, bool valid, double (&)[1], const WaveRows<T, 1>&) const {
    if (valid) p[i] = fma(beta, p[i], z[i]);  // P = beta*P + 1*Z
  }
  __device__ void finalize(const double (&)[1]) const {}
};

// ---- CG: q = A p; partials (p, q) for the update's alpha (CG.f90:86-88) ---------
template <typename T, int B>
struct OpCgSpmv {
  static constexpr int NV = 1;
  const T* __restrict__ p;
  T* __restrict__ q;
  KState* st;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void gather(int c, T (&xv)[B]) const {
#pragma unroll
    for (int k = 0; k < B; ++k) xv[k] = p[(size_t)c * B + k];
  }
  __device__ void epilogue(int, bool valid, const T (&acc)[B], double (&part)[1], const WaveRows<T, B>& io) const {
    io.store(q, acc);
    T pv[B];
    io.load(p, pv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) part[0] += (double)pv[k] * (double)acc[k];
    }
  }
  __device__ void finalize(const double (&)[1]) const {}
};

// ---- CG: alpha = rho / (p, q); x += alpha p; r -= alpha q; z = M^-1 r ---------
// (CG.f90:86-100, then the next iteration's preconditioner CG.f90:71.)  The
// prologue (with F_UPD or F_XONLY) takes the SpMV's (p, q) partials; the
// partials (r, r) and (r, z) are left for the next SpMV's prologue.
template <typename T, int B, bool PREC>
struct OpCgUpdate {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = T;
  static constexpr int BS = B;
  static constexpr int MIN_WAVES = B <= 3 ? 8 : 1;  // two 1024-thread workgroups per CU
  static constexpr int NV = 2;
  static constexpr int NPRE = 1;
  T* __restrict__ x;
  T* __restrict__ r;
  T* __restrict__ z;
  const T* __restrict__ p;
  const T* __restrict__ q;
  const T* __restrict__ D;
  int n;
  KState* st;
  int it;
  int flags;
  const double* pre;  // SpMV partials (p, q), one per tile; null: no x / r update
  int npre;
  T alpha;
  const double* pre2 = nullptr;  // unused: every sum comes from `pre`
  int npre2 = 0;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ bool prologue(const double (&tot)[1], bool writer) {
    const double a = st->rhok[it & 1] / tot[0];  // CG.f90:87-88
    alpha = (T)a;
    if (writer) {
      st->pq = tot[0];
      st->alpha = a;
    }
    return true;
  }
  // The loads of x, p, r and q are issued before the first store; the
  // elementwise updates run in the coalesced order of the loads.
  // The usual flags (F_UPD) get a copy of the body with them as constants, so
  // its loads are straight-line code.
  __device__ void row(int i, bool valid, double (&part)[2], const WaveRows<T, B>& io) const {
    if (flags == F_UPD) row_f<F_UPD>(i, valid, part, io);
    else row_f<-1>(i, valid, part, io);
  }
  template <int FL>
  __device__ __forceinline__ void row_f(int i, bool valid, double (&part)[2], const WaveRows<T, B>& io) const {
    const int f = FL >= 0 ? FL : flags;
    const bool upd = f & F_UPD, xupd = f & (F_UPD | F_XONLY), xonly = f & F_XONLY;
    const bool prec = !xonly && !(f & F_EXTPREC);
    T xc[B], pc[B], rc[B], qc[B];
    if (xupd) {
      io.fetch(x, xc);
      io.fetch(p, pc);
    }
    if (!xonly) io.fetch(r, rc);
    if (upd) io.fetch(q, qc);
    if (xupd) {
#pragma unroll
      for (int j = 0; j < B; ++j) xc[j] = fma(alpha, pc[j], xc[j]);  // X = alpha*P + X  (CG.f90:89)
      io.put_flat(x, xc);
    }
    // the diagonal block's loads follow the x store, in the registers x and p
    // leave (loading it with the vectors would cost two workgroups per CU)
    DiagRegs<T, T, B, PREC> dg;
    if (prec && valid) dg.load(D, n, i);
    if (xonly) return;
    if (upd) {
#pragma unroll
      for (int j = 0; j < B; ++j) rc[j] = fma(-alpha, qc[j], rc[j]);  // R = -alpha*Q + R  (CG.f90:94)
      io.put_flat(r, rc);
    }
    T v[B];
    io.transpose(rc, v);
    if (!prec) {
      if (valid) {
#pragma unroll
        for (int k = 0; k < B; ++k) part[0] += (double)v[k] * (double)v[k];
      }
      return;
    }
    T w[B];
#pragma unroll
    for (int k = 0; k < B; ++k) w[k] = v[k];
    if (valid) {
      dg.apply(D, n, i, w);
#pragma unroll
      for (int k = 0; k < B; ++k) {
        part[0] += (double)v[k] * (double)v[k];
        part[1] += (double)v[k] * (double)w[k];
      }
    }
    io.store(z, w);
  }
  __device__ void finalize(const double (&)[2]) const {}
};

// ---- BiCGSTAB ---------------------------------------------------------------
// Five launches per iteration and no finalize kernels: each kernel's prologue
// takes the partial sums the kernel before it stored.  A resumed stage (after
// the host's sorted-sum fallback) runs without its prologue (pre = null) and
// takes its scalars from the state the host wrote.
//
// a: test of iteration it-1 and rho = (r, rt), beta (prologue, from e's
//    partials); p = r + beta (p - omega v) (first: p = r); pt = M^-1 p
template <typename T, int B, bool PREC>
struct OpBiA {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = T;
  static constexpr int BS = B;
  static constexpr int MIN_WAVES = B <= 3 ? 8 : 1;  // two 1024-thread workgroups per CU
  static constexpr int NV = 0;
  static constexpr int NPRE = 2;
  const T* __restrict__ r;
  T* __restrict__ p;
  const T* __restrict__ v;
  T* __restrict__ pt;
  const T* __restrict__ D;
  int n;
  KState* st;
  double* hist;
  const double* pre;  // e's partials (r, r), (r, rt)
  int npre;
  int it;
  T beta, omega;
  const double* pre2 = nullptr;  // unused: every sum comes from `pre`
  int npre2 = 0;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {
    beta = (T)st->beta;
    omega = (T)st->omega;
  }
  __device__ bool prologue(const double (&tot)[2], bool writer) {
    if (check_deferred(st, hist, it - 1, tot[0], writer)) return false;
    const double rho = tot[1];
    if (rho == 0.0) {  // BiCGSTAB.f90:89-93: the host recomputes it with the sorted sum
      if (writer) {
        st->breakdown |= 1;
        st->bd_stage = 5;
        st->bd_iter = it - 1;
        st->done = 1;
      }
      return false;
    }
    const double rho1 = st->rhok[(it - 1) & 1];
    const double bt = (rho / rho1) * (st->alpha / st->omega);  // BiCGSTAB.f90:96
    beta = (T)bt;
    if (writer) {
      st->beta = bt;
      st->rho_prev = rho1;
      st->rho = rho;
      st->rhok[it & 1] = rho;
    }
    return true;
  }
  __device__ void row(int i, bool valid, double (&part)[1], const WaveRows<T, B>& io) const {
    if (it > 1) row_f<true>(i, valid, part, io);
    else row_f<false>(i, valid, part, io);
  }
  template <bool REC>
  __device__ __forceinline__ void row_f(int i, bool valid, double (&)[1], const WaveRows<T, B>& io) const {
    T pc[B], po[B], vc[B];
    io.fetch(r, pc);
    if (REC) {
      io.fetch(p, po);
      io.fetch(v, vc);
    }
    DiagRegs<T, T, B, PREC> dg;
    if (valid) dg.load(D, n, i);
    if (REC) {
#pragma unroll
      for (int j = 0; j < B; ++j) pc[j] = pc[j] + beta * (po[j] - omega * vc[j]);  // BiCGSTAB.f90:103
    }
    io.put_flat(p, pc);
    if constexpr (PREC) {
      T pv[B];
      io.transpose(pc, pv);
      if (valid) dg.apply(D, n, i, pv);
      io.store(pt, pv);
    } else {
      io.put_flat(pt, pc);
    }
  }
  __device__ void finalize(const double (&)[1]) const {}
};

// b: v = A pt; partials (rt, v) for c
template <typename T, int B>
struct OpBiB {
  static constexpr int NV = 1;
  const T* __restrict__ pt;
  T* __restrict__ v;
  const T* __restrict__ rt;
  KState* st;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void gather(int c, T (&xv)[B]) const {
#pragma unroll
    for (int k = 0; k < B; ++k) xv[k] = pt[(size_t)c * B + k];
  }
  __device__ void epilogue(int, bool valid, const T (&acc)[B], double (&part)[1], const WaveRows<T, B>& io) const {
    io.store(v, acc);
    T rtv[B];
    io.load(rt, rtv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) part[0] += (double)rtv[k] * (double)acc[k];
    }
  }
  __device__ void finalize(const double (&)[1]) const {}
};

// c: alpha = rho / (rt, v) (prologue, from b's partials); s = r - alpha v; st = M^-1 s
template <typename T, int B, bool PREC>
struct OpBiC {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = T;
  static constexpr int BS = B;
  static constexpr int MIN_WAVES = B <= 3 ? 8 : 1;  // two 1024-thread workgroups per CU
  static constexpr int NV = 0;
  static constexpr int NPRE = 1;
  const T* __restrict__ r;
  const T* __restrict__ v;
  T* __restrict__ s;
  T* __restrict__ st_;
  const T* __restrict__ D;
  int n;
  KState* st;
  const double* pre;  // b's partials (rt, v)
  int npre;
  int it;
  T alpha;
  const double* pre2 = nullptr;  // unused: every sum comes from `pre`
  int npre2 = 0;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() { alpha = (T)st->alpha; }
  __device__ bool prologue(const double (&tot)[1], bool writer) {
    const double c2 = tot[0];
    if (c2 == 0.0) {  // BiCGSTAB.f90:110-114: the host recomputes it with the sorted sum
      if (writer) {
        st->pq = c2;
        st->breakdown |= 2;
        st->bd_stage = 2;
        st->bd_iter = it;
        st->done = 1;
      }
      return false;
    }
    const double a = st->rhok[it & 1] / c2;  // BiCGSTAB.f90:116
    alpha = (T)a;
    if (writer) {
      st->pq = c2;
      st->alpha = a;
    }
    return true;
  }
  __device__ void row(int i, bool valid, double (&)[1], const WaveRows<T, B>& io) const {
    T rc[B], vc[B];
    io.fetch(r, rc);
    io.fetch(v, vc);
    DiagRegs<T, T, B, PREC> dg;
    if (valid) dg.load(D, n, i);
#pragma unroll
    for (int j = 0; j < B; ++j) rc[j] = fma(-alpha, vc[j], rc[j]);  // S = -alpha*V + R  (BiCGSTAB.f90:118)
    io.put_flat(s, rc);
    if constexpr (PREC) {
      T rv[B];
      io.transpose(rc, rv);
      if (valid) dg.apply(D, n, i, rv);
      io.store(st_, rv);
    } else {
      io.put_flat(st_, rc);
    }
  }
  __device__ void finalize(const double (&)[1]) const {}
};

// d: t = A st; partials (t, s), (t, t) for e
template <typename T, int B>
struct OpBiD {
  static constexpr int NV = 2;
  const T* __restrict__ stv;
  T* __restrict__ t;
  const T* __restrict__ s;
  KState* st;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void gather(int c, T (&xv)[B]) const {
#pragma unroll
    for (int k = 0; k < B; ++k) xv[k] = stv[(size_t)c * B + k];
  }
  __device__ void epilogue(int, bool valid, const T (&acc)[B], double (&part)[2], const WaveRows<T, B>& io) const {
    io.store(t, acc);
    T sv[B];
    io.load(s, sv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        part[0] += (double)acc[k] * (double)sv[k];
        part[1] += (double)acc[k] * (double)acc[k];
      }
    }
  }
  __device__ void finalize(const double (&)[2]) const {}
};

// e: omega = (t, s) / (t, t) (prologue, from d's partials); x += alpha pt +
//    omega st; r = s - omega t; partials (r, r), (r, rt) for the next a
template <typename T, int B>
struct OpBiE {
  using VT = T;
  static constexpr int BS = B;
  static constexpr int MIN_WAVES = B <= 3 ? 8 : 1;  // two 1024-thread workgroups per CU
  static constexpr int NV = 2;
  static constexpr int NPRE = 2;
  T* __restrict__ x;
  T* __restrict__ r;
  const T* __restrict__ pt;
  const T* __restrict__ stv;
  const T* __restrict__ s;
  const T* __restrict__ t;
  const T* __restrict__ rt;
  KState* st;
  int it;
  int flags;
  const double* pre;  // d's partials (t, s), (t, t); null after a true residual
  int npre;
  T alpha, omega;
  const double* pre2 = nullptr;  // unused: every sum comes from `pre`
  int npre2 = 0;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {
    alpha = (T)st->alpha;
    omega = (T)st->omega;
  }
  __device__ bool prologue(const double (&tot)[2], bool writer) {
    if (tot[1] == 0.0) {  // BiCGSTAB.f90:132-137: the host recomputes (t, t) with the sorted sum
      if (writer) {
        st->ts = tot[0];
        st->tt = tot[1];
        st->breakdown |= 4;
        st->bd_stage = 4;
        st->bd_iter = it;
        st->done = 1;
      }
      return false;
    }
    const double w = tot[0] / tot[1];  // BiCGSTAB.f90:139
    omega = (T)w;
    if (writer) {
      st->ts = tot[0];
      st->tt = tot[1];
      st->omega = w;
    }
    return true;
  }
  // Every load is issued before the first store; the elementwise updates run in
  // the coalesced order of the loads, the dots in row order.  The usual flags
  // (F_UPD) get a copy of the body with them as constants.
  __device__ void row(int i, bool valid, double (&part)[2], const WaveRows<T, B>& io) const {
    if (flags == F_UPD) row_f<F_UPD>(i, valid, part, io);
    else row_f<-1>(i, valid, part, io);
  }
  template <int FL>
  __device__ __forceinline__ void row_f(int, bool valid, double (&part)[2], const WaveRows<T, B>& io) const {
    const int f = FL >= 0 ? FL : flags;
    const bool upd = f & F_UPD, xupd = f & (F_UPD | F_XONLY), xonly = f & F_XONLY;
    T xc[B], pc[B], sc[B], rc[B], tc[B], rtc[B];
    if (xupd) {
      io.fetch(x, xc);
      io.fetch(pt, pc);
      io.fetch(stv, sc);
    }
    if (upd) {
      io.fetch(s, rc);
      io.fetch(t, tc);
    } else if (!xonly) {
      io.fetch(r, rc);
    }
    if (!xonly) io.fetch(rt, rtc);
    if (xupd) {
#pragma unroll
      for (int j = 0; j < B; ++j) xc[j] = xc[j] + alpha * pc[j] + omega * sc[j];  // BiCGSTAB.f90:153
      io.put_flat(x, xc);
    }
    if (xonly) return;
    if (upd) {
#pragma unroll
      for (int j = 0; j < B; ++j) rc[j] = fma(-omega, tc[j], rc[j]);  // R = -omega*T + S  (BiCGSTAB.f90:166)
      io.put_flat(r, rc);
    }
    T rv[B], rtv[B];
    io.transpose(rc, rv);
    io.transpose(rtc, rtv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        part[0] += (double)rv[k] * (double)rv[k];
        part[1] += (double)rv[k] * (double)rtv[k];
      }
    }
  }
  __device__ void finalize(const double (&)[2]) const {}
};

// ---- PipeCG (src/iterative/PipeCG.f90) ---------------------------------------
// a: gamma = (r, u), delta = (v, u); m = M^-1 v; then the scalar recurrences
//    gamma1 = 1/gamma_old, alpha1 = 1/alpha_old, beta = gamma*gamma1,
//    alpha = gamma / (delta - beta*gamma*alpha1) (first: alpha = gamma/delta, beta = 0).
template <typename T, int B, bool PREC>
struct OpPipeA {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 2;
  const T* __restrict__ r;
  const T* __restrict__ u;
  const T* __restrict__ v;
  T* __restrict__ m;
  const T* __restrict__ D;
  int n;
  KState* st;
  int first;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void row(int i, bool valid, double (&part)[2], const WaveRows<T, B>& io) const {
    T rv[B], uv[B], vv[B];
    io.load(r, rv);
    io.load(u, uv);
    io.load(v, vv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        part[0] += (double)rv[k] * (double)uv[k];
        part[1] += (double)vv[k] * (double)uv[k];
      }
      if constexpr (PREC) diag_apply<T, T, B>(D, n, i, vv);
    }
    io.store(m, vv);
  }
  __device__ void finalize(const double (&tot)[2]) const {
    const double gamma1 = 1.0 / st->gamma;
    const double alpha1 = 1.0 / st->alpha;
    const double gamma = tot[0], delta = tot[1];
    st->gamma = gamma;
    st->delta = delta;
    if (!first) {
      st->beta = gamma * gamma1;
      st->alpha = gamma / (delta - st->beta * gamma * alpha1);
    } else {
      st->alpha = gamma / delta;
      st->beta = 0.0;
    }
  }
};

// y = A x (no reduction), skipped once the solve is done.
template <typename T, int B>
struct OpSpmvSolve {
  static constexpr int NV = 0;
  const T* __restrict__ x;
  T* __restrict__ y;
  const KState* st;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void gather(int c, T (&xv)[B]) const {
#pragma unroll
    for (int k = 0; k < B; ++k) xv[k] = x[(size_t)c * B + k];
  }
  __device__ void epilogue(int, bool, const T (&acc)[B], double (&)[1], const WaveRows<T, B>& io) const {
    io.store(y, acc);
  }
  __device__ void finalize(const double (&)[1]) const {}
};

// b: z = beta z + l; q = beta q + m; s = beta s + v; p = beta p + u; then
//    r -= alpha s; u -= alpha q; v -= alpha z; x += alpha p and the test on
//    (r, r) (PipeCG.f90 loop body), or x += alpha p only (F_XONLY: the
//    residual-replacement iteration).
template <typename T, int B>
struct OpPipeB {
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 1;
  T *__restrict__ z, *__restrict__ q, *__restrict__ s, *__restrict__ p;
  T *__restrict__ r, *__restrict__ u, *__restrict__ v, *__restrict__ x;
  const T* __restrict__ l;
  const T* __restrict__ m;
  KState* st;
  double* hist;
  int it;
  int flags;
  T alpha, beta;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {
    alpha = (T)st->alpha;
    beta = (T)st->beta;
  }
  __device__ void row(int, bool valid, double (&part)[1], const WaveRows<T, B>& io) const {
    T zv[B], lv[B], qv[B], mv[B], sv[B], vv[B], pv[B], uv[B];
    io.load(z, zv);
    io.load(l, lv);
    io.load(q, qv);
    io.load(m, mv);
    io.load(s, sv);
    io.load(v, vv);
    io.load(p, pv);
    io.load(u, uv);
#pragma unroll
    for (int k = 0; k < B; ++k) {
      zv[k] = fma(beta, zv[k], lv[k]);  // Z = beta*Z + L
      qv[k] = fma(beta, qv[k], mv[k]);  // Q = beta*Q + M
      sv[k] = fma(beta, sv[k], vv[k]);  // S = beta*S + V
      pv[k] = fma(beta, pv[k], uv[k]);  // P = beta*P + U
    }
    io.store(z, zv);
    io.store(q, qv);
    io.store(s, sv);
    io.store(p, pv);
    T xv[B];
    io.load(x, xv);
#pragma unroll
    for (int k = 0; k < B; ++k) xv[k] = fma(alpha, pv[k], xv[k]);  // X = alpha*P + X
    io.store(x, xv);
    if (flags & F_XONLY) return;
    T rv[B];
    io.load(r, rv);
#pragma unroll
    for (int k = 0; k < B; ++k) {
      rv[k] = fma(-alpha, sv[k], rv[k]);  // R = -alpha*S + R
      uv[k] = fma(-alpha, qv[k], uv[k]);  // U = -alpha*Q + U
      vv[k] = fma(-alpha, zv[k], vv[k]);  // V = -alpha*Z + V
    }
    io.store(r, rv);
    io.store(u, uv);
    io.store(v, vv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) part[0] += (double)rv[k] * (double)rv[k];
    }
  }
  __device__ void finalize(const double (&tot)[1]) const {
    if (flags & F_XONLY) return;
    finalize_check(st, hist, it, tot[0]);
  }
};

// u = M^-1 r, with the test on (r, r) when F_CHECK (after a true residual).
template <typename T, int B, bool PREC>
struct OpPrecCheck {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 1;
  const T* __restrict__ r;
  T* __restrict__ u;
  const T* __restrict__ D;
  int n;
  KState* st;
  double* hist;
  int it;
  int flags;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void row(int i, bool valid, double (&part)[1], const WaveRows<T, B>& io) const {
    T rv[B];
    io.load(r, rv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) part[0] += (double)rv[k] * (double)rv[k];
      if constexpr (PREC) diag_apply<T, T, B>(D, n, i, rv);
    }
    io.store(u, rv);
  }
  __device__ void finalize(const double (&tot)[1]) const {
    if (flags & F_CHECK) finalize_check(st, hist, it, tot[0]);
  }
};

// ---- PipeCR (src/iterative/PipeCR.f90) ---------------------------------------
// init: U2_0 = (u, u) of u = M^-1 r0; phi = sqrt(R2/U2_0), R2 = (r0, r0) = B2;
// the test threshold utol = tol/phi (PipeCR.f90:72-79).  KState::aux holds
// phi, utol, 1/gamma and 1/alpha.
template <typename T, int B>
struct OpCrPhi {
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 1;
  const T* __restrict__ u;
  KState* st;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void row(int, bool valid, double (&part)[1], const WaveRows<T, B>& io) const {
    T uv[B];
    io.load(u, uv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) part[0] += (double)uv[k] * (double)uv[k];
    }
  }
  __device__ void finalize(const double (&tot)[1]) const {
    const double phi = sqrt(st->B2 / tot[0]);
    st->aux[0] = phi;
    st->aux[1] = st->tol / phi;
  }
};

// a: m = M^-1 v (or m given: MGIVEN, the variable-block apply ran before);
//    gamma = (v, u), delta = (v, m), U2 = (u, u); then the test
//    sqrt(U2/B2) <= utol and, when it fails, beta = gamma/gamma_old and
//    alpha = gamma/(delta - beta*gamma/alpha_old) (first: beta = 0,
//    alpha = gamma/delta), PipeCR.f90:82-105.  cur_resid is reported as
//    sqrt(U2/B2)*phi.
template <typename T, int B, bool PREC, bool MGIVEN>
struct OpCrA {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 3;
  const T* __restrict__ u;
  const T* __restrict__ v;
  T* __restrict__ m;
  const T* __restrict__ D;
  int n;
  KState* st;
  double* hist;
  int it;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void row(int i, bool valid, double (&part)[3], const WaveRows<T, B>& io) const {
    T uv[B], vv[B], mv[B];
    io.load(u, uv);
    io.load(v, vv);
    if constexpr (MGIVEN) {
      io.load(m, mv);
    } else {
#pragma unroll
      for (int k = 0; k < B; ++k) mv[k] = vv[k];
      if constexpr (PREC) {
        if (valid) diag_apply<T, T, B>(D, n, i, mv);
      }
      io.store(m, mv);
    }
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        part[0] += (double)vv[k] * (double)uv[k];
        part[1] += (double)vv[k] * (double)mv[k];
        part[2] += (double)uv[k] * (double)uv[k];
      }
    }
  }
  __device__ void finalize(const double (&tot)[3]) const {
    const double gamma = tot[0], delta = tot[1], U2 = tot[2];
    const double resid = sqrt(U2 / st->B2);
    const double phi = st->aux[0];
    st->iter = it;
    st->rr = U2;
    st->resid = resid * phi;
    if (hist) hist[it] = resid * phi;
    if (resid <= st->aux[1]) {
      st->converged = 1;
      st->done = 1;
      return;
    }
    double alpha, beta;
    if (it > 1) {
      beta = gamma * st->aux[2];
      alpha = gamma / (delta - beta * gamma * st->aux[3]);
    } else {
      beta = 0.0;
      alpha = gamma / delta;
    }
    st->gamma = gamma;
    st->delta = delta;
    st->alpha = alpha;
    st->beta = beta;
    st->aux[2] = 1.0 / gamma;  // gamma1, PipeCR.f90:121
    st->aux[3] = 1.0 / alpha;  // alpha1, PipeCR.f90:122
  }
};

// b: z = beta z + l; q = beta q + m; p = beta p + u; x += alpha p; then
//    u -= alpha q; v -= alpha z, or x only (F_XONLY: the residual-replacement
//    iteration), PipeCR.f90:107-119.
template <typename T, int B>
struct OpCrB {
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 0;
  T *__restrict__ z, *__restrict__ q, *__restrict__ p;
  T *__restrict__ u, *__restrict__ v, *__restrict__ x;
  const T* __restrict__ l;
  const T* __restrict__ m;
  KState* st;
  int flags;
  T alpha, beta;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {
    alpha = (T)st->alpha;
    beta = (T)st->beta;
  }
  __device__ void row(int, bool, double (&)[1], const WaveRows<T, B>& io) const {
    T zv[B], lv[B], qv[B], mv[B], pv[B], uv[B], xv[B];
    io.load(z, zv);
    io.load(l, lv);
    io.load(q, qv);
    io.load(m, mv);
    io.load(p, pv);
    io.load(u, uv);
    io.load(x, xv);
#pragma unroll
    for (int k = 0; k < B; ++k) {
      zv[k] = fma(beta, zv[k], lv[k]);   // Z = beta*Z + L
      qv[k] = fma(beta, qv[k], mv[k]);   // Q = beta*Q + M
      pv[k] = fma(beta, pv[k], uv[k]);   // P = beta*P + U
      xv[k] = fma(alpha, pv[k], xv[k]);  // X = alpha*P + X
    }
    io.store(z, zv);
    io.store(q, qv);
    io.store(p, pv);
    io.store(x, xv);
    if (flags & F_XONLY) return;
    T vv[B];
    io.load(v, vv);
#pragma unroll
    for (int k = 0; k < B; ++k) {
      uv[k] = fma(-alpha, qv[k], uv[k]);  // U = -alpha*Q + U
      vv[k] = fma(-alpha, zv[k], vv[k]);  // V = -alpha*Z + V
    }
    io.store(u, uv);
    io.store(v, vv);
  }
  __device__ void finalize(const double (&)[1]) const {}
};

// ---- GropCG (src/iterative/GropCG.f90) ---------------------------------------
// init: u = M^-1 r; gamma = (r, u)
template <typename T, int B, bool PREC>
struct OpGropInit {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 1;
  const T* __restrict__ r;
  T* __restrict__ u;
  const T* __restrict__ D;
  int n;
  KState* st;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void row(int i, bool valid, double (&part)[1], const WaveRows<T, B>& io) const {
    T rv[B], w[B];
    io.load(r, rv);
#pragma unroll
    for (int k = 0; k < B; ++k) w[k] = rv[k];
    if (valid) {
      if constexpr (PREC) diag_apply<T, T, B>(D, n, i, w);
#pragma unroll
      for (int k = 0; k < B; ++k) part[0] += (double)rv[k] * (double)w[k];
    }
    io.store(u, w);
  }
  __device__ void finalize(const double (&tot)[1]) const { st->gamma = tot[0]; }
};

// x += alpha p; r -= alpha s; u -= alpha q; test on (r, r); beta = (r, u) / gamma,
// gamma = (r, u)  (GropCG.f90 loop body)
template <typename T, int B>
struct OpGropUpd {
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 2;
  T *__restrict__ x, *__restrict__ r, *__restrict__ u;
  const T *__restrict__ p, *__restrict__ s, *__restrict__ q;
  KState* st;
  double* hist;
  int it;
  T alpha;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() { alpha = (T)st->alpha; }
  __device__ void row(int, bool valid, double (&part)[2], const WaveRows<T, B>& io) const {
    T xv[B], pv[B], rv[B], sv[B], uv[B], qv[B];
    io.load(x, xv);
    io.load(p, pv);
    io.load(r, rv);
    io.load(s, sv);
    io.load(u, uv);
    io.load(q, qv);
#pragma unroll
    for (int k = 0; k < B; ++k) {
      xv[k] = fma(alpha, pv[k], xv[k]);   // X = alpha*P + X
      rv[k] = fma(-alpha, sv[k], rv[k]);  // R = -alpha*S + R
      uv[k] = fma(-alpha, qv[k], uv[k]);  // U = -alpha*Q + U
    }
    io.store(x, xv);
    io.store(r, rv);
    io.store(u, uv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        part[0] += (double)rv[k] * (double)uv[k];
        part[1] += (double)rv[k] * (double)rv[k];
      }
    }
  }
  __device__ void finalize(const double (&tot)[2]) const {
    if (finalize_check(st, hist, it, tot[1])) return;
    st->beta = tot[0] / st->gamma;
    st->gamma = tot[0];
  }
};

// v = A u fused with p = beta p + u; s = beta s + v; q = M^-1 s; delta = (p, s),
// alpha = gamma / delta for the next iteration.  v itself is never stored.
template <typename T, int B, bool PREC>
struct OpGropSpmv {
  static constexpr int NV = 1;
  const T* __restrict__ u;
  T* __restrict__ p;
  T* __restrict__ s;
  T* __restrict__ q;
  const T* __restrict__ D;
  int n;
  KState* st;
  T beta;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() { beta = (T)st->beta; }
  __device__ void gather(int c, T (&xv)[B]) const {
#pragma unroll
    for (int k = 0; k < B; ++k) xv[k] = u[(size_t)c * B + k];
  }
  __device__ void epilogue(int row, bool valid, const T (&acc)[B], double (&part)[1],
                           const WaveRows<T, B>& io) const {
    T pv[B], uv[B], sv[B];
    io.load(p, pv);
    io.load(u, uv);
    io.load(s, sv);
#pragma unroll
    for (int k = 0; k < B; ++k) {
      pv[k] = fma(beta, pv[k], uv[k]);   // P = beta*P + U
      sv[k] = fma(beta, sv[k], acc[k]);  // S = beta*S + V
    }
    io.store(p, pv);
    io.store(s, sv);
    T w[B];
#pragma unroll
    for (int k = 0; k < B; ++k) w[k] = sv[k];
    if (valid) {
      if constexpr (PREC) diag_apply<T, T, B>(D, n, row, w);
#pragma unroll
      for (int k = 0; k < B; ++k) part[0] += (double)pv[k] * (double)sv[k];
    }
    io.store(q, w);
  }
  __device__ void finalize(const double (&tot)[1]) const {
    st->delta = tot[0];
    st->alpha = st->gamma / tot[0];
  }
};

// ---- PipeBiCGSTAB (src/iterative/PipeBiCGSTAB.f90) ----------------------------
// a: the seven recurrences of the loop head, (Q, Y), (Y, Y) and ZT = M^-1 Z;
//    omega1 = QY / YY (0 when YY = 0).
template <typename T, int B, bool PREC>
struct OpPbA {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 2;
  T *__restrict__ pt, *__restrict__ s_, *__restrict__ st_, *__restrict__ z, *__restrict__ zt;
  T *__restrict__ q, *__restrict__ qt, *__restrict__ y;
  const T *__restrict__ rt, *__restrict__ w0, *__restrict__ wt, *__restrict__ t, *__restrict__ v,
      *__restrict__ r;
  const T* __restrict__ D;
  int n;
  KState* st;
  T alpha, beta, omega;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {
    alpha = (T)st->alpha;
    beta = (T)st->beta;
    omega = (T)st->omega;
  }
  __device__ void row(int i, bool valid, double (&part)[2], const WaveRows<T, B>& io) const {
    T ptv[B], sv[B], stv[B], zv[B], ztv[B], rtv[B], w0v[B], wtv[B], tv[B], vv[B], rv[B];
    io.load(pt, ptv);
    io.load(s_, sv);
    io.load(st_, stv);
    io.load(z, zv);
    io.load(zt, ztv);
    io.load(rt, rtv);
    io.load(w0, w0v);
    io.load(wt, wtv);
    io.load(t, tv);
    io.load(v, vv);
    io.load(r, rv);
    T qv[B], qtv[B], yv[B];
#pragma unroll
    for (int k = 0; k < B; ++k) {
      ptv[k] = rtv[k] + beta * (ptv[k] - omega * stv[k]);
      sv[k] = w0v[k] + beta * (sv[k] - omega * zv[k]);
      stv[k] = wtv[k] + beta * (stv[k] - omega * ztv[k]);
      zv[k] = tv[k] + beta * (zv[k] - omega * vv[k]);
      qv[k] = rv[k] - alpha * sv[k];
      qtv[k] = rtv[k] - alpha * stv[k];
      yv[k] = w0v[k] - alpha * zv[k];
    }
    io.store(pt, ptv);
    io.store(s_, sv);
    io.store(st_, stv);
    io.store(z, zv);
    io.store(q, qv);
    io.store(qt, qtv);
    io.store(y, yv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        part[0] += (double)qv[k] * (double)yv[k];
        part[1] += (double)yv[k] * (double)yv[k];
      }
      if constexpr (PREC) diag_apply<T, T, B>(D, n, i, zv);
    }
    io.store(zt, zv);
  }
  __device__ void finalize(const double (&tot)[2]) const {
    st->aux[0] = (tot[1] != 0.0) ? tot[0] / tot[1] : 0.0;  // omega1 (PipeBiCGSTAB.f90 YY guard)
  }
};

// b: (mode 0) X += alpha PT + omega1 QT; R = Q - omega1 Y; RT = QT - omega1 (WT - alpha ZT);
//    W0 = Y - omega1 (T - alpha V); then (mode 0 and 2) WT = M^-1 W0 and the dots
//    (R0,R), (R0,W0), (R0,S), (R0,Z), (R,R): test, beta, alpha, omega, RR.
//    mode 1: X += alpha PT + omega1 T only (the residual-replacement iteration,
//    as the reference writes it).
template <typename T, int B, bool PREC>
struct OpPbB {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 5;
  T *__restrict__ x, *__restrict__ r, *__restrict__ rt, *__restrict__ w0, *__restrict__ wt;
  const T *__restrict__ pt, *__restrict__ qt, *__restrict__ q, *__restrict__ y, *__restrict__ zt,
      *__restrict__ t, *__restrict__ v, *__restrict__ r0, *__restrict__ s_, *__restrict__ z;
  const T* __restrict__ D;
  int n;
  KState* st;
  double* hist;
  int it;
  int mode;
  T alpha, omega1;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {
    alpha = (T)st->alpha;
    omega1 = (T)st->aux[0];
  }
  __device__ void row(int i, bool valid, double (&part)[5], const WaveRows<T, B>& io) const {
    if (mode != 2) {
      T xv[B], ptv[B], ov[B];
      io.load(x, xv);
      io.load(pt, ptv);
      io.load(mode == 1 ? t : qt, ov);
#pragma unroll
      for (int k = 0; k < B; ++k) xv[k] = xv[k] + alpha * ptv[k] + omega1 * ov[k];
      io.store(x, xv);
      if (mode == 1) return;
    }
    T rv[B], w0v[B];
    if (mode == 0) {
      T qv[B], yv[B], qtv[B], wtv[B], ztv[B], tv[B], vv[B], rtv[B];
      io.load(q, qv);
      io.load(y, yv);
      io.load(qt, qtv);
      io.load(wt, wtv);
      io.load(zt, ztv);
      io.load(t, tv);
      io.load(v, vv);
#pragma unroll
      for (int k = 0; k < B; ++k) {
        rv[k] = qv[k] - omega1 * yv[k];
        rtv[k] = qtv[k] - omega1 * (wtv[k] - alpha * ztv[k]);
        w0v[k] = yv[k] - omega1 * (tv[k] - alpha * vv[k]);
      }
      io.store(r, rv);
      io.store(rt, rtv);
      io.store(w0, w0v);
    } else {
      io.load(r, rv);
      io.load(w0, w0v);
    }
    T r0v[B], sv[B], zv[B];
    io.load(r0, r0v);
    io.load(s_, sv);
    io.load(z, zv);
    T wn[B];
#pragma unroll
    for (int k = 0; k < B; ++k) wn[k] = w0v[k];
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        part[0] += (double)r0v[k] * (double)rv[k];
        part[1] += (double)r0v[k] * (double)w0v[k];
        part[2] += (double)r0v[k] * (double)sv[k];
        part[3] += (double)r0v[k] * (double)zv[k];
        part[4] += (double)rv[k] * (double)rv[k];
      }
      if constexpr (PREC) diag_apply<T, T, B>(D, n, i, wn);
    }
    io.store(wt, wn);
  }
  __device__ void finalize(const double (&tot)[5]) const {
    if (mode == 1) return;
    if (finalize_check(st, hist, it, tot[4])) return;
    const double RR1 = tot[0], RW = tot[1], RS = tot[2], RZ = tot[3];
    const double alpha_ = st->alpha, omega1_ = st->aux[0];
    const double beta_ = (alpha_ * RR1) / (st->rho * omega1_);
    st->beta = beta_;
    st->alpha = RR1 / (RW + beta_ * RS - beta_ * omega1_ * RZ);
    st->omega = omega1_;
    st->rho = RR1;
  }
};

// setup: R0 = R and RT = M^-1 R (mode 0); WT = M^-1 W0 with (R, R), (R, W0) ->
// alpha = RR / RW, beta = omega = 0 (mode 1); RT = M^-1 R only (mode 2, after a
// true residual).
template <typename T, int B, bool PREC>
struct OpPbInit {
  static constexpr int KCLASS = kClsPrec;  // applies the preconditioner
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 2;
  const T* __restrict__ r;
  T* __restrict__ r0;
  T* __restrict__ rt;
  const T* __restrict__ w0;
  T* __restrict__ wt;
  const T* __restrict__ D;
  int n;
  KState* st;
  int mode;
  __device__ bool skip() const { return st->done != 0; }
  __device__ void begin() {}
  __device__ void row(int i, bool valid, double (&part)[2], const WaveRows<T, B>& io) const {
    T rv[B];
    io.load(r, rv);
    if (mode != 1) {
      if (mode == 0) io.store(r0, rv);
      if constexpr (PREC) {
        if (valid) diag_apply<T, T, B>(D, n, i, rv);
      }
      io.store(rt, rv);
      return;
    }
    T wv[B];
    io.load(w0, wv);
    if (valid) {
#pragma unroll
      for (int k = 0; k < B; ++k) {
        part[0] += (double)rv[k] * (double)rv[k];
        part[1] += (double)rv[k] * (double)wv[k];
      }
      if constexpr (PREC) diag_apply<T, T, B>(D, n, i, wv);
    }
    io.store(wt, wv);
  }
  __device__ void finalize(const double (&tot)[2]) const {
    if (mode != 1) return;
    st->rho = tot[0];
    st->alpha = tot[0] / tot[1];
    st->beta = 0.0;
    st->omega = 0.0;
  }
};

__global__ void k_state_init(KState* st, double tol, int max_iter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    KState z;
    memset(&z, 0, sizeof(z));
    z.tol = tol;
    z.max_iter = max_iter;
    z.gamma = 1.0;  // PipeCG.f90 starts with gamma = alpha = 1
    z.alpha = 1.0;
    *st = z;
  }
}


template <typename T>
size_t cg_bytes(const Matrix& M, int nvec, int max_iter) {
  const size_t vec = (size_t)M.np * M.b;
  const size_t parts = partials_needed((long long)M.n) * 4 + partials_needed((long long)M.nslices * kWave) * 4;
  return nvec * (vec * sizeof(T) + 256) + parts * sizeof(double) + sizeof(KState) + 256 +
         ((size_t)max_iter + 2) * sizeof(double) + 4096;
}


template <typename T, int B, bool PREC>
int cg_impl(Matrix& M, const SolveParams& P, const T* b, T* x, Workspace& W, SolveResult* out, hipStream_t s) {
  Context* c = nullptr;
  MH_TRY(context(&c));
  const size_t vec = (size_t)M.np * B;
  MH_TRY(W.ensure(cg_bytes<T>(M, 5, P.max_iter)));
  Carve cv{W.raw.p};
  T* r = cv.take<T>(vec);
  T* z = cv.take<T>(vec);
  T* q = cv.take<T>(vec);
  T* p = cv.take<T>(vec);
  const size_t pn = partials_needed((long long)M.n);
  const size_t nparts = pn * 4 + partials_needed((long long)M.nslices * kWave) * 4;
  double* parts = cv.take<double>(nparts);
  double* parts_u = parts;           // update: (r, r), (r, z) per tile
  double* parts_v = parts + 2 * pn;  // variable-block preconditioner: (r, z) per tile
  double* parts_s = parts + 4 * pn;  // SpMV: (p, q) per workgroup
  KState* st = cv.take<KState>(1);
  double* hist = cv.take<double>((size_t)P.max_iter + 2);
  double* hp = P.keep_hist ? hist : nullptr;
  const T* D = nullptr;
  if constexpr (PREC) D = (sizeof(T) == 8) ? (const T*)M.D64.p : (const T*)M.D32.p;
  const int n = M.n;
  const size_t m = (size_t)n * B;
  const int var = sizeof(T) == 8 ? M.variant64 : M.variant32;
  int nwg_s = 0, spw = 0;
  if (M.nslices > 0) sell_geometry(M, var, &nwg_s, &spw);
  const int ntile_u = rows_grid<T, B>(n);  // partial sets of the update

  k_state_init<<<1, 64, 0, s>>>(st, P.tol, P.max_iter);
  MH_HIP(hipGetLastError());
  if (P.init_x) MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
  MH_HIP(hipMemsetAsync(p, 0, sizeof(T) * vec, s));
  if (vec > m) {  // halo part of z stays defined for gathers
    MH_HIP(hipMemsetAsync(z + m, 0, sizeof(T) * (vec - m), s));
  }
  // R = B - A X; B2 = (R, R)   (CG.f90:61-63, mat_converge.f90:38)
  {
    OpResidSolve<T, B> op{x, b, r, nullptr, &st->B2, st};
    MH_TRY((launch_sell<T, B>(M, op, parts_s, var, s)));
  }
  double B2 = 0.0;
  MH_TRY(read_B2(P, st, &B2, s));
  out->hist.clear();
  if (B2 == 0.0) {  // mat_converge.f90:40-48: X = 0, converged, no iteration
    MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
    MH_HIP(hipStreamSynchronize(s));
    out->zero_rhs = 1;
    out->converged = 1;
    out->iter = -1;
    out->resid = 0.0;
    return 0;
  }
  // Three launches per iteration and no finalize kernels: the p step's
  // prologue runs the previous iteration's test and derives rho and beta from
  // the update's partials; the update's prologue derives alpha from the
  // SpMV's partials.  Variable block sizes: diag_V runs after the update as
  // its own kernel and leaves (r, z) in parts_v.
  const bool vprec = PREC && B == 1 && M.vb != nullptr;
  const int ntile_v = vprec ? rows_grid<T, 1>(M.vb->nb) : 0;
  auto upd = [&](int it, int flags) -> int {
    const double* pre = (flags & (F_UPD | F_XONLY)) ? parts_s : nullptr;
    if constexpr (PREC && B == 1) {
      if (vprec) {
        OpCgUpdate<T, B, false> op{x, r, z, p, q, nullptr, n, st, it, flags | F_EXTPREC, pre, nwg_s, (T)0};
        MH_TRY((launch_rows<decltype(op), false>(n, op, parts_u, s)));
        if (flags & F_XONLY) return 0;
        const VBlock& V = *M.vb;
        const T* DV = sizeof(T) == 8 ? (const T*)V.D64.p : (const T*)V.D32.p;
        OpVPrec<T, true> vp{r, z, V.list.p, V.dof_index.p, V.doff.p, DV, st, 0};
        return launch_rows<decltype(vp), false>(V.nb, vp, parts_v, s);
      }
    }
    OpCgUpdate<T, B, PREC> op{x, r, z, p, q, D, n, st, it, flags, pre, nwg_s, (T)0};
    return launch_rows<decltype(op), false>(n, op, parts_u, s);
  };
  auto p_op = [&](int it) {
    return OpCgP<T>{z, p, st, hp, parts_u, ntile_u, vprec ? parts_v : nullptr, ntile_v, it, (T)0};
  };
  MH_TRY(upd(0, 0));  // z = M^-1 r and the partials of the first rho
  Poller poll{c, &st->resid, P.tol};
  int done = 0;
  int it = 1;
  while (it <= P.max_iter && !done) {
    const int end = std::min(it + poll.next_chunk(it, P.chunk) - 1, P.max_iter);
    for (; it <= end; ++it) {
      MH_TRY(launch_rows((int)m, p_op(it), nullptr, s));
      OpCgSpmv<T, B> sp{p, q, st};
      MH_TRY((launch_sell<T, B, OpCgSpmv<T, B>, false>(M, sp, parts_s, var, s)));
      if (it % kIterRR == 0) {  // true residual (CG.f90:91-93)
        MH_TRY(upd(it, F_XONLY));
        OpResidSolve<T, B> rs{x, b, r, nullptr, nullptr, st};
        MH_TRY((launch_sell<T, B, OpResidSolve<T, B>, false>(M, rs, parts_s, var, s)));
        MH_TRY(upd(it, 0));
      } else {
        MH_TRY(upd(it, F_UPD));
      }
    }
    MH_TRY(poll.issue(s));
    MH_TRY(poll.check(1, &done));
  }
  // test of the last enqueued iteration (a no-op once the solve is done)
  k_prologue_only<row_block<T, 1>()><<<1, row_block<T, 1>(), 0, s>>>(p_op(it));
  MH_HIP(hipGetLastError());
  ktimer_mark(kClsDot);
  MH_HIP(hipStreamSynchronize(s));
  KState hs;
  MH_HIP(hipMemcpy(&hs, st, sizeof(KState), hipMemcpyDeviceToHost));
  out->iter = hs.iter;
  out->resid = hs.resid;
  out->converged = hs.converged;
  out->breakdown = hs.breakdown;
  if (P.keep_hist && hs.iter > 0) {
    out->hist.resize((size_t)hs.iter + 1);
    MH_HIP(hipMemcpy(out->hist.data(), hist, sizeof(double) * ((size_t)hs.iter + 1), hipMemcpyDeviceToHost));
  }
  return 0;
}

// GropCG (monolis method 2, src/iterative/GropCG.f90): two fused launches per
// iteration -- the x / r / u update with (r, u), (r, r) and the test; the SpMV
// A u with the p, s and q = M^-1 s updates and (p, s) in its epilogue.  No
// residual replacement (the reference has none for this method).
template <typename T, int B, bool PREC>
int gropcg_impl(Matrix& M, const SolveParams& P, const T* b, T* x, Workspace& W, SolveResult* out,
                hipStream_t s) {
  Context* c = nullptr;
  MH_TRY(context(&c));
  const size_t vec = (size_t)M.np * B;
  MH_TRY(W.ensure(cg_bytes<T>(M, 5, P.max_iter)));
  Carve cv{W.raw.p};
  T* r = cv.take<T>(vec);
  T* u = cv.take<T>(vec);
  T* q = cv.take<T>(vec);
  T* p = cv.take<T>(vec);
  T* sv = cv.take<T>(vec);
  const size_t nparts = partials_needed((long long)M.n) * 4 + partials_needed((long long)M.nslices * kWave) * 4;
  double* parts = cv.take<double>(nparts);
  KState* st = cv.take<KState>(1);
  double* hist = cv.take<double>((size_t)P.max_iter + 2);
  const T* D = nullptr;
  if constexpr (PREC) D = (sizeof(T) == 8) ? (const T*)M.D64.p : (const T*)M.D32.p;
  const int n = M.n;
  const size_t m = (size_t)n * B;
  const int var = sizeof(T) == 8 ? M.variant64 : M.variant32;
  double* hp = P.keep_hist ? hist : nullptr;
  const bool vprec = PREC && B == 1 && M.vb != nullptr;
  auto vop = [&](const T* src, T* dst, bool dot) -> int {
    if constexpr (PREC && B == 1) {
      const VBlock& V = *M.vb;
      const T* DV = sizeof(T) == 8 ? (const T*)V.D64.p : (const T*)V.D32.p;
      if (dot) {
        OpVPrec<T, true> vp{src, dst, V.list.p, V.dof_index.p, V.doff.p, DV, st, 2};
        return launch_rows(V.nb, vp, parts, s);
      }
      OpVPrec<T, false> vp{src, dst, V.list.p, V.dof_index.p, V.doff.p, DV, st, 0};
      return launch_rows(V.nb, vp, nullptr, s);
    }
    (void)src;
    (void)dst;
    (void)dot;
    return 0;
  };
  auto spmv_step = [&]() -> int {
    if (vprec) {
      OpGropSpmv<T, B, false> op{u, p, sv, q, nullptr, n, st, (T)0};
      MH_TRY((launch_sell<T, B>(M, op, parts, var, s)));
      return vop(sv, q, false);
    }
    OpGropSpmv<T, B, PREC> op{u, p, sv, q, D, n, st, (T)0};
    return launch_sell<T, B>(M, op, parts, var, s);
  };

  k_state_init<<<1, 64, 0, s>>>(st, P.tol, P.max_iter);
  MH_HIP(hipGetLastError());
  if (P.init_x) MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
  for (T* w : {u, q, p, sv}) MH_HIP(hipMemsetAsync(w, 0, sizeof(T) * vec, s));
  {
    OpResidSolve<T, B> op{x, b, r, nullptr, &st->B2, st};
    MH_TRY((launch_sell<T, B>(M, op, parts, var, s)));
  }
  double B2 = 0.0;
  MH_TRY(read_B2(P, st, &B2, s));
  out->hist.clear();
  if (B2 == 0.0) {  // mat_converge.f90:40-48
    MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
    MH_HIP(hipStreamSynchronize(s));
    out->zero_rhs = 1;
    out->converged = 1;
    out->iter = -1;
    out->resid = 0.0;
    return 0;
  }
  // U = M^-1 R, gamma = (R, U); P = U, S = A P, Q = M^-1 S, alpha = gamma / (P, S)
  if (vprec) {
    MH_TRY(vop(r, u, true));
  } else {
    OpGropInit<T, B, PREC> op{r, u, D, n, st};
    MH_TRY(launch_rows(n, op, parts, s));
  }
  MH_TRY(spmv_step());  // beta = 0 from the state init, p = s = 0
  Poller poll{c, &st->resid, P.tol};
  int done = 0;
  for (int it = 1; it <= P.max_iter && !done;) {
    const int end = std::min(it + poll.next_chunk(it, P.chunk) - 1, P.max_iter);
    for (; it <= end; ++it) {
      OpGropUpd<T, B> up{x, r, u, p, sv, q, st, hp, it, (T)0};
      MH_TRY(launch_rows(n, up, parts, s));
      MH_TRY(spmv_step());
    }
    MH_TRY(poll.issue(s));
    MH_TRY(poll.check(1, &done));
  }
  MH_HIP(hipStreamSynchronize(s));
  KState hs;
  MH_HIP(hipMemcpy(&hs, st, sizeof(KState), hipMemcpyDeviceToHost));
  out->iter = hs.iter;
  out->resid = hs.resid;
  out->converged = hs.converged;
  if (P.keep_hist && hs.iter > 0) {
    out->hist.resize((size_t)hs.iter + 1);
    MH_HIP(hipMemcpy(out->hist.data(), hist, sizeof(double) * ((size_t)hs.iter + 1), hipMemcpyDeviceToHost));
  }
  return 0;
}

// PipeCG (monolis method 3, src/iterative/PipeCG.f90): three fused launches per
// iteration -- (r, u), (v, u) and m = M^-1 v; l = A m; the eight vector
// recurrences with (r, r) and the convergence test -- and the true residual
// every 50 iterations (PipeCG.f90 iter_RR = 50).
template <typename T, int B, bool PREC>
int pipecg_impl(Matrix& M, const SolveParams& P, const T* b, T* x, Workspace& W, SolveResult* out,
                hipStream_t s) {
  constexpr int kPipeRR = 50;
  Context* c = nullptr;
  MH_TRY(context(&c));
  const size_t vec = (size_t)M.np * B;
  MH_TRY(W.ensure(cg_bytes<T>(M, 9, P.max_iter)));
  Carve cv{W.raw.p};
  T* r = cv.take<T>(vec);
  T* u = cv.take<T>(vec);
  T* v = cv.take<T>(vec);
  T* q = cv.take<T>(vec);
  T* p = cv.take<T>(vec);
  T* z = cv.take<T>(vec);
  T* l = cv.take<T>(vec);
  T* mm = cv.take<T>(vec);
  T* sv = cv.take<T>(vec);
  const size_t nparts = partials_needed((long long)M.n) * 4 + partials_needed((long long)M.nslices * kWave) * 4;
  double* parts = cv.take<double>(nparts);
  KState* st = cv.take<KState>(1);
  double* hist = cv.take<double>((size_t)P.max_iter + 2);
  const T* D = nullptr;
  if constexpr (PREC) D = (sizeof(T) == 8) ? (const T*)M.D64.p : (const T*)M.D32.p;
  const int n = M.n;
  const size_t m = (size_t)n * B;
  const int var = sizeof(T) == 8 ? M.variant64 : M.variant32;
  double* hp = P.keep_hist ? hist : nullptr;
  const bool vprec = PREC && B == 1 && M.vb != nullptr;
  auto vapply = [&](const T* src, T* dst) -> int {
    if constexpr (PREC && B == 1) {
      const VBlock& V = *M.vb;
      const T* DV = sizeof(T) == 8 ? (const T*)V.D64.p : (const T*)V.D32.p;
      OpVPrec<T, false> vp{src, dst, V.list.p, V.dof_index.p, V.doff.p, DV, st, 0};
      return launch_rows(V.nb, vp, nullptr, s);
    }
    (void)src;
    (void)dst;
    return 0;
  };
  // u = M^-1 r (+ test when flags has F_CHECK)
  auto prec_check = [&](int it, int flags) -> int {
    if (vprec) {
      OpPrecCheck<T, B, false> op{r, u, nullptr, n, st, hp, it, flags};
      MH_TRY(launch_rows(n, op, parts, s));
      return vapply(r, u);
    }
    OpPrecCheck<T, B, PREC> op{r, u, D, n, st, hp, it, flags};
    return launch_rows(n, op, parts, s);
  };

  k_state_init<<<1, 64, 0, s>>>(st, P.tol, P.max_iter);
  MH_HIP(hipGetLastError());
  if (P.init_x) MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
  for (T* w : {q, p, z, sv, l, mm, u, v}) MH_HIP(hipMemsetAsync(w, 0, sizeof(T) * vec, s));
  {
    OpResidSolve<T, B> op{x, b, r, nullptr, &st->B2, st};
    MH_TRY((launch_sell<T, B>(M, op, parts, var, s)));
  }
  double B2 = 0.0;
  MH_TRY(read_B2(P, st, &B2, s));
  out->hist.clear();
  if (B2 == 0.0) {  // mat_converge.f90:40-48
    MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
    MH_HIP(hipStreamSynchronize(s));
    out->zero_rhs = 1;
    out->converged = 1;
    out->iter = -1;
    out->resid = 0.0;
    return 0;
  }
  // U = M^-1 R; V = A U
  MH_TRY(prec_check(0, 0));
  {
    OpSpmvSolve<T, B> sp{u, v, st};
    MH_TRY((launch_sell<T, B>(M, sp, parts, var, s)));
  }
  Poller poll{c, &st->resid, P.tol};
  int done = 0;
  for (int it = 1; it <= P.max_iter && !done;) {
    const int end = std::min(it + poll.next_chunk(it, P.chunk) - 1, P.max_iter);
    for (; it <= end; ++it) {
      if (vprec) {
        OpPipeA<T, B, false> a{r, u, v, mm, nullptr, n, st, it == 1};
        MH_TRY(launch_rows(n, a, parts, s));
        MH_TRY(vapply(v, mm));
      } else {
        OpPipeA<T, B, PREC> a{r, u, v, mm, D, n, st, it == 1};
        MH_TRY(launch_rows(n, a, parts, s));
      }
      OpSpmvSolve<T, B> sl{mm, l, st};
      MH_TRY((launch_sell<T, B>(M, sl, parts, var, s)));
      const bool rr = it % kPipeRR == 0;
      OpPipeB<T, B> ub{z, q, sv, p, r, u, v, x, l, mm, st, hp, it, rr ? F_XONLY : 0, (T)0, (T)0};
      MH_TRY(launch_rows(n, ub, parts, s));
      if (rr) {
        OpResidSolve<T, B> rs{x, b, r, nullptr, nullptr, st};
        MH_TRY((launch_sell<T, B>(M, rs, parts, var, s)));
        MH_TRY(prec_check(it, F_CHECK));
        OpSpmvSolve<T, B> sp{u, v, st};
        MH_TRY((launch_sell<T, B>(M, sp, parts, var, s)));
      }
    }
    MH_TRY(poll.issue(s));
    MH_TRY(poll.check(1, &done));
  }
  MH_HIP(hipStreamSynchronize(s));
  KState hs;
  MH_HIP(hipMemcpy(&hs, st, sizeof(KState), hipMemcpyDeviceToHost));
  out->iter = hs.iter;
  out->resid = hs.resid;
  out->converged = hs.converged;
  if (P.keep_hist && hs.iter > 0) {
    out->hist.resize((size_t)hs.iter + 1);
    MH_HIP(hipMemcpy(out->hist.data(), hist, sizeof(double) * ((size_t)hs.iter + 1), hipMemcpyDeviceToHost));
  }
  return 0;
}

// PipeCR (monolis method 4, src/iterative/PipeCR.f90): three fused launches per
// iteration -- m = M^-1 v with (v, u), (v, m), (u, u) and the test on the
// preconditioned residual; l = A m; the five vector recurrences -- and the
// true residual every 50 iterations (PipeCR.f90 iter_RR = 50).
template <typename T, int B, bool PREC>
int pipecr_impl(Matrix& M, const SolveParams& P, const T* b, T* x, Workspace& W, SolveResult* out,
                hipStream_t s) {
  constexpr int kPipeRR = 50;
  Context* c = nullptr;
  MH_TRY(context(&c));
  const size_t vec = (size_t)M.np * B;
  MH_TRY(W.ensure(cg_bytes<T>(M, 8, P.max_iter)));
  Carve cv{W.raw.p};
  T* r = cv.take<T>(vec);
  T* u = cv.take<T>(vec);
  T* v = cv.take<T>(vec);
  T* q = cv.take<T>(vec);
  T* p = cv.take<T>(vec);
  T* z = cv.take<T>(vec);
  T* l = cv.take<T>(vec);
  T* mm = cv.take<T>(vec);
  const size_t nparts = partials_needed((long long)M.n) * 4 + partials_needed((long long)M.nslices * kWave) * 4;
  double* parts = cv.take<double>(nparts);
  KState* st = cv.take<KState>(1);
  double* hist = cv.take<double>((size_t)P.max_iter + 2);
  const T* D = nullptr;
  if constexpr (PREC) D = (sizeof(T) == 8) ? (const T*)M.D64.p : (const T*)M.D32.p;
  const int n = M.n;
  const size_t m = (size_t)n * B;
  const int var = sizeof(T) == 8 ? M.variant64 : M.variant32;
  double* hp = P.keep_hist ? hist : nullptr;
  const bool vprec = PREC && B == 1 && M.vb != nullptr;
  auto vapply = [&](const T* src, T* dst) -> int {
    if constexpr (PREC && B == 1) {
      const VBlock& V = *M.vb;
      const T* DV = sizeof(T) == 8 ? (const T*)V.D64.p : (const T*)V.D32.p;
      OpVPrec<T, false> vp{src, dst, V.list.p, V.dof_index.p, V.doff.p, DV, st, 0};
      return launch_rows(V.nb, vp, nullptr, s);
    }
    (void)src;
    (void)dst;
    return 0;
  };
  // u = M^-1 r
  auto prec = [&]() -> int {
    if (vprec) {
      OpPrecCheck<T, B, false> op{r, u, nullptr, n, st, nullptr, 0, 0};
      MH_TRY(launch_rows(n, op, parts, s));
      return vapply(r, u);
    }
    OpPrecCheck<T, B, PREC> op{r, u, D, n, st, nullptr, 0, 0};
    return launch_rows(n, op, parts, s);
  };

  k_state_init<<<1, 64, 0, s>>>(st, P.tol, P.max_iter);
  MH_HIP(hipGetLastError());
  if (P.init_x) MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
  for (T* w : {q, p, z, l, mm, u, v}) MH_HIP(hipMemsetAsync(w, 0, sizeof(T) * vec, s));
  {
    OpResidSolve<T, B> op{x, b, r, nullptr, &st->B2, st};
    MH_TRY((launch_sell<T, B>(M, op, parts, var, s)));
  }
  double B2 = 0.0;
  MH_TRY(read_B2(P, st, &B2, s));
  out->hist.clear();
  if (B2 == 0.0) {  // as mat_converge.f90:40-48 for the other methods (PipeCR.f90 itself divides 0/0)
    MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
    MH_HIP(hipStreamSynchronize(s));
    out->zero_rhs = 1;
    out->converged = 1;
    out->iter = -1;
    out->resid = 0.0;
    return 0;
  }
  // U = M^-1 R; phi = sqrt(R2/U2); V = A U
  MH_TRY(prec());
  {
    OpCrPhi<T, B> ph{u, st};
    MH_TRY(launch_rows(n, ph, parts, s));
    OpSpmvSolve<T, B> sp{u, v, st};
    MH_TRY((launch_sell<T, B>(M, sp, parts, var, s)));
  }
  Poller poll{c, &st->resid, P.tol};
  int done = 0;
  for (int it = 1; it <= P.max_iter && !done;) {
    const int end = std::min(it + poll.next_chunk(it, P.chunk) - 1, P.max_iter);
    for (; it <= end; ++it) {
      if (vprec) {
        MH_TRY(vapply(v, mm));
        OpCrA<T, B, false, true> a{u, v, mm, nullptr, n, st, hp, it};
        MH_TRY(launch_rows(n, a, parts, s));
      } else {
        OpCrA<T, B, PREC, false> a{u, v, mm, D, n, st, hp, it};
        MH_TRY(launch_rows(n, a, parts, s));
      }
      OpSpmvSolve<T, B> sl{mm, l, st};
      MH_TRY((launch_sell<T, B>(M, sl, parts, var, s)));
      const bool rr = it % kPipeRR == 0;
      OpCrB<T, B> ub{z, q, p, u, v, x, l, mm, st, rr ? F_XONLY : 0, (T)0, (T)0};
      MH_TRY(launch_rows(n, ub, parts, s));
      if (rr) {
        OpResidSolve<T, B> rs{x, b, r, nullptr, nullptr, st};
        MH_TRY((launch_sell<T, B>(M, rs, parts, var, s)));
        MH_TRY(prec());
        OpSpmvSolve<T, B> sp{u, v, st};
        MH_TRY((launch_sell<T, B>(M, sp, parts, var, s)));
      }
    }
    MH_TRY(poll.issue(s));
    MH_TRY(poll.check(1, &done));
  }
  MH_HIP(hipStreamSynchronize(s));
  KState hs;
  MH_HIP(hipMemcpy(&hs, st, sizeof(KState), hipMemcpyDeviceToHost));
  out->iter = hs.iter;
  out->resid = hs.resid;
  out->converged = hs.converged;
  if (P.keep_hist && hs.iter > 0) {
    out->hist.resize((size_t)hs.iter + 1);
    MH_HIP(hipMemcpy(out->hist.data(), hist, sizeof(double) * ((size_t)hs.iter + 1), hipMemcpyDeviceToHost));
  }
  return 0;
}

// PipeBiCGSTAB (monolis method 6; method 8 = no preconditioner): per iteration
// the recurrence pass with (Q, Y), (Y, Y) and ZT = M^-1 Z; V = A ZT; the update
// pass with five dots, WT = M^-1 W0 and the test; T = A WT.  True residual
// every 200 iterations (PipeBiCGSTAB.f90 iter_RR).
template <typename T, int B, bool PREC>
int pipebicgstab_impl(Matrix& M, const SolveParams& P, const T* b, T* x, Workspace& W, SolveResult* out,
                      hipStream_t s) {
  constexpr int kPbRR = 200;
  Context* c = nullptr;
  MH_TRY(context(&c));
  const size_t vec = (size_t)M.np * B;
  MH_TRY(W.ensure(cg_bytes<T>(M, 15, P.max_iter)));
  Carve cv{W.raw.p};
  T *r = cv.take<T>(vec), *r0 = cv.take<T>(vec), *rt = cv.take<T>(vec), *w0 = cv.take<T>(vec);
  T *wt = cv.take<T>(vec), *t = cv.take<T>(vec), *pt = cv.take<T>(vec), *sv = cv.take<T>(vec);
  T *stv = cv.take<T>(vec), *z = cv.take<T>(vec), *zt = cv.take<T>(vec), *q = cv.take<T>(vec);
  T *qt = cv.take<T>(vec), *y = cv.take<T>(vec), *v = cv.take<T>(vec);
  const size_t nparts = partials_needed((long long)M.n) * 4 + partials_needed((long long)M.nslices * kWave) * 4;
  double* parts = cv.take<double>(nparts);
  KState* st = cv.take<KState>(1);
  double* hist = cv.take<double>((size_t)P.max_iter + 2);
  const T* D = nullptr;
  if constexpr (PREC) D = (sizeof(T) == 8) ? (const T*)M.D64.p : (const T*)M.D32.p;
  const int n = M.n;
  const size_t m = (size_t)n * B;
  const int var = sizeof(T) == 8 ? M.variant64 : M.variant32;
  double* hp = P.keep_hist ? hist : nullptr;
  const bool vprec = PREC && B == 1 && M.vb != nullptr;
  auto vapply = [&](const T* src, T* dst) -> int {
    if constexpr (PREC && B == 1) {
      if (vprec) {
        const VBlock& V = *M.vb;
        const T* DV = sizeof(T) == 8 ? (const T*)V.D64.p : (const T*)V.D32.p;
        OpVPrec<T, false> vp{src, dst, V.list.p, V.dof_index.p, V.doff.p, DV, st, 0};
        return launch_rows(V.nb, vp, nullptr, s);
      }
    }
    (void)src;
    (void)dst;
    return 0;
  };
  // the fused block-Jacobi runs only for fixed blocks; variable blocks take vapply
  auto rows_init = [&](int mode) -> int {
    if (vprec) {
      OpPbInit<T, B, false> op{r, r0, rt, w0, wt, nullptr, n, st, mode};
      MH_TRY(launch_rows(n, op, parts, s));
      return mode == 1 ? vapply(w0, wt) : vapply(r, rt);
    }
    OpPbInit<T, B, PREC> op{r, r0, rt, w0, wt, D, n, st, mode};
    return launch_rows(n, op, parts, s);
  };
  auto rows_b = [&](int it, int mode) -> int {
    if (vprec) {
      OpPbB<T, B, false> op{x, r, rt, w0, wt, pt, qt, q, y, zt, t, v, r0, sv, z, nullptr, n, st, hp, it, mode,
                            (T)0, (T)0};
      MH_TRY(launch_rows(n, op, parts, s));
      return mode == 1 ? 0 : vapply(w0, wt);
    }
    OpPbB<T, B, PREC> op{x, r, rt, w0, wt, pt, qt, q, y, zt, t, v, r0, sv, z, D, n, st, hp, it, mode, (T)0, (T)0};
    return launch_rows(n, op, parts, s);
  };
  auto spmv = [&](const T* in, T* outv) -> int {
    OpSpmvSolve<T, B> op{in, outv, st};
    return launch_sell<T, B>(M, op, parts, var, s);
  };

  k_state_init<<<1, 64, 0, s>>>(st, P.tol, P.max_iter);
  MH_HIP(hipGetLastError());
  if (P.init_x) MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
  for (T* w : {r0, rt, w0, wt, t, pt, sv, stv, z, zt, q, qt, y, v}) MH_HIP(hipMemsetAsync(w, 0, sizeof(T) * vec, s));
  {
    OpResidSolve<T, B> op{x, b, r, nullptr, &st->B2, st};
    MH_TRY((launch_sell<T, B>(M, op, parts, var, s)));
  }
  double B2 = 0.0;
  MH_TRY(read_B2(P, st, &B2, s));
  out->hist.clear();
  if (B2 == 0.0) {  // mat_converge.f90:40-48
    MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
    MH_HIP(hipStreamSynchronize(s));
    out->zero_rhs = 1;
    out->converged = 1;
    out->iter = -1;
    out->resid = 0.0;
    return 0;
  }
  MH_TRY(rows_init(0));     // R0 = R, RT = M^-1 R
  MH_TRY(spmv(rt, w0));     // W0 = A RT
  MH_TRY(rows_init(1));     // WT = M^-1 W0, RR, RW, alpha
  MH_TRY(spmv(wt, t));      // T = A WT
  Poller poll{c, &st->resid, P.tol};
  int done = 0;
  for (int it = 1; it <= P.max_iter && !done;) {
    const int end = std::min(it + poll.next_chunk(it, P.chunk) - 1, P.max_iter);
    for (; it <= end; ++it) {
      if (vprec) {
        OpPbA<T, B, false> a{pt, sv, stv, z, zt, q, qt, y, rt, w0, wt, t, v, r, nullptr, n, st, (T)0, (T)0, (T)0};
        MH_TRY(launch_rows(n, a, parts, s));
        MH_TRY(vapply(z, zt));
      } else {
        OpPbA<T, B, PREC> a{pt, sv, stv, z, zt, q, qt, y, rt, w0, wt, t, v, r, D, n, st, (T)0, (T)0, (T)0};
        MH_TRY(launch_rows(n, a, parts, s));
      }
      MH_TRY(spmv(zt, v));  // V = A ZT
      if (it % kPbRR == 0) {
        MH_TRY(rows_b(it, 1));  // X += alpha PT + omega1 T
        OpResidSolve<T, B> rs{x, b, r, nullptr, nullptr, st};
        MH_TRY((launch_sell<T, B>(M, rs, parts, var, s)));
        MH_TRY(rows_init(2));  // RT = M^-1 R (mode 2 = mode 0 without touching R0)
        MH_TRY(spmv(rt, w0));
        MH_TRY(rows_b(it, 2));  // dots, WT, test, scalars
      } else {
        MH_TRY(rows_b(it, 0));
      }
      MH_TRY(spmv(wt, t));  // T = A WT
    }
    MH_TRY(poll.issue(s));
    MH_TRY(poll.check(1, &done));
  }
  MH_HIP(hipStreamSynchronize(s));
  KState hs;
  MH_HIP(hipMemcpy(&hs, st, sizeof(KState), hipMemcpyDeviceToHost));
  out->iter = hs.iter;
  out->resid = hs.resid;
  out->converged = hs.converged;
  if (P.keep_hist && hs.iter > 0) {
    out->hist.resize((size_t)hs.iter + 1);
    MH_HIP(hipMemcpy(out->hist.data(), hist, sizeof(double) * ((size_t)hs.iter + 1), hipMemcpyDeviceToHost));
  }
  return 0;
}

// rr_period: true-residual replacement period (200 for BiCGSTAB.f90, 50 for
// BiCGSTAB_noprec.f90, which also runs without a preconditioner).
template <typename T, int B, bool PREC>
int bicgstab_impl(Matrix& M, const SolveParams& P, const T* b, T* x, Workspace& W, SolveResult* out,
                  hipStream_t s, int rr_period) {
  Context* c = nullptr;
  MH_TRY(context(&c));
  const size_t vec = (size_t)M.np * B;
  MH_TRY(W.ensure(cg_bytes<T>(M, 8, P.max_iter)));
  Carve cv{W.raw.p};
  T* r = cv.take<T>(vec);
  T* rt = cv.take<T>(vec);
  T* p = cv.take<T>(vec);
  T* pt = cv.take<T>(vec);
  T* sv = cv.take<T>(vec);
  T* stv = cv.take<T>(vec);
  T* t = cv.take<T>(vec);
  T* v = cv.take<T>(vec);
  const size_t pn = partials_needed((long long)M.n);
  const size_t pn_s = partials_needed((long long)M.nslices * kWave);
  const size_t nparts = pn * 4 + pn_s * 4;
  double* parts = cv.take<double>(nparts);
  double* parts_e = parts;               // e: (r, r), (r, rt) per tile
  double* parts_b = parts + 4 * pn;      // b: (rt, v) per workgroup
  double* parts_d = parts_b + pn_s;      // d: (t, s), (t, t) per workgroup
  double* dsum = parts_d + 2 * pn_s;     // sorted-sum result
  KState* st = cv.take<KState>(1);
  double* hist = cv.take<double>((size_t)P.max_iter + 2);
  const T* D = nullptr;
  if constexpr (PREC) D = (sizeof(T) == 8) ? (const T*)M.D64.p : (const T*)M.D32.p;
  const int n = M.n;
  const size_t m = (size_t)n * B;
  const int var = sizeof(T) == 8 ? M.variant64 : M.variant32;
  int nwg_s = 0, spw = 0;
  if (M.nslices > 0) sell_geometry(M, var, &nwg_s, &spw);
  const int ntile_u = rows_grid<T, B>(n);  // partial sets of the update

  k_state_init<<<1, 64, 0, s>>>(st, P.tol, P.max_iter);
  MH_HIP(hipGetLastError());
  if (P.init_x) MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
  if (vec > m) {
    MH_HIP(hipMemsetAsync(pt + m, 0, sizeof(T) * (vec - m), s));
    MH_HIP(hipMemsetAsync(stv + m, 0, sizeof(T) * (vec - m), s));
  }
  MH_HIP(hipMemsetAsync(v, 0, sizeof(T) * vec, s));
  // R = B - A X, RT = R, B2 = (R, R); rho_1 = (R, RT) = B2
  {
    OpResidSolve<T, B> op{x, b, r, rt, &st->B2, st};
    MH_TRY((launch_sell<T, B>(M, op, parts_b, var, s)));
  }
  double B2 = 0.0;
  MH_TRY(read_B2(P, st, &B2, s));
  out->hist.clear();
  if (B2 == 0.0) {
    MH_HIP(hipMemsetAsync(x, 0, sizeof(T) * m, s));
    MH_HIP(hipStreamSynchronize(s));
    out->zero_rhs = 1;
    out->converged = 1;
    out->iter = -1;
    out->resid = 0.0;
    return 0;
  }
  // rho of iteration 1 is (R, RT) with RT == R: the same sum as B2.
  MH_HIP(hipMemcpyAsync(&st->rho, &st->B2, sizeof(double), hipMemcpyDeviceToDevice, s));
  MH_HIP(hipMemcpyAsync(&st->rhok[1], &st->B2, sizeof(double), hipMemcpyDeviceToDevice, s));
  // variable block sizes: diag_V runs as its own kernel after steps A and C
  const bool vprec = PREC && B == 1 && M.vb != nullptr;
  auto vapply = [&](const T* src, T* dst) -> int {
    if constexpr (PREC && B == 1) {
      const VBlock& V = *M.vb;
      const T* DV = sizeof(T) == 8 ? (const T*)V.D64.p : (const T*)V.D32.p;
      OpVPrec<T, false> vp{src, dst, V.list.p, V.dof_index.p, V.doff.p, DV, st, 0};
      return launch_rows(V.nb, vp, nullptr, s);
    }
    (void)src;
    (void)dst;
    return 0;
  };
  double* hp = P.keep_hist ? hist : nullptr;
  auto a_op = [&](int it, bool with_pre) {
    return OpBiA<T, B, PREC>{r, p, v, pt, D, n, st, hp, with_pre ? parts_e : nullptr, ntile_u, it, (T)0, (T)0};
  };
  // Stage cursor: an iteration is A (test of the previous iteration, rho,
  // beta; p, pt), B (v, (rt, v)), C (alpha; s, st), D (t, (t, s), (t, t)),
  // E (omega; x, r, (r, r), (r, rt)).  A breakdown stops the device loop in
  // the prologue of A, C or E; the host then recomputes that dot with the
  // sorted sum and resumes at that kernel without its prologue.
  int it = 1, stage = 0;
  int resumed = -1;  // stage whose prologue the next pass skips
  KState hs;
  out->fallbacks = 0;
  for (;;) {
    Poller poll{c, &st->resid, P.tol};
    int done = 0;
    while (it <= P.max_iter && !done) {
      const int end = std::min(it + poll.next_chunk(it, P.chunk) - 1, P.max_iter);
      for (; it <= end; ++it, stage = 0, resumed = -1) {
        if (stage <= 0) {
          const bool pre_a = it > 1 && resumed != 0;
          if (vprec) {
            OpBiA<T, B, false> a{r, p, v, pt, nullptr, n, st, hp, pre_a ? parts_e : nullptr, ntile_u, it, (T)0, (T)0};
            MH_TRY(launch_rows(n, a, nullptr, s));
            MH_TRY(vapply(p, pt));
          } else {
            MH_TRY(launch_rows(n, a_op(it, pre_a), nullptr, s));
          }
          OpBiB<T, B> bb{pt, v, rt, st};
          MH_TRY((launch_sell<T, B, OpBiB<T, B>, false>(M, bb, parts_b, var, s)));
        }
        if (stage <= 2) {
          const double* pre_c = resumed == 2 ? nullptr : parts_b;
          if (vprec) {
            OpBiC<T, B, false> cc{r, v, sv, stv, nullptr, n, st, pre_c, nwg_s, it, (T)0};
            MH_TRY(launch_rows(n, cc, nullptr, s));
            MH_TRY(vapply(sv, stv));
          } else {
            OpBiC<T, B, PREC> cc{r, v, sv, stv, D, n, st, pre_c, nwg_s, it, (T)0};
            MH_TRY(launch_rows(n, cc, nullptr, s));
          }
          OpBiD<T, B> dd{stv, t, sv, st};
          MH_TRY((launch_sell<T, B, OpBiD<T, B>, false>(M, dd, parts_d, var, s)));
        }
        const double* pre_e = resumed == 4 ? nullptr : parts_d;
        if (it % rr_period == 0) {  // true residual (BiCGSTAB.f90:163-165)
          OpBiE<T, B> e1{x, r, pt, stv, sv, t, rt, st, it, F_XONLY, pre_e, nwg_s, (T)0, (T)0};
          MH_TRY((launch_rows<OpBiE<T, B>, false>(n, e1, parts_e, s)));
          OpResidSolve<T, B> rs{x, b, r, nullptr, nullptr, st};
          MH_TRY((launch_sell<T, B, OpResidSolve<T, B>, false>(M, rs, parts_b, var, s)));
          OpBiE<T, B> e2{x, r, pt, stv, sv, t, rt, st, it, 0, nullptr, 0, (T)0, (T)0};
          MH_TRY((launch_rows<OpBiE<T, B>, false>(n, e2, parts_e, s)));
        } else {
          OpBiE<T, B> e{x, r, pt, stv, sv, t, rt, st, it, F_UPD, pre_e, nwg_s, (T)0, (T)0};
          MH_TRY((launch_rows<OpBiE<T, B>, false>(n, e, parts_e, s)));
        }
      }
      MH_TRY(poll.issue(s));
      MH_TRY(poll.check(1, &done));
    }
    // test of the last enqueued iteration (a no-op once the solve is done)
    if (it > 1) {
      k_prologue_only<row_block<T, B>()><<<1, row_block<T, B>(), 0, s>>>(a_op(it, true));
      MH_HIP(hipGetLastError());
      ktimer_mark(kClsDot);
    }
    MH_HIP(hipStreamSynchronize(s));
    MH_HIP(hipMemcpy(&hs, st, sizeof(KState), hipMemcpyDeviceToHost));
    if (hs.bd_stage == 0 || hs.converged) break;
    // sorted-sum fallback (monolis_global_sorted_inner_product_main_R), then resume
    double val = 0.0;
    out->fallbacks++;
    if (hs.bd_stage == 2) {
      MH_TRY(sorted_dot<T>((long long)m, rt, v, dsum, s));
      MH_HIP(hipMemcpy(&val, dsum, sizeof(double), hipMemcpyDeviceToHost));
      hs.pq = val;
      hs.alpha = hs.rho / val;
      it = hs.bd_iter;
      stage = 2;
    } else if (hs.bd_stage == 4) {
      MH_TRY(sorted_dot<T>((long long)m, t, t, dsum, s));
      MH_HIP(hipMemcpy(&val, dsum, sizeof(double), hipMemcpyDeviceToHost));
      hs.tt = val;
      hs.omega = (val == 0.0) ? 0.0 : hs.ts / val;  // BiCGSTAB.f90:139-145
      it = hs.bd_iter;
      stage = 4;
    } else {
      MH_TRY(sorted_dot<T>((long long)m, r, rt, dsum, s));
      MH_HIP(hipMemcpy(&val, dsum, sizeof(double), hipMemcpyDeviceToHost));
      hs.beta = (val / hs.rho) * (hs.alpha / hs.omega);
      hs.rho_prev = hs.rho;
      hs.rho = val;
      it = hs.bd_iter + 1;
      hs.rhok[it & 1] = val;
      stage = 0;
    }
    resumed = stage;
    hs.bd_stage = 0;
    hs.done = 0;
    MH_HIP(hipMemcpy(st, &hs, sizeof(KState), hipMemcpyHostToDevice));
    if (it > P.max_iter) break;
  }
  out->iter = hs.iter;
  out->resid = hs.resid;
  out->converged = hs.converged;
  out->breakdown = hs.breakdown;
  if (P.keep_hist && hs.iter > 0) {
    out->hist.resize((size_t)hs.iter + 1);
    MH_HIP(hipMemcpy(out->hist.data(), hist, sizeof(double) * ((size_t)hs.iter + 1), hipMemcpyDeviceToHost));
  }
  return 0;
}

}  // namespace

template <typename T>
int krylov_solve(Matrix& M, const SolveParams& P, const T* b, T* x, Workspace& W, SolveResult* out,
                 hipStream_t s) {
  if (P.precond != 0 && P.precond != 1) {
    set_error("preconditioner %d is not implemented on the GPU path (NONE=0, DIAG=1)", P.precond);
    return -3;
  }
  if (P.method < 1 || P.method > 8) {
    set_error("method %d is not implemented on the GPU path (CG=1, GropCG=2, PipeCG=3, PipeCR=4, BiCGSTAB=5, "
              "PipeBiCGSTAB=6, BiCGSTAB_noprec=7, PipeBiCGSTAB_noprec=8)", P.method);
    return -4;
  }
  if (P.max_iter < 0) {
    set_error("max_iter must be >= 0");
    return -2;
  }
  if (sizeof(T) == 4) MH_TRY(matrix_make_fp32(&M, s));
  if (P.precond == 1 && P.method != 7 && P.method != 8) MH_TRY(diag_setup<T>(M, s));
  const auto t0 = std::chrono::steady_clock::now();
  int rc = 0;
  if (P.method == 1) {
    if (P.precond == 1) {
      MH_DISPATCH_B(M.b, rc = (cg_impl<T, B, true>(M, P, b, x, W, out, s)));
    } else {
      MH_DISPATCH_B(M.b, rc = (cg_impl<T, B, false>(M, P, b, x, W, out, s)));
    }
  } else if (P.method == 2) {
    if (P.precond == 1) {
      MH_DISPATCH_B(M.b, rc = (gropcg_impl<T, B, true>(M, P, b, x, W, out, s)));
    } else {
      MH_DISPATCH_B(M.b, rc = (gropcg_impl<T, B, false>(M, P, b, x, W, out, s)));
    }
  } else if (P.method == 3) {
    if (P.precond == 1) {
      MH_DISPATCH_B(M.b, rc = (pipecg_impl<T, B, true>(M, P, b, x, W, out, s)));
    } else {
      MH_DISPATCH_B(M.b, rc = (pipecg_impl<T, B, false>(M, P, b, x, W, out, s)));
    }
  } else if (P.method == 4) {
    if (P.precond == 1) {
      MH_DISPATCH_B(M.b, rc = (pipecr_impl<T, B, true>(M, P, b, x, W, out, s)));
    } else {
      MH_DISPATCH_B(M.b, rc = (pipecr_impl<T, B, false>(M, P, b, x, W, out, s)));
    }
  } else if (P.method == 6) {
    if (P.precond == 1) {
      MH_DISPATCH_B(M.b, rc = (pipebicgstab_impl<T, B, true>(M, P, b, x, W, out, s)));
    } else {
      MH_DISPATCH_B(M.b, rc = (pipebicgstab_impl<T, B, false>(M, P, b, x, W, out, s)));
    }
  } else if (P.method == 8) {  // PipeBiCGSTAB_noprec.f90
    MH_DISPATCH_B(M.b, rc = (pipebicgstab_impl<T, B, false>(M, P, b, x, W, out, s)));
  } else if (P.method == 7) {  // BiCGSTAB_noprec.f90: no preconditioner whatever Iarray[2] says
    MH_DISPATCH_B(M.b, rc = (bicgstab_impl<T, B, false>(M, P, b, x, W, out, s, 50)));
  } else {
    if (P.precond == 1) {
      MH_DISPATCH_B(M.b, rc = (bicgstab_impl<T, B, true>(M, P, b, x, W, out, s, kIterRR)));
    } else {
      MH_DISPATCH_B(M.b, rc = (bicgstab_impl<T, B, false>(M, P, b, x, W, out, s, kIterRR)));
    }
  }
  out->t_solve = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return rc;
}
template int krylov_solve<double>(Matrix&, const SolveParams&, const double*, double*, Workspace&, SolveResult*,
                                  hipStream_t);
template int krylov_solve<float>(Matrix&, const SolveParams&, const float*, float*, Workspace&, SolveResult*,
                                 hipStream_t);

// ---------------------------------------------------------------------------
// FP32-inner / FP64-outer iterative refinement.
// ---------------------------------------------------------------------------
__global__ void k_scale_to_fp32(size_t m, const double* __restrict__ r, double s, float* __restrict__ r32) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x)
    r32[i] = (float)(r[i] * s);
}
__global__ void k_add_scaled(size_t m, double* __restrict__ x, const float* __restrict__ d, double s) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x)
    x[i] = fma(s, (double)d[i], x[i]);
}

static int grid_1d(size_t m) {
  size_t g = (m + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  return (int)(g > 0 ? g : 1);
}

int refine_solve(Matrix& M, const RefineParams& P, const double* b, double* x, Workspace& W, Workspace& Wo,
                 RefineResult* out, hipStream_t s) {
  const size_t vec = (size_t)M.np * M.b;
  const size_t m = (size_t)M.n * M.b;
  // outer buffers live in their own workspace (kept across calls); W serves the inner solve
  const size_t nparts = partials_needed((long long)M.nslices * kWave) * 2;
  MH_TRY(Wo.ensure(vec * (sizeof(double) + 2 * sizeof(float)) + (nparts + 64) * sizeof(double) + 4 * 256));
  Carve cv{Wo.raw.p};
  double* r64 = cv.take<double>(vec);
  float* r32 = cv.take<float>(vec);
  float* d32 = cv.take<float>(vec);
  double* parts = cv.take<double>(nparts + 64);
  double* rr_dev = parts + nparts;
  MH_TRY(matrix_make_fp32(&M, s));
  if (P.precond == 1) {
    MH_TRY(diag_setup<double>(M, s));
    MH_TRY(diag_setup<float>(M, s));
  }
  if (P.init_x) MH_HIP(hipMemsetAsync(x, 0, sizeof(double) * m, s));
  if (vec > m) {
    MH_HIP(hipMemsetAsync(r32 + m, 0, sizeof(float) * (vec - m), s));
    MH_HIP(hipMemsetAsync(d32 + m, 0, sizeof(float) * (vec - m), s));
  }
  MH_TRY(residual<double>(M, x, b, r64, parts, rr_dev, s));
  double rr = 0.0;
  MH_HIP(hipMemcpyAsync(&rr, rr_dev, sizeof(double), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  out->outer_resid.clear();
  out->inner_iters.clear();
  out->inner_total = 0;
  out->outer = 0;
  const double B2 = rr;
  if (B2 == 0.0) {
    MH_HIP(hipMemsetAsync(x, 0, sizeof(double) * m, s));
    MH_HIP(hipStreamSynchronize(s));
    out->converged = 1;
    out->resid = 0.0;
    return 0;
  }
  double resid = 1.0;
  out->outer_resid.push_back(resid);
  SolveParams ip;
  ip.method = P.method;
  ip.precond = P.precond;
  ip.max_iter = P.inner_max_iter;
  ip.init_x = 1;
  ip.tol = P.inner_tol;
  ip.keep_hist = 0;
  ip.rhs_nonzero = 1;  // r / ||r|| with ||r|| > 0
  for (int k = 1; k <= P.max_outer; ++k) {
    const double nr = std::sqrt(rr);
    // A correction only needs to take the residual from `resid` to below tol:
    // once that is a smaller reduction than inner_tol asks for, the inner
    // solve stops at half of it (tol / resid / 2) instead of running to
    // inner_tol.  A correction that falls short is followed by another.
    ip.tol = std::max(P.inner_tol, 0.5 * P.tol / resid);
    k_scale_to_fp32<<<grid_1d(m), kBlock, 0, s>>>(m, r64, 1.0 / nr, r32);
    MH_HIP(hipGetLastError());
    SolveResult ir;
    MH_TRY(krylov_solve<float>(M, ip, r32, d32, W, &ir, s));
    out->inner_iters.push_back(ir.iter);
    out->inner_total += ir.iter > 0 ? ir.iter : 0;
    k_add_scaled<<<grid_1d(m), kBlock, 0, s>>>(m, x, d32, nr);
    MH_HIP(hipGetLastError());
    MH_TRY(residual<double>(M, x, b, r64, parts, rr_dev, s));
    MH_HIP(hipMemcpyAsync(&rr, rr_dev, sizeof(double), hipMemcpyDeviceToHost, s));
    MH_HIP(hipStreamSynchronize(s));
    resid = std::sqrt(rr / B2);
    out->outer_resid.push_back(resid);
    out->outer = k;
    if (resid < P.tol) {
      out->converged = 1;
      break;
    }
    if (!(resid == resid)) break;  // NaN: stop
  }
  out->resid = resid;
  return 0;
}

}  // namespace mh
