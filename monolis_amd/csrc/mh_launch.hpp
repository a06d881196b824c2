// Launch helpers shared by the translation units: SELL SpMV variants and
// row-parallel kernels, each followed by its reduction finalize kernel.
#pragma once

#include <cmath>
#include <deque>
#include <type_traits>
#include <vector>

#include "mh_internal.hpp"
#include "mh_kernels.hpp"

namespace mh {

// Compute units of the current device (context init reads it once).
int cu_count();

// Launch geometry of a SELL variant: logical workgroups and the
// slices-per-wave mode (0: one slice per wavefront, 1: contiguous ranges).
inline void sell_geometry(const Matrix& M, int variant, int* nwg, int* spw) {
  const SellVariant v = sell_variant(variant);
  const int full = (M.nslices + kWpb - 1) / kWpb;
  if (v.wpc == 0) {
    *nwg = full;
    *spw = 0;
  } else {
    int g = cu_count() * v.wpc / kWpb;
    if (g > full) g = full;
    if (g < 1) g = 1;
    *nwg = g;
    *spw = 1;
  }
}

// ---------------------------------------------------------------------------
// Solver time split by kernel class (monolis Rarray slots time_spmv,
// time_dotp, time_prec; src/define/def_solver_prm.f90:151-163).  While a
// KTimer is active on the calling thread, every launch helper records one
// event after its kernel; the interval since the previous event is charged to
// that kernel's class: SpMV kernels (with their fused dot products) to spmv,
// row kernels that apply the preconditioner to prec, the other row kernels
// and reduction kernels (vector updates and dot products) to dotp.  The
// classes therefore cover the solve's device time without overlap.
// ---------------------------------------------------------------------------
enum : int { kClsSpmv = 0, kClsDot = 1, kClsPrec = 2 };

template <class Op, int DEF, class = void>
struct op_kclass {
  static constexpr int value = DEF;
};
template <class Op, int DEF>
struct op_kclass<Op, DEF, std::void_t<decltype(Op::KCLASS)>> {
  static constexpr int value = Op::KCLASS;
};

struct KTimer {
  struct Mark {
    hipEvent_t ev;
    int cls;
  };
  hipStream_t s = nullptr;
  std::deque<Mark> marks;        // marks.front() is the last harvested event
  std::vector<hipEvent_t> pool;  // recycled events
  double ms[3] = {0.0, 0.0, 0.0};
  int err = 0;
  explicit KTimer(hipStream_t st) : s(st) { mark(kClsDot); }
  ~KTimer() {
    for (auto& m : marks) (void)hipEventDestroy(m.ev);
    for (auto e : pool) (void)hipEventDestroy(e);
  }
  KTimer(const KTimer&) = delete;
  KTimer& operator=(const KTimer&) = delete;
  void mark(int cls) {
    hipEvent_t e = nullptr;
    if (!pool.empty()) {
      e = pool.back();
      pool.pop_back();
    } else if (hipEventCreate(&e) != hipSuccess) {
      err = 1;
      return;
    }
    if (hipEventRecord(e, s) != hipSuccess) {
      err = 1;
      pool.push_back(e);
      return;
    }
    marks.push_back(Mark{e, cls});
    if (marks.size() > 256) harvest(false);
  }
  // Charges every completed interval; with `wait` first waits for the last event.
  void harvest(bool wait) {
    if (wait && !marks.empty() && hipEventSynchronize(marks.back().ev) != hipSuccess) err = 1;
    while (marks.size() >= 2) {
      if (!wait && hipEventQuery(marks[1].ev) != hipSuccess) break;
      float t = 0.0f;
      if (hipEventElapsedTime(&t, marks[0].ev, marks[1].ev) != hipSuccess) err = 1;
      ms[marks[1].cls] += t;
      pool.push_back(marks[0].ev);
      marks.pop_front();
    }
  }
};

// The timer launches on this thread report to (null: none).
KTimer*& active_ktimer();

inline void ktimer_mark(int cls) {
  if (KTimer* t = active_ktimer()) t->mark(cls);
}

// Partials a SELL launch of this matrix may write (per value of NV).
inline size_t sell_partials(const Matrix& M) { return (size_t)(M.nslices + kWpb - 1) / kWpb + 8; }

// FIN = false leaves the partial sums for the next kernel's prologue (an Op
// with NPRE consumes them there) instead of launching the finalize kernel.
template <typename T, int B, class Op, bool FIN = true>
int launch_sell(const Matrix& M, const Op& op, double* partials, int variant, hipStream_t s) {
  int nwg = 0, spw = 0;
  if (M.nslices > 0) {
    sell_geometry(M, variant, &nwg, &spw);
    const int u = sell_variant(variant).unroll;
    if constexpr (B == 1 || B == 3) {
      MH_SELL_UNROLL(u, k_sell<T, T, B, U, Op><<<padded_grid(nwg), kBlock, 0, s>>>(view<T>(M), op, partials,
                                                                                   nwg, spw));
    } else {
      (void)u;
      k_sell<T, T, B, 3, Op><<<padded_grid(nwg), kBlock, 0, s>>>(view<T>(M), op, partials, nwg, spw);
    }
    MH_HIP(hipGetLastError());
    ktimer_mark(op_kclass<Op, kClsSpmv>::value);
  }
  if constexpr (FIN && Op::NV > 0) {
    k_finalize<Op><<<1, kBlock, 0, s>>>(op, partials, nwg);
    MH_HIP(hipGetLastError());
    ktimer_mark(kClsDot);
  }
  return 0;
}

// Logical workgroups of a row-kernel launch over nrows block rows: one per
// tile of row_block rows, at most 32 wavefronts' worth per CU (the persistent
// grid of k_rows); also the number of partial sets the launch leaves.
template <typename VT, int BS>
int rows_grid(long long nrows) {
  constexpr int NT = row_block<VT, BS>();
  const long long cap = (long long)cu_count() * (32 * kWave / NT);
  const long long t = rows_tiles<VT, BS>(nrows);
  return (int)(t < cap ? t : cap);
}

template <class Op, bool FIN = true>
int launch_rows(int nrows, const Op& op, double* partials, hipStream_t s) {
  constexpr int NT = row_block<typename Op::VT, Op::BS>();
  const int ntiles = rows_tiles<typename Op::VT, Op::BS>(nrows);
  const int nwg = rows_grid<typename Op::VT, Op::BS>(nrows);
  if (ntiles > 0) {
    k_rows<Op><<<padded_grid(nwg), NT, 0, s>>>(nrows, op, partials, nwg, ntiles);
    MH_HIP(hipGetLastError());
    ktimer_mark(op_kclass<Op, kClsDot>::value);
  }
  if constexpr (FIN && Op::NV > 0) {
    k_finalize<Op><<<1, kBlock, 0, s>>>(op, partials, nwg);
    MH_HIP(hipGetLastError());
    ktimer_mark(kClsDot);
  }
  return 0;
}

// Sub-allocation of a solver workspace, 256-byte aligned pieces.
struct Carve {
  unsigned char* base;
  size_t off = 0;
  template <typename U>
  U* take(size_t count) {
    off = (off + 255) / 256 * 256;
    U* p = (U*)(base + off);
    off += count * sizeof(U);
    return p;
  }
};

// Convergence polling one chunk behind the device: after each chunk of
// iterations the solver enqueues a 16-byte copy of its PollView and waits for
// the previous chunk's.  The chunk length adapts: from the last two polled
// residuals the geometric rate predicts the iterations still needed for `tol`,
// and the next chunk covers no more than those (at least kMinChunk), so few
// iterations of no-op launches follow convergence.
struct Poller {
  static constexpr int kMinChunk = 2;
  Context* c;
  const void* view;  // device PollView (KState::resid / ZState::resid_abs)
  double tol;
  std::deque<int> pending;
  int next = 0;
  int n_seen = 0;
  PollView seen[2] = {};  // the last two polled views, oldest first
  int issue(hipStream_t s) {
    const int slot = next;
    next = (next + 1) % kFlagSlots;
    c->h_poll[slot] = PollView{0.0, 0, 0};
    MH_HIP(hipMemcpyAsync(&c->h_poll[slot], view, sizeof(PollView), hipMemcpyDeviceToHost, s));
    MH_HIP(hipEventRecord(c->ev[slot], s));
    pending.push_back(slot);
    return 0;
  }
  // waits for the polls beyond the newest `keep`; *done = 1 when one reports done
  int check(size_t keep, int* done) {
    *done = 0;
    while (pending.size() > keep) {
      const int slot = pending.front();
      pending.pop_front();
      MH_HIP(hipEventSynchronize(c->ev[slot]));
      const PollView v = c->h_poll[slot];
      if (v.done) {
        *done = 1;
        pending.clear();
        return 0;
      }
      if (v.iter > 0 && (n_seen == 0 || v.iter > seen[1].iter)) {
        seen[0] = seen[1];
        seen[1] = v;
        ++n_seen;
      }
    }
    return 0;
  }
  // Iterations to enqueue next, the first of them being `it`.
  int next_chunk(int it, int max_chunk) const {
    if (max_chunk <= kMinChunk || n_seen < 2 || !(tol > 0.0)) return max_chunk;
    const PollView& a = seen[0];
    const PollView& b = seen[1];
    if (!(b.resid > 0.0) || !(a.resid > b.resid) || b.iter <= a.iter) return max_chunk;
    if (b.resid <= tol) return kMinChunk;
    const double rate = std::log(b.resid / a.resid) / (double)(b.iter - a.iter);  // < 0
    const double need = std::log(tol / b.resid) / rate;                           // iterations after b
    const double left = std::ceil(need) - (double)(it - 1 - b.iter);
    if (!(left < (double)max_chunk)) return max_chunk;
    return left < (double)kMinChunk ? kMinChunk : (int)left;
  }
};

#define MH_DISPATCH_B(b, ...)                                            \
  switch (b) {                                                           \
    case 1: { constexpr int B = 1; __VA_ARGS__; } break;                  \
    case 2: { constexpr int B = 2; __VA_ARGS__; } break;                  \
    case 3: { constexpr int B = 3; __VA_ARGS__; } break;                  \
    case 4: { constexpr int B = 4; __VA_ARGS__; } break;                  \
    case 5: { constexpr int B = 5; __VA_ARGS__; } break;                  \
    case 6: { constexpr int B = 6; __VA_ARGS__; } break;                  \
    case 7: { constexpr int B = 7; __VA_ARGS__; } break;                  \
    case 8: { constexpr int B = 8; __VA_ARGS__; } break;                  \
    default: set_error("unsupported block size %d", b); return -2;       \
  }

}  // namespace mh
