// Launch helpers shared by the translation units: SELL SpMV variants and
// row-parallel kernels, each followed by its reduction finalize kernel.
#pragma once

#include <cmath>
#include <deque>

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

// Partials a SELL launch of this matrix may write (per value of NV).
inline size_t sell_partials(const Matrix& M) { return (size_t)(M.nslices + kWpb - 1) / kWpb + 8; }

template <typename T, int B, class Op>
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
  }
  if constexpr (Op::NV > 0) {
    k_finalize<Op><<<1, kBlock, 0, s>>>(op, partials, nwg);
    MH_HIP(hipGetLastError());
  }
  return 0;
}

template <class Op>
int launch_rows(int nrows, const Op& op, double* partials, hipStream_t s) {
  const int ntiles = (nrows + kBlock - 1) / kBlock;
  if (ntiles > 0) {
    k_rows<Op><<<padded_grid(ntiles), kBlock, 0, s>>>(nrows, op, partials, ntiles);
    MH_HIP(hipGetLastError());
  }
  if constexpr (Op::NV > 0) {
    k_finalize<Op><<<1, kBlock, 0, s>>>(op, partials, ntiles);
    MH_HIP(hipGetLastError());
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
