// Device-side building blocks: the sliced-ELL block layout, the SpMV row
// loop, the block-Jacobi apply, and the generic fused SpMV kernel.
//
// Layout ("SELL-64", built once per matrix from monolis's BCSR arrays):
//   rows are grouped in slices of 64 consecutive block rows, one wavefront per
//   slice and one lane per block row.  Slice s holds w_s = max row length slots;
//   slot k of lane l stores the k-th block of row 64*s+l (CSR order, so columns
//   stay ascending).  For slot t = slice_ptr[s] + k:
//     col[t*64 + l]                 0-based block column
//     val[t*B*B*64 + sell_vofs(e)]  block entry e = r*B + c (row-major block);
//                                   entries 2q, 2q+1 of a lane side by side
//                                   (real values), the odd last one alone
//   so every load of a slot is one contiguous 256 B (int), 512 B (double) or
//   1 KiB (double pair) line-aligned request for the whole wavefront.  Slots past a row's length
//   are padding (col = own row, val = 0); a padded slot is loaded but masked
//   off by the row length taken from monolis's `index` array.  When every
//   |col - row| fits in 16 bits the SpMV reads the columns from a 16-bit delta
//   copy (SellView::col16), which halves the index stream.
#pragma once

#include <type_traits>

#include "mh_common.hpp"

namespace mh {

template <typename TV>
struct SellView {
  int nrows;               // block rows that are multiplied (monolis N)
  int nslices;             // ceil(nrows / 64)
  const int* slice_ptr;    // [nslices + 1] slot offsets
  const int* rowptr;       // [nrows + 1] monolis `index` (0-based offsets)
  const int* col;          // [nslots * 64]
  const TV* val;           // [nslots * 64 * B * B]
  // Optional 16-bit copy of `col` as deltas col - row, two slots per 32-bit
  // word: word (t / 2) * 64 + l holds slot t of lane l in its low half when t
  // is even, high half when odd.  Null when some |col - row| exceeds 32767.
  const unsigned* col16;
};

// Column id of slot t (global slot index) for lane `lane` (block row `row`).
// The 16-bit form halves the index stream; t's parity is wave-uniform, so the
// half select is scalar.  Its words are loaded without the non-temporal hint:
// each word serves two consecutive slots, and a streaming load would let the
// line leave L2 before the second slot reads it (measured: the index stream
// was then fetched twice).
template <bool NT, bool C16, typename TV>
__device__ __forceinline__ int sell_col(const SellView<TV>& M, size_t t, int lane, int row, bool valid) {
  if constexpr (C16) {
    const unsigned w = ld_val<false>(M.col16 + (t >> 1) * kWave + lane);
    const int d = (t & 1) ? (int)(short)(w >> 16) : (int)(short)(w & 0xffffu);
    return valid ? row + d : 0;
  } else {
    return ld_val<NT>(M.col + t * kWave + lane);
  }
}

// acc[r] += sum_k sum_c val(k, r, c) * x(col_k, c) for the lane's row.  `xg`
// gathers the B entries of the input vector at a block column.  Products are
// accumulated in the order of monolis's matvec_33_R (matvec.f90:805-823):
// block by block, and inside a block c = 1..B.
//
// The slot loop is wave-uniform (one slice per wavefront) and processes U slots
// per step with every load issued up front: U column ids, U*B*B value planes and
// then U gathers.  Padding slots are always safe to load (col = own row,
// val = 0), so no lane branches; a select keeps a padded slot from touching the
// accumulator, which preserves IEEE semantics (0 * inf never enters a sum).
// Row length of a lane, either a plain int or the row's two offsets as
// loaded from monolis's `index`.  The offsets pass through an empty asm
// statement where the length is first needed (after the step's matrix loads
// are issued): the wait for their load then overlaps those loads instead of
// stalling the start of every slice (a loop-invariant length would otherwise
// be computed, and waited for, before the slot loop).
struct RowLen {
  int r0, r1;
  bool valid;
};
__device__ __forceinline__ int row_len(int len, int) { return len; }
__device__ __forceinline__ int row_len(const RowLen& r, int k) {
  // not volatile (no side effect, so uniform loads keep using the scalar
  // cache); the slot index as an input keeps it inside the loop
  int a = r.r0, b = r.r1;
  asm("" : "+v"(a), "+v"(b) : "s"(k));
  return r.valid ? b - a : 0;
}

template <typename TV, typename TA, int B, int U, bool NT, bool C16, class XG, class LEN>
__device__ __forceinline__ void sell_row_impl(const SellView<TV>& M, int s, int lane, const LEN& len, const XG& xg,
                                              TA (&acc)[B]) {
  const int k0 = M.slice_ptr[s];
  const int w = M.slice_ptr[s + 1] - k0;
  const int row = s * kWave + lane;
  const bool valid = row < M.nrows;
  const TV* vals = M.val + (size_t)k0 * kWave * (B * B);
  int k = 0;
  for (; k + U <= w; k += U) {
    int c[U];
    TV a[U][B * B];
    TA xv[U][B];
#pragma unroll
    for (int u = 0; u < U; ++u) c[u] = sell_col<NT, C16>(M, (size_t)(k0 + k + u), lane, row, valid);
#pragma unroll
    for (int u = 0; u < U; ++u) ld_slot<B * B, NT>(vals + (size_t)(k + u) * (B * B) * kWave, lane, a[u]);
#pragma unroll
    for (int u = 0; u < U; ++u) xg(c[u], xv[u]);
    const int lim = row_len(len, k);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool on = (k + u) < lim;
#pragma unroll
      for (int r = 0; r < B; ++r) {
        TA t = acc[r];
#pragma unroll
        for (int cc = 0; cc < B; ++cc) t = madd((TA)a[u][r * B + cc], xv[u][cc], t);
        acc[r] = pick(on, t, acc[r]);
      }
    }
  }
  for (; k < w; ++k) {
    const int c = sell_col<NT, C16>(M, (size_t)(k0 + k), lane, row, valid);
    TV a[B * B];
    ld_slot<B * B, NT>(vals + (size_t)k * (B * B) * kWave, lane, a);
    TA xv[B];
    xg(c, xv);
    const bool on = k < row_len(len, k);
#pragma unroll
    for (int r = 0; r < B; ++r) {
      TA t = acc[r];
#pragma unroll
      for (int cc = 0; cc < B; ++cc) t = madd((TA)a[r * B + cc], xv[cc], t);
      acc[r] = pick(on, t, acc[r]);
    }
  }
}

template <typename TV, typename TA, int B, int U, bool NT, class XG, class LEN>
__device__ __forceinline__ void sell_row(const SellView<TV>& M, int s, int lane, const LEN& len, const XG& xg,
                                         TA (&acc)[B]) {
  if (M.col16)
    sell_row_impl<TV, TA, B, U, NT, true>(M, s, lane, len, xg, acc);
  else
    sell_row_impl<TV, TA, B, U, NT, false>(M, s, lane, len, xg, acc);
}

// Software-pipelined form of sell_row: the column ids and values of step j+1
// are issued right after step j's gathers and before step j's products, so a
// wavefront keeps the next step of the matrix stream in flight while it waits
// on the (L2-resident) x gathers.  The wait counter of vector loads is in
// order, so the gathers must be issued before the prefetch (scheduling
// barriers keep that order), and the two step buffers alternate by name
// (ping-pong) instead of being copied, which would wait for the whole
// prefetch at the top of every step.  Same arithmetic order as sell_row.
template <typename TV, int B, int U>
struct SellStep {
  int c[U];
  TV a[U][B * B];
};

template <typename TV, typename TA, int B, int U, bool NT, bool C16, class XG, class LEN>
__device__ __forceinline__ void sell_row_pipe_impl(const SellView<TV>& M, int s, int lane, const LEN& len,
                                                   const XG& xg, TA (&acc)[B]) {
  const int k0 = M.slice_ptr[s];
  const int w = M.slice_ptr[s + 1] - k0;
  const int row = s * kWave + lane;
  const bool valid = row < M.nrows;
  const TV* vals = M.val + (size_t)k0 * kWave * (B * B);
  const int ns = w / U;  // full steps
  auto fetch = [&](SellStep<TV, B, U>& st, int j) {
    const int k = j * U;
#pragma unroll
    for (int u = 0; u < U; ++u) st.c[u] = sell_col<NT, C16>(M, (size_t)(k0 + k + u), lane, row, valid);
#pragma unroll
    for (int u = 0; u < U; ++u) ld_slot<B * B, NT>(vals + (size_t)(k + u) * (B * B) * kWave, lane, st.a[u]);
  };
  // step j from buffer `cur`; if nxt != nullptr, prefetch step jn into it
  auto step = [&](const SellStep<TV, B, U>& cur, int j, SellStep<TV, B, U>* nxt, int jn) {
    TA xv[U][B];
#pragma unroll
    for (int u = 0; u < U; ++u) xg(cur.c[u], xv[u]);
    __builtin_amdgcn_sched_barrier(0);
    if (nxt) fetch(*nxt, jn);
    __builtin_amdgcn_sched_barrier(0);
    const int k = j * U;
    const int lim = row_len(len, k);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const bool on = (k + u) < lim;
#pragma unroll
      for (int r = 0; r < B; ++r) {
        TA t = acc[r];
#pragma unroll
        for (int cc = 0; cc < B; ++cc) t = madd((TA)cur.a[u][r * B + cc], xv[u][cc], t);
        acc[r] = pick(on, t, acc[r]);
      }
    }
  };
  SellStep<TV, B, U> b0, b1;
  int j = 0;
  if (ns > 0) {
    fetch(b0, 0);
    for (; j + 2 < ns; j += 2) {
      step(b0, j, &b1, j + 1);
      step(b1, j + 1, &b0, j + 2);
    }
    if (ns - j == 2) {
      step(b0, j, &b1, j + 1);
      step(b1, j + 1, nullptr, 0);
    } else {
      step(b0, j, nullptr, 0);
    }
  }
  for (int k = ns * U; k < w; ++k) {
    const int c = sell_col<NT, C16>(M, (size_t)(k0 + k), lane, row, valid);
    TV a[B * B];
    ld_slot<B * B, NT>(vals + (size_t)k * (B * B) * kWave, lane, a);
    TA xv[B];
    xg(c, xv);
    const bool on = k < row_len(len, k);
#pragma unroll
    for (int r = 0; r < B; ++r) {
      TA t = acc[r];
#pragma unroll
      for (int cc = 0; cc < B; ++cc) t = madd((TA)a[r * B + cc], xv[cc], t);
      acc[r] = pick(on, t, acc[r]);
    }
  }
}

template <typename TV, typename TA, int B, int U, bool NT, class XG, class LEN>
__device__ __forceinline__ void sell_row_pipe(const SellView<TV>& M, int s, int lane, const LEN& len, const XG& xg,
                                              TA (&acc)[B]) {
  if (M.col16)
    sell_row_pipe_impl<TV, TA, B, U, NT, true>(M, s, lane, len, xg, acc);
  else
    sell_row_pipe_impl<TV, TA, B, U, NT, false>(M, s, lane, len, xg, acc);
}

// Kernel variants of the SELL SpMV.  A variant fixes
//   unroll: slots per step in sell_row (loads in flight per wave vs. VGPRs), and
//   wpc:    0 = one wavefront per slice over a grid covering every slice, or
//           w > 0 = a persistent grid of w wavefronts per CU, each wavefront
//           walking a contiguous range of slices (fewer, longer HBM streams).
//   pipe:   1 = software-pipelined slot loop (sell_row_pipe), 0 = sell_row.
// Matrix streams are always non-temporal.  Variant 4 is the default before
// tuning; monolis_hip_matrix_autotune times all of them per matrix.
struct SellVariant {
  int unroll;  // slots per step, plus kSellPipe when pipelined
  int wpc;
};
constexpr int kSellPipe = 16;
constexpr int kNumVariants = 22;
__host__ __device__ constexpr SellVariant sell_variant(int v) {
  constexpr int P = kSellPipe;
  constexpr int u[kNumVariants] = {3, 2, 4, 1, 3, 2, 4, 1, 3, 2, 4, 3, 2, 4, 3, 4,
                                   P + 1, P + 1, P + 1, P + 1, P + 1, P + 1};
  constexpr int w[kNumVariants] = {8, 0, 0, 0, 0, 8, 8, 8, 4, 4, 4, 12, 12, 12, 16, 16,
                                   0, 4, 8, 12, 16, 24};
  return SellVariant{u[v], w[v]};
}
// B > 3 blocks carry B*B values per slot: cap the unroll to bound VGPRs.
template <int B, int U0>
struct SellCfg {
  static constexpr int U = B == 1 ? U0 * 2 : B <= 3 ? U0 : B <= 5 ? (U0 < 2 ? U0 : 2) : 1;
};

// Wave-local ordering of LDS accesses between lanes of one wavefront (the LDS
// region is private to the wavefront, so no workgroup barrier is needed).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-row vector I/O of one wavefront covering 64 consecutive block rows
// (row0 .. row0+63).  A lane owns one block row, i.e. B consecutive entries, so
// direct per-lane access would be a 8*B-byte strided pattern that writes
// partial cache lines.  For B > 1 the values go through a 64*B LDS tile and
// are read / written by the wavefront as B contiguous 64-entry runs.  Every
// lane of the wavefront must call load/store (rows >= nrows are skipped).
// Vector loads of the row kernels are plain: non-temporal loads measured the
// same on the 100^3 BCSR3 solves (CG 20.8 vs 21.0 ms, BiCGSTAB 32.7 vs 31.9 ms,
// refined CG 12.7 vs 12.9 ms, plain vs non-temporal).
template <typename T, int B>
struct WaveRows {
  T* lds;     // 64*B entries private to this wavefront (unused for B == 1)
  int row0;
  int nrows;
  int lane;
  __device__ __forceinline__ int count() const {
    const int c = nrows - row0;
    return (c < kWave ? c : kWave) * B;
  }
  __device__ __forceinline__ void load(const T* __restrict__ src, T (&v)[B]) const {
    if constexpr (B == 1) {
      v[0] = (T)0;
      if (row0 + lane < nrows) v[0] = ld_val<false>(src + (size_t)row0 + lane);
    } else {
      const int cnt = count();
      const T* base = src + (size_t)row0 * B;
#pragma unroll
      for (int j = 0; j < B; ++j) {
        const int e = j * kWave + lane;
        if (e < cnt) lds[e] = ld_val<false>(base + e);
      }
      wave_sync();
#pragma unroll
      for (int k = 0; k < B; ++k) v[k] = lds[lane * B + k];
      wave_sync();
    }
  }
  // The wavefront's 64*B entries of src in coalesced order (lane takes entries
  // j*64 + lane).  A row kernel fetches every vector it reads this way before
  // its first store, so all of its loads are in flight together; elementwise
  // updates work in this order directly (put_flat), block operations take
  // transpose() to row order.  Entries past the last row re-read the last entry
  // (no branch around the load, which would make the compiler wait for the
  // loads before it): they are never stored and masked out of every sum.
  __device__ __forceinline__ void fetch(const T* __restrict__ src, T (&c)[B]) const {
    const int last = count() - 1;
    const T* base = src + (size_t)row0 * B;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int e = j * kWave + lane;
      c[j] = ld_val<false>(base + (e < last ? e : last));
    }
  }
  __device__ __forceinline__ void transpose(const T (&c)[B], T (&v)[B]) const {
    if constexpr (B == 1) {
      v[0] = c[0];
    } else {
#pragma unroll
      for (int j = 0; j < B; ++j) lds[j * kWave + lane] = c[j];
      wave_sync();
#pragma unroll
      for (int k = 0; k < B; ++k) v[k] = lds[lane * B + k];
      wave_sync();
    }
  }
  __device__ __forceinline__ void put_flat(T* __restrict__ dst, const T (&c)[B]) const {
    const int cnt = count();
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const int e = j * kWave + lane;
      if (e < cnt) put(dst, (size_t)row0 * B + e, c[j]);
    }
  }
  // Outputs are written non-temporally: they are consumed by a later kernel,
  // and streaming them out during the kernel avoids leaving megabytes of dirty
  // L2 lines for the end-of-kernel write-back.
  __device__ __forceinline__ void put(T* __restrict__ dst, size_t i, T v) const { st_stream(dst + i, v); }
  __device__ __forceinline__ void store(T* __restrict__ dst, const T (&v)[B]) const {
    if constexpr (B == 1) {
      if (row0 + lane < nrows) put(dst, (size_t)row0 + lane, v[0]);
    } else {
      const int cnt = count();
#pragma unroll
      for (int k = 0; k < B; ++k) lds[lane * B + k] = v[k];
      wave_sync();
#pragma unroll
      for (int j = 0; j < B; ++j) {
        const int e = j * kWave + lane;
        if (e < cnt) put(dst, (size_t)row0 * B + e, lds[e]);
      }
      wave_sync();
    }
  }
};

// Block-Jacobi apply on one block row, in place.  D is the factored diagonal
// block in structure-of-arrays form: entry e of row i at D[e * nrows + i].
// The arithmetic order is that of diag_33.f90:123-135 for B = 3 and of
// diag_nn.f90:207-228 otherwise (B = 1 reduces to x * D).
template <typename T, typename TD, int B>
__device__ __forceinline__ void diag_apply(const TD* __restrict__ D, int nrows, int i, T (&x)[B]) {
  if constexpr (B == 1) {
    x[0] = (T)D[i] * x[0];
  } else if constexpr (B == 3) {
    const size_t n = (size_t)nrows;
    const T l3 = (T)D[3 * n + i], l6 = (T)D[6 * n + i], l7 = (T)D[7 * n + i];
    const T l8 = (T)D[8 * n + i], l4 = (T)D[4 * n + i], l5 = (T)D[5 * n + i];
    const T l0 = (T)D[i], l1 = (T)D[1 * n + i], l2 = (T)D[2 * n + i];
    T x1 = x[0], x2 = x[1], x3 = x[2];
    x2 = x2 - l3 * x1;
    x3 = x3 - l6 * x1 - l7 * x2;
    x3 = l8 * x3;
    x2 = l4 * (x2 - l5 * x3);
    x1 = l0 * (x1 - l2 * x3 - l1 * x2);
    x[0] = x1;
    x[1] = x2;
    x[2] = x3;
  } else {
    const size_t n = (size_t)nrows;
#pragma unroll
    for (int j = 1; j < B; ++j)
#pragma unroll
      for (int k = 0; k < j; ++k) x[j] = x[j] - (T)D[(size_t)(B * j + k) * n + i] * x[k];
#pragma unroll
    for (int j = B - 1; j >= 0; --j) {
#pragma unroll
      for (int k = B - 1; k > j; --k) x[j] = x[j] - (T)D[(size_t)(B * j + k) * n + i] * x[k];
      x[j] = (T)D[(size_t)((B + 1) * j) * n + i] * x[j];
    }
  }
}

// Block-Jacobi apply with the block's B*B entries loaded up front (so their
// loads can be issued with a row kernel's other loads), for B <= 4; larger
// blocks stream them in diag_apply.  Same arithmetic order as diag_apply.
template <typename T, typename TD, int B, bool ON>
struct DiagRegs {
  static constexpr bool PRE = ON && B <= 4;
  TD d[PRE ? B * B : 1];
  __device__ __forceinline__ void load(const TD* __restrict__ D, int nrows, int i) {
    if constexpr (PRE) {
#pragma unroll
      for (int e = 0; e < B * B; ++e) d[e] = D[(size_t)e * nrows + i];
    }
  }
  __device__ __forceinline__ void apply(const TD* __restrict__ D, int nrows, int i, T (&x)[B]) const {
    if constexpr (!ON) {
      return;
    } else if constexpr (!PRE) {
      diag_apply<T, TD, B>(D, nrows, i, x);
    } else if constexpr (B == 1) {
      x[0] = (T)d[0] * x[0];
    } else if constexpr (B == 3) {
      T x1 = x[0], x2 = x[1], x3 = x[2];
      x2 = x2 - (T)d[3] * x1;
      x3 = x3 - (T)d[6] * x1 - (T)d[7] * x2;
      x3 = (T)d[8] * x3;
      x2 = (T)d[4] * (x2 - (T)d[5] * x3);
      x1 = (T)d[0] * (x1 - (T)d[2] * x3 - (T)d[1] * x2);
      x[0] = x1;
      x[1] = x2;
      x[2] = x3;
    } else {
#pragma unroll
      for (int j = 1; j < B; ++j)
#pragma unroll
        for (int k = 0; k < j; ++k) x[j] = x[j] - (T)d[B * j + k] * x[k];
#pragma unroll
      for (int j = B - 1; j >= 0; --j) {
#pragma unroll
        for (int k = B - 1; k > j; --k) x[j] = x[j] - (T)d[B * j + k] * x[k];
        x[j] = (T)d[(B + 1) * j] * x[j];
      }
    }
  }
};

// Deferred reduction finalize.  An Op may take the partial sums that the
// previous kernel on the stream stored, instead of a one-workgroup finalize
// kernel launched between the two.  Such an Op has `static constexpr int
// NPRE > 0` and members `pre`, `npre` (NPRE values per tile) and `pre2`,
// `npre2` (optional: one value per tile that replaces the last of the NPRE
// sums).  With a non-null `pre`, every workgroup sums them in the fixed order
// of sum_partials (so every workgroup derives the same bits) and calls
// `bool prologue(const double (&tot)[NPRE], bool writer)`.  The writer
// (logical workgroup 0, thread 0) records the derived scalars in the solver
// state; false ends the launch (the solve is done).  The kernel boundary makes
// the previous kernel's partials visible, so no device-scope fence is needed.
template <class Op, class = void>
struct op_npre {
  static constexpr int value = 0;
};
template <class Op>
struct op_npre<Op, std::void_t<decltype(Op::NPRE)>> {
  static constexpr int value = Op::NPRE;
};

template <int NT, class Op>
__device__ __forceinline__ bool run_prologue(Op& op, int wg) {
  constexpr int NP = op_npre<Op>::value;
  if constexpr (NP > 0) {
    if (op.pre == nullptr) return true;
    double tot[NP];
    sum_partials<NP, NT>(op.pre, op.npre, tot);
    if (op.pre2) {
      double t1[1];
      sum_partials<1, NT>(op.pre2, op.npre2, t1);
      tot[NP - 1] = t1[0];
    }
    return op.prologue(tot, wg == 0 && threadIdx.x == 0);
  } else {
    (void)op;
    (void)wg;
    return true;
  }
}

// Generic fused SpMV kernel: one lane per block row, one wavefront per 64-row
// slice at a time.  With spw == 0 wavefront w of the grid handles slice w; with
// spw == 1 the nwg*4 wavefronts split the slices into contiguous ranges.
// Logical workgroups are laid out XCD by XCD (xcd_tile), so each XCD streams one
// contiguous part of the matrix.  `Op` supplies
//   bool skip()                                  whole launch is a no-op (solver done)
//   void begin()                                 per-thread setup (reads device scalars)
//   void gather(int col, TA (&xv)[B])            input vector at a block column
//   void epilogue(int row, bool valid, const TA (&acc)[B], double (&part)[NV],
//                 const WaveRows<TA, B>& io)     (every lane calls it; vectors via io)
//   void finalize(const double (&tot)[NV])       run once by k_finalize (thread 0)
// and NV >= 0 partial sums per workgroup, stored for the finalize kernel.
template <typename TV, typename TA, int B, int U, class Op>
__global__ void __launch_bounds__(kBlock) k_sell(SellView<TV> M, Op op0, double* partials, int nwg, int spw) {
  const int wg = xcd_tile<false>(nwg);
  if (wg < 0) return;
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  if (!run_prologue<kBlock>(op, wg)) return;
  const int lane = threadIdx.x & (kWave - 1);
  const int lw = __builtin_amdgcn_readfirstlane(wg * kWpb + (int)(threadIdx.x / kWave));
  int s0 = lw, s1 = lw + 1;
  if (spw) {
    const long long W = (long long)nwg * kWpb;
    s0 = (int)((long long)lw * M.nslices / W);
    s1 = (int)((long long)(lw + 1) * M.nslices / W);
  }
  if (s1 > M.nslices) s1 = M.nslices;
  constexpr int NV = Op::NV;
  double part[NV > 0 ? NV : 1];
#pragma unroll
  for (int k = 0; k < (NV > 0 ? NV : 1); ++k) part[k] = 0.0;
  __shared__ TA lds[kWpb * kWave * B];
  for (int s = s0; s < s1; ++s) {
    const int row = s * kWave + lane;
    const bool valid = row < M.nrows;
    // The row length is loaded unconditionally (lanes past the last row read
    // the last row's offsets and are masked): a load under a branch makes the
    // compiler wait for it at the branch's end, which would stall every slice
    // for a full memory latency before its first matrix load.
    const int rowc = valid ? row : M.nrows - 1;
    const RowLen len{M.rowptr[rowc], M.rowptr[rowc + 1], valid};
    TA acc[B];
#pragma unroll
    for (int r = 0; r < B; ++r) acc[r] = (TA)0;
    auto xg = [&](int c, TA(&xv)[B]) { op.gather(c, xv); };
    if constexpr (U >= kSellPipe)
      sell_row_pipe<TV, TA, B, SellCfg<B, U - kSellPipe>::U, true>(M, s, lane, len, xg, acc);
    else
      sell_row<TV, TA, B, SellCfg<B, U>::U, true>(M, s, lane, len, xg, acc);
    const WaveRows<TA, B> io{lds + (threadIdx.x / kWave) * kWave * B, s * kWave, M.nrows, lane};
    op.epilogue(row, valid, acc, part, io);
  }
  if constexpr (NV > 0) {
    block_sum<NV>(part);
    store_partials<NV>(part, partials, wg);
  }
}

// Dispatches a runtime unroll to the k_sell instantiation.
#define MH_SELL_UNROLL(u, ...)                                      \
  switch (u) {                                                      \
    case 1: { constexpr int U = 1; __VA_ARGS__; } break;            \
    case 2: { constexpr int U = 2; __VA_ARGS__; } break;            \
    case 4: { constexpr int U = 4; __VA_ARGS__; } break;            \
    case kSellPipe + 1: { constexpr int U = kSellPipe + 1; __VA_ARGS__; } break; \
    default: { constexpr int U = 3; __VA_ARGS__; } break;           \
  }

// Row-parallel (non-SpMV) fused vector kernel over tiles of NT block rows: in
// a tile a thread takes one block row (wavefront w the rows NT*t + 64w .. +63).
// The grid is persistent -- nwg logical workgroups, workgroup g taking the
// contiguous tiles [g*ntiles/nwg, (g+1)*ntiles/nwg) -- so a prologue runs once
// per workgroup and a launch leaves nwg partial sets.  Rows of up to 24 bytes
// (B <= 3 in FP64) run 1024-thread workgroups; wider rows keep 256 threads,
// which bounds the LDS staging (64 * B entries per wavefront).
//   using VT (vector type) and BS (block size);
//   bool skip(); void begin();
//   void row(int i, bool valid, double (&part)[NV], const WaveRows<VT, BS>& io)
//       (every lane calls it; per-row vectors via io)
//   void finalize(...)
template <typename VT, int BS>
__host__ __device__ constexpr int row_block() {
  return BS * sizeof(VT) <= 24 ? 1024 : kBlock;
}
template <typename VT, int BS>
__host__ __device__ constexpr int rows_tiles(long long nrows) {
  return (int)((nrows + row_block<VT, BS>() - 1) / row_block<VT, BS>());
}

template <class Op>
__host__ __device__ constexpr int op_row_block() {
  return row_block<typename Op::VT, Op::BS>();
}

// An Op may ask for `static constexpr int MIN_WAVES` wavefronts per SIMD: 8 (at
// most 64 VGPRs) keeps two 1024-thread workgroups on a CU where a row body
// with all of its loads in flight first would otherwise take a few more VGPRs
// and halve the resident waves.
template <class Op, class = void>
struct op_min_waves {
  static constexpr int value = 1;
};
template <class Op>
struct op_min_waves<Op, std::void_t<decltype(Op::MIN_WAVES)>> {
  static constexpr int value = Op::MIN_WAVES;
};

template <class Op>
__global__ void __launch_bounds__(op_row_block<Op>())
    __attribute__((amdgpu_waves_per_eu(op_row_block<Op>() == 1024 ? op_min_waves<Op>::value : 1, 8)))
    k_rows(int nrows, Op op0, double* partials, int nwg, int ntiles) {
  using VT = typename Op::VT;
  constexpr int BS = Op::BS;
  constexpr int NT = row_block<VT, BS>();
  const int g = xcd_tile<true>(nwg);
  if (g < 0) return;
  Op op = op0;
  if (op.skip()) return;
  op.begin();
  if (!run_prologue<NT>(op, g)) return;
  __shared__ VT lds[NT * BS];
  constexpr int NV = Op::NV;
  double part[NV > 0 ? NV : 1];
#pragma unroll
  for (int k = 0; k < (NV > 0 ? NV : 1); ++k) part[k] = 0.0;
  const int lane = (int)(threadIdx.x & (kWave - 1));
  const int wofs = (int)(threadIdx.x / kWave) * kWave;
  const int t0 = (int)((long long)g * ntiles / nwg), t1 = (int)((long long)(g + 1) * ntiles / nwg);
  for (int tile = t0; tile < t1; ++tile) {
    const int row0 = tile * NT + wofs;
    if (row0 < nrows) {
      const WaveRows<VT, BS> io{lds + wofs * BS, row0, nrows, lane};
      op.row(row0 + lane, row0 + lane < nrows, part, io);
    }
  }
  if constexpr (NV > 0) {
    block_sum<NV, NT>(part);
    store_partials<NV>(part, partials, g);
  }
}

// The prologue of an Op alone (one workgroup of kBlock threads): consumes the
// partials of the last kernel of a loop when no further kernel follows.
template <int NT, class Op>
__global__ void __launch_bounds__(NT) k_prologue_only(Op op) {
  if (op.skip()) return;
  op.begin();
  run_prologue<NT>(op, 0);
}

// Second phase of a reduction: one workgroup, fixed-order sum, op.finalize().
template <class Op>
__global__ void __launch_bounds__(kBlock) k_finalize(Op op, const double* partials, int ntiles) {
  if (op.skip()) return;  // every thread reads the same flag: uniform
  constexpr int NV = Op::NV;
  double tot[NV];
  sum_partials<NV>(partials, ntiles, tot);
  if (threadIdx.x == 0) op.finalize(tot);
}

}  // namespace mh
