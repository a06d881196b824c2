// Shared device/host helpers for the monolis HIP path (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>

namespace mh {

constexpr int kWave = 64;     // CDNA wavefront width
constexpr int kBlock = 256;   // 4 wavefronts per workgroup
constexpr int kWpb = kBlock / kWave;
constexpr int kXcd = 8;       // MI355X: 8 XCDs, workgroups dealt round-robin

// ---------------------------------------------------------------------------
// Host-side error state.  Every internal entry point returns 0 on success or a
// negative code; the message of the last failure is kept per thread.
// ---------------------------------------------------------------------------
void set_error(const char* fmt, ...);
const char* last_error();
void clear_error();

#define MH_HIP(expr)                                                                 \
  do {                                                                               \
    hipError_t mh_e_ = (expr);                                                       \
    if (mh_e_ != hipSuccess) {                                                       \
      ::mh::set_error("%s:%d: %s failed: %s", __FILE__, __LINE__, #expr,              \
                      hipGetErrorString(mh_e_));                                     \
      return -1;                                                                     \
    }                                                                                \
  } while (0)

#define MH_TRY(expr)              \
  do {                            \
    int mh_rc_ = (expr);          \
    if (mh_rc_ != 0) return mh_rc_; \
  } while (0)

// Launch geometry with optional XCD-aware remapping: the grid is padded to a
// multiple of 8 workgroups and, with REMAP, logical tile t = (xcd, k) is given
// to hardware workgroup k*8 + xcd, so each XCD (and its private L2) walks one
// contiguous eighth of the rows.  Without it consecutive tiles are dealt
// round-robin to the XCDs, so all XCDs stream neighbouring parts of the matrix
// at the same time.  Measured on the 100^3 BCSR3 SpMV the round-robin order is
// 2.6 % faster (333 vs 342 us; the gathered x window is then shared by all
// XCDs and served from MALL), so the SELL kernels use it; the row kernels
// keep the remap (no difference measured).
inline int padded_grid(long long nwg) {
  long long g = (nwg + kXcd - 1) / kXcd * kXcd;
  return (int)(g > 0 ? g : kXcd);
}

template <bool REMAP = true>
__device__ __forceinline__ int xcd_tile(int nwg_logical) {
  const int b = blockIdx.x;
  if constexpr (!REMAP) return b < nwg_logical ? b : -1;
  const int g = gridDim.x;            // multiple of kXcd
  const int per = g / kXcd;
  const int t = (b % kXcd) * per + b / kXcd;
  return t < nwg_logical ? t : -1;
}

// ---------------------------------------------------------------------------
// Reductions.  All sums are deterministic: fixed butterfly order inside a
// wavefront, fixed wave order inside a workgroup, and a fixed-order final pass
// over per-workgroup partials done by the last workgroup to arrive.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
  return v;
}

__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    int w = __shfl_xor(v, o, kWave);
    v = v > w ? v : w;
  }
  return v;
}

// Sum NV values over a workgroup of NT threads; result valid in thread 0.
template <int NV, int NT = kBlock>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
  constexpr int NW = NT / kWave;
  __shared__ double red[NW][NV];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) red[w][k] = v[k];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = red[0][k];
      for (int i = 1; i < NW; ++i) s += red[i][k];
      v[k] = s;
    }
  }
  __syncthreads();
}

// Cross-workgroup reductions are two-phase: every workgroup stores its NV
// partial sums with plain stores, and a one-workgroup finalize kernel launched
// next on the same stream sums them in a fixed order.  The kernel boundary
// provides the visibility across XCDs; no device-scope fence is executed inside
// a kernel (on a multi-XCD GPU that fence writes back and invalidates the XCD's
// L2, which would also evict the gathered vector).
template <int NV>
__device__ __forceinline__ void store_partials(const double (&v)[NV], double* partials, int tile) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)tile * NV + k] = v[k];
  }
}

// Fixed-order sum of ntiles x NV partials by one workgroup of NT threads;
// valid in every thread.  A pass covers NT * PER tiles (4096 for NV <= 2 with
// 256 threads, 8192 with 1024: the 8192 slices' partials of a scalar SpMV on
// 2 M rows in one pass) and
// issues all of its loads before the first add (out-of-range slots read the
// last tile and add 0), so a sum of a few thousand tiles costs about one
// memory latency: kernels that take the previous kernel's partials in their
// prologue run it in every workgroup.  The order depends on NT, so every
// derivation of one scalar uses the same NT.
template <int NV, int NT = kBlock>
__device__ __forceinline__ void sum_partials(const double* partials, int ntiles, double (&total)[NV]) {
  constexpr int PER = NV <= 2 ? (NT > kBlock ? 8 : 16) : (NT > kBlock ? 4 : 8);
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (int base = 0; base < ntiles; base += NT * PER) {
    double v[PER][NV];
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int t = base + u * NT + (int)threadIdx.x;
      const int tc = t < ntiles ? t : ntiles - 1;
#pragma unroll
      for (int k = 0; k < NV; ++k) v[u][k] = partials[(size_t)tc * NV + k];
    }
#pragma unroll
    for (int u = 0; u < PER; ++u) {
      const int t = base + u * NT + (int)threadIdx.x;
#pragma unroll
      for (int k = 0; k < NV; ++k) acc[k] += t < ntiles ? v[u][k] : 0.0;
    }
  }
  block_sum<NV, NT>(acc);
  __shared__ double tot[NV];
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) tot[k] = acc[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NV; ++k) total[k] = tot[k];
}

// ---------------------------------------------------------------------------
// Complex FP64 value (monolis complex(kdouble): real part, then imaginary).
// Trivially default-constructible so it can live in LDS; the constructor from
// a real value is explicit so no real code converts into it by accident.
// ---------------------------------------------------------------------------
struct alignas(16) zd {
  double re, im;
  zd() = default;
  __host__ __device__ constexpr explicit zd(double r, double i = 0.0) : re(r), im(i) {}
};
__host__ __device__ __forceinline__ zd operator+(zd a, zd b) { return zd(a.re + b.re, a.im + b.im); }
__host__ __device__ __forceinline__ zd operator-(zd a, zd b) { return zd(a.re - b.re, a.im - b.im); }
__host__ __device__ __forceinline__ zd operator-(zd a) { return zd(-a.re, -a.im); }
__host__ __device__ __forceinline__ zd operator*(zd a, zd b) {
  return zd(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re);
}
// Smith's algorithm for a / b (no intermediate overflow for |b| near the range ends).
__host__ __device__ __forceinline__ zd zdiv(zd a, zd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return zd((a.re + a.im * r) / d, (a.im - a.re * r) / d);
  }
  const double r = b.re / b.im, d = b.im + b.re * r;
  return zd((a.re * r + a.im) / d, (a.im * r - a.re) / d);
}

// on ? a : b.  The complex form selects each part: a select of two structs
// becomes a select of their addresses, which keeps them in scratch memory.
template <typename T>
__device__ __forceinline__ T pick(bool on, T a, T b) {
  return on ? a : b;
}
__device__ __forceinline__ zd pick(bool on, zd a, zd b) { return zd(on ? a.re : b.re, on ? a.im : b.im); }

// t + a * x, the multiply-add of the SpMV row loops.  The complex form keeps
// the real and imaginary parts as two fused chains.
__device__ __forceinline__ double madd(double a, double x, double t) { return fma(a, x, t); }
__device__ __forceinline__ float madd(float a, float x, float t) { return fmaf(a, x, t); }
__device__ __forceinline__ zd madd(zd a, zd x, zd t) {
  return zd(fma(a.re, x.re, fma(-a.im, x.im, t.re)), fma(a.re, x.im, fma(a.im, x.re, t.im)));
}

// Streaming loads: matrix values and column ids are read once per SpMV, so
// they are fetched non-temporally to keep L2 / MALL for the gathered vector.
template <typename T>
__device__ __forceinline__ T ld_stream(const T* p) {
  return __builtin_nontemporal_load(p);
}
template <bool NT, typename T>
__device__ __forceinline__ T ld_val(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
using d2v = double __attribute__((ext_vector_type(2)));
template <bool NT>
__device__ __forceinline__ zd ld_val(const zd* p) {
  d2v v;
  if constexpr (NT) v = __builtin_nontemporal_load((const d2v*)p);
  else v = *(const d2v*)p;
  return zd(v.x, v.y);
}

// SELL value layout inside one slot (B*B block entries x 64 lanes).  FP64
// values are stored in pairs: entries (2q, 2q+1) of a lane side by side, so one
// 16-byte load per lane fetches both and a wavefront's pair load is one
// contiguous 1 KiB request; for an odd B*B the last entry has a plane of its
// own after the pairs.  Measured on the 100^3 BCSR3 SpMV: 325 -> 319 us (FP64).
// FP32 values keep one plane per entry: pairs of floats measured 1 % slower
// (162.5 vs 160.8 us).  16-byte complex values keep one plane per entry.
template <typename TV>
__host__ __device__ constexpr bool sell_paired() {
  return sizeof(TV) == 8;
}
// Offset of entry e of lane `lane` from the start of its slot.
template <typename TV>
__host__ __device__ __forceinline__ size_t sell_vofs(int e, int bb, int lane) {
  if constexpr (!sell_paired<TV>()) {
    return (size_t)e * kWave + lane;
  } else {
    const int np2 = bb & ~1;  // entries held in pairs
    return e < np2 ? (size_t)(e >> 1) * (2 * kWave) + 2 * lane + (e & 1) : (size_t)np2 * kWave + lane;
  }
}
// Position of entry e of slot `slot`, lane `lane`, in the value array.
template <typename TV>
__host__ __device__ __forceinline__ size_t sell_vpos(size_t slot, int e, int bb, int lane) {
  return slot * (size_t)bb * kWave + sell_vofs<TV>(e, bb, lane);
}

// Loads the B*B entries of one slot for lane `lane` (slot: start of the slot).
template <int BB, bool NT, typename TV>
__device__ __forceinline__ void ld_slot(const TV* __restrict__ slot, int lane, TV (&a)[BB]) {
  if constexpr (sell_paired<TV>()) {
    using P2 = TV __attribute__((ext_vector_type(2)));
    constexpr int NP = BB / 2;
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      const P2* p = (const P2*)(slot + (size_t)q * (2 * kWave)) + lane;
      P2 v;
      if constexpr (NT) v = __builtin_nontemporal_load(p);
      else v = *p;
      a[2 * q] = v.x;
      a[2 * q + 1] = v.y;
    }
    if constexpr (BB & 1) a[BB - 1] = ld_val<NT>(slot + (size_t)(2 * NP) * kWave + lane);
  } else {
#pragma unroll
    for (int e = 0; e < BB; ++e) a[e] = ld_val<NT>(slot + (size_t)e * kWave + lane);
  }
}

// Non-temporal store of one element (16-byte complex values as one vector store).
template <typename T>
__device__ __forceinline__ void st_stream(T* p, T v) {
  if constexpr (sizeof(T) == 16) {
    d2v w;
    w.x = v.re;
    w.y = v.im;
    __builtin_nontemporal_store(w, (d2v*)p);
  } else {
    __builtin_nontemporal_store(v, p);
  }
}

}  // namespace mh
