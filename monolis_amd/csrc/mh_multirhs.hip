// Multi-right-hand-side CG: k <= 64 independent monolis CG solves
// (src/iterative/CG.f90:18-119) on one scalar matrix, sharing every SpMM pass.
//
// Vectors are row-major np x k (entry (i, j) at i*k + j: right-hand sides
// innermost).  A wavefront owns one 64-row SELL slice and maps its lanes to
// right-hand sides: the slice's column ids and values are wavefront-uniform
// (scalar loads) and every gather of x(col, :) is one contiguous k-element load,
// so the matrix streams from HBM once per SpMM instead of once per right-hand
// side.  Each right-hand side keeps its own rho, alpha, beta, residual,
// convergence flag and iteration count on the device; once it converges its x
// and r stay frozen, as if its own solve had returned.  The launch loop stops
// when every right-hand side is done.
#include <algorithm>
#include <cmath>
#include <vector>

#include "mh_multirhs.hpp"

namespace mh {

int cu_count();

namespace {

constexpr int kIterRR = 200;       // true-residual replacement period (CG.f90:97)

struct RhsState {
  double rho, alpha, beta, B2, resid;
  int done, conv, iter, zero;
};

struct MState {
  int all_done;
  int pad[63];
};

// Lane layout: kk = 2^ks >= k lanes per row, 64/kk rows per wavefront step;
// lane = rsub * kk + j handles row (step base + rsub), right-hand side j (< k).
struct Lanes {
  int ks, kk, rpw, rsub, j;
  bool on;
  __device__ Lanes(int k, int ks_) : ks(ks_), kk(1 << ks_), rpw(kWave >> ks_) {
    const int lane = threadIdx.x & (kWave - 1);
    rsub = lane >> ks;
    j = lane & (kk - 1);
    on = j < k;
  }
};

// Contiguous unit range of this workgroup.  Logical workgroups are numbered so
// that each XCD (blockIdx % 8) owns a contiguous part of the rows; the four
// wavefronts of a workgroup interleave over its range.
__device__ __forceinline__ void unit_range(long long units, long long* u0, long long* u1) {
  const int g = gridDim.x;
  const int per = g / kXcd;
  const int t = (blockIdx.x % kXcd) * per + blockIdx.x / kXcd;
  *u0 = units * t / g;
  *u1 = units * (t + 1) / g;
}

// Unit order of the SpMM: each XCD owns a contiguous part of the units, every
// workgroup a contiguous sub-range of it, its wavefronts interleaved.  (An
// order that interleaves an XCD's workgroups unit by unit over one narrow band
// fetched 8 % fewer bytes from beyond L2 and ran in the same time.)
template <int NW = kWpb>
__device__ __forceinline__ void spmm_units(long long units, int w, long long* first, long long* stride,
                                           long long* end) {
  long long u0, u1;
  unit_range(units, &u0, &u1);
  *first = u0 + w;
  *stride = NW;
  *end = u1;
}

// Per right-hand side sum of the lanes' partials over the workgroup (fixed
// order: wavefronts, then row sub-slots); wave 0 writes 64 slots per value.
template <int NV>
__device__ __forceinline__ void rhs_partials(int kk, double (&v)[NV], double* __restrict__ partials) {
  __shared__ double red[NV][kWpb][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int q = 0; q < NV; ++q) red[q][w][lane] = v[q];
  __syncthreads();
  if (w == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double a = 0.0;
      if (lane < kk)
        for (int u = 0; u < kWpb; ++u)
          for (int r = 0; r < kWave / kk; ++r) a += red[q][u][r * kk + lane];
      partials[((size_t)blockIdx.x * NV + q) * kWave + lane] = a;
    }
  }
}

// Per right-hand side sum over the workgroup of the SpMM's partials when the
// right-hand sides are processed in NCH chunks of KK: v[ch] of lane
// (row sub-slot r, j) belongs to right-hand side ch*KK + j.  Fixed order
// (wavefronts, then row sub-slots); wave 0 writes 64 slots, lane L = right-hand
// side L (0 past k).  With one chunk this is rhs_partials<1>.
template <int NCH, int KK, int NW = kWpb>
__device__ __forceinline__ void rhs_partials_chunked(int k, const double (&v)[NCH], double* __restrict__ partials) {
  __shared__ double red[NCH][NW][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) red[ch][w][lane] = v[ch];
  __syncthreads();
  if (w == 0) {
    const int ch = lane / KK, j = lane % KK;
    double a = 0.0;
    if (lane < k)
      for (int u = 0; u < NW; ++u)
        for (int r = 0; r < kWave / KK; ++r) a += red[ch][u][r * KK + j];
    partials[(size_t)blockIdx.x * kWave + lane] = a;
  }
}

// Broadcast of lane `src` of v to the wavefront (uniform src: a readlane).
template <bool UNIFORM>
__device__ __forceinline__ int bcast(int v, int src) {
  if constexpr (UNIFORM) return __builtin_amdgcn_readlane(v, src);
  else return __shfl(v, src, kWave);
}
template <bool UNIFORM>
__device__ __forceinline__ double bcast(double v, int src) {
  if constexpr (UNIFORM) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), src);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), src);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
  } else {
    return __shfl(v, src, kWave);
  }
}
template <bool UNIFORM>
__device__ __forceinline__ float bcast(float v, int src) {
  return __int_as_float(bcast<UNIFORM>(__float_as_int(v), src));
}

// SpMM over units of 8 rows of one SELL slice, lane = (row sub-slot, right-hand
// side) with KK = 2^ks (>= 8) lanes per row.  The right-hand sides are taken in
// chunks of KK columns (one chunk when KK >= k): each workgroup walks its unit
// range once per chunk, so the rows of X gathered while a chunk is in flight
// are KK * sizeof(T) bytes wide and the neighbour planes of a chunk fit in the
// XCD's L2, at the price of reading the matrix once per chunk.  For each chunk
// of 8 slots one vector load brings the unit's 8 x 8 column ids and one the
// values (lane q*8 + r holds slot q of row r); every row then broadcasts its
// entries from those lanes and issues all its gathers x(col, :) back to back,
// so no gather waits on a dependent index load.  Padded slots (col = own row,
// val = 0) and slots past the slice width (clamped to the last slot) are
// masked by the row's length with a select.
//   MODE 0: y = A x
//   MODE 1: y = A x and partial (x, y) per right-hand side      (Q = A P, (P, Q))
//   MODE 2: y = b - A x and partial (y, y) per right-hand side  (R = B - A X, (R, R))
template <typename T, int MODE, int KK, int SB, int NT = kBlock>
__global__ void __launch_bounds__(NT) k_spmm(int n, int k, int nslices, const int* __restrict__ slice_ptr,
                                                 const int* __restrict__ rowptr, const int* __restrict__ col,
                                                 const T* __restrict__ val, const T* __restrict__ x,
                                                 const T* __restrict__ b, T* __restrict__ y,
                                                 double* __restrict__ partials, const MState* ms) {
  if (ms && ms->all_done) return;
  constexpr int RPW = kWave / KK;              // rows per step
  constexpr int STEPS = 8 / RPW;               // steps per unit of 8 rows
  constexpr int BS = SB < STEPS ? SB : STEPS;  // steps whose gathers are in flight together
  constexpr bool UNI = RPW == 1;
  constexpr int NCH = kWave / KK;              // chunks of KK right-hand sides (k <= 64)
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int rsub = lane / KK, j = lane % KK;
  const int lt = lane >> 3, lr = lane & 7;  // slot / row of the loaded tile entry
  double part[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) part[ch] = 0.0;
  long long ufirst, ustride, uend;
  spmm_units<NT / kWave>((long long)nslices * 8, w, &ufirst, &ustride, &uend);
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    const int c0 = ch * KK;
    if (c0 >= k) break;
    const bool on = c0 + j < k;
    const int xj = on ? c0 + j : c0;
    for (long long u = ufirst; u < uend; u += ustride) {
      const int s = (int)(u >> 3);
      const int rbase = (int)(u & 7) * 8;
      const int row0 = s * kWave + rbase;
      if (row0 >= n) continue;
      const int k0 = slice_ptr[s];
      const int wd = slice_ptr[s + 1] - k0;
      const int li = row0 + lr;
      const int lenv = (lane < 8 && li < n) ? rowptr[li + 1] - rowptr[li] : 0;
      int len[STEPS];
#pragma unroll
      for (int st = 0; st < STEPS; ++st) len[st] = bcast<UNI>(lenv, st * RPW + rsub);
      T acc[STEPS];
#pragma unroll
      for (int st = 0; st < STEPS; ++st) acc[st] = (T)0;
      for (int t0 = 0; t0 < wd; t0 += 8) {
        const int t = min(t0 + lt, wd - 1);
        const size_t slot = (size_t)(k0 + t) * kWave + rbase + lr;
        const int cv = col[slot];
        const T av = val[slot];
#pragma unroll
        for (int sb = 0; sb < STEPS; sb += BS) {
          T xv[BS][8];
#pragma unroll
          for (int st = 0; st < BS; ++st)
#pragma unroll
            for (int q = 0; q < 8; ++q)
              xv[st][q] = x[(size_t)bcast<UNI>(cv, q * 8 + (sb + st) * RPW + rsub) * k + xj];
#pragma unroll
          for (int st = 0; st < BS; ++st)
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              const T a = bcast<UNI>(av, q * 8 + (sb + st) * RPW + rsub);
              acc[sb + st] = (t0 + q < len[sb + st]) ? fma(a, xv[st][q], acc[sb + st]) : acc[sb + st];
            }
        }
      }
#pragma unroll
      for (int st = 0; st < STEPS; ++st) {
        const int i = row0 + st * RPW + rsub;
        if (i < n && on) {
          const size_t e = (size_t)i * k + c0 + j;
          if constexpr (MODE == 2) {
            const T rv = b[e] - acc[st];
            y[e] = rv;
            part[ch] += (double)rv * (double)rv;
          } else {
            y[e] = acc[st];
            if constexpr (MODE == 1) part[ch] += (double)x[e] * (double)acc[st];
          }
        }
      }
    }
  }
  if constexpr (MODE != 0) rhs_partials_chunked<NCH, KK, NT / kWave>(k, part, partials);
}

// One (column id, value) entry of a wavefront's 8 x 8 tile in LDS.
template <typename T>
struct TileEnt;
template <>
struct alignas(16) TileEnt<double> {
  int c, pad;
  double v;
};
template <>
struct alignas(8) TileEnt<float> {
  int c;
  float v;
};

// Buffer resource over x (word 3 = 0x00020000: 32-bit data format, gfx9):
// lanes whose byte offset is at or past `bytes` load 0 without a memory access.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t x_rsrc(const void* x, unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(x), (short)0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ double buf_ld(__amdgpu_buffer_rsrc_t r, unsigned off, double) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float buf_ld(__amdgpu_buffer_rsrc_t r, unsigned off, float) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}

// k_spmm with the tile's column ids and values staged through LDS (two LDS reads
// per gathered entry instead of three cross-lane shuffles) and x read through a
// buffer resource with 32-bit byte offsets (x must be under 4 GiB).  Entries past
// a row's length get an offset past the end of x: they load 0 without touching
// memory and their FMA is skipped, so every row sums exactly its own entries in
// slot order, as k_spmm does.
template <typename T, int MODE, int KK, int SB, int NT>
__global__ void __launch_bounds__(NT) k_spmm_b(int n, int k, int nslices, const int* __restrict__ slice_ptr,
                                                   const int* __restrict__ rowptr, const int* __restrict__ col,
                                                   const T* __restrict__ val, const T* __restrict__ x,
                                                   const T* __restrict__ b, T* __restrict__ y,
                                                   double* __restrict__ partials, const MState* ms) {
  if (ms && ms->all_done) return;
  constexpr int RPW = kWave / KK;
  constexpr int STEPS = 8 / RPW;
  constexpr int BS = SB < STEPS ? SB : STEPS;
  constexpr int NCH = kWave / KK;
  constexpr int NW = NT / kWave;
  __shared__ TileEnt<T> tile[NW][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int rsub = lane / KK, j = lane % KK;
  const int lr = lane & 7;
  const unsigned xbytes = (unsigned)((size_t)n * k * sizeof(T));
  const __amdgpu_buffer_rsrc_t xr = x_rsrc(x, xbytes);
  const unsigned rowb = (unsigned)(k * sizeof(T));
  double part[NCH];
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) part[ch] = 0.0;
  long long ufirst, ustride, uend;
  spmm_units<NW>((long long)nslices * 8, w, &ufirst, &ustride, &uend);
#pragma unroll
  for (int ch = 0; ch < NCH; ++ch) {
    const int c0 = ch * KK;
    if (c0 >= k) break;
    const bool on = c0 + j < k;
    const unsigned jb = (unsigned)((on ? c0 + j : c0) * sizeof(T));
    for (long long u = ufirst; u < uend; u += ustride) {
      const int s = (int)(u >> 3);
      const int rbase = (int)(u & 7) * 8;
      const int row0 = s * kWave + rbase;
      if (row0 >= n) continue;
      const int k0 = slice_ptr[s];
      const int wd = slice_ptr[s + 1] - k0;
      const int li = row0 + lr;
      const int lenv = (lane < 8 && li < n) ? rowptr[li + 1] - rowptr[li] : 0;
      int len[STEPS];
#pragma unroll
      for (int st = 0; st < STEPS; ++st) len[st] = __shfl(lenv, st * RPW + rsub, kWave);
      T acc[STEPS];
#pragma unroll
      for (int st = 0; st < STEPS; ++st) acc[st] = (T)0;
      for (int t0 = 0; t0 < wd; t0 += 8) {
        {  // stage the 8 slots x 8 rows tile (slot-major, as stored)
          const int t = min(t0 + (lane >> 3), wd - 1);
          const size_t slot = (size_t)(k0 + t) * kWave + rbase + lr;
          TileEnt<T> e;
          e.c = col[slot];
          e.v = val[slot];
          __builtin_amdgcn_wave_barrier();
          tile[w][lane] = e;
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
#pragma unroll
        for (int sb = 0; sb < STEPS; sb += BS) {
          T xv[BS][8];
#pragma unroll
          for (int st = 0; st < BS; ++st)
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              const int c = tile[w][q * 8 + (sb + st) * RPW + rsub].c;
              const unsigned off = (t0 + q < len[sb + st]) ? (unsigned)c * rowb + jb : 0xffffffffu;
              xv[st][q] = buf_ld(xr, off, T());
            }
#pragma unroll
          for (int st = 0; st < BS; ++st)
#pragma unroll
            for (int q = 0; q < 8; ++q) {
              const T a = tile[w][q * 8 + (sb + st) * RPW + rsub].v;
              acc[sb + st] = (t0 + q < len[sb + st]) ? fma(a, xv[st][q], acc[sb + st]) : acc[sb + st];
            }
        }
      }
#pragma unroll
      for (int st = 0; st < STEPS; ++st) {
        const int i = row0 + st * RPW + rsub;
        if (i < n && on) {
          const size_t e = (size_t)i * k + c0 + j;
          if constexpr (MODE == 2) {
            const T rv = b[e] - acc[st];
            y[e] = rv;
            part[ch] += (double)rv * (double)rv;
          } else {
            y[e] = acc[st];
            if constexpr (MODE == 1) part[ch] += (double)x[e] * (double)acc[st];
          }
        }
      }
    }
  }
  if constexpr (MODE != 0) rhs_partials_chunked<NCH, KK, NW>(k, part, partials);
}

// Elementwise steps over row groups of 64/kk rows, lane = (row, right-hand side).
// Z = M^-1 R is never stored: the diagonal scaling is recomputed where it is
// used, with the same single multiplication, so it is bit-identical.
//   MODE 0: P = Z + beta P                                        (CG.f90:80-84)
//   MODE 1: X += alpha P; R -= alpha Q; (R,R), (R, M^-1 R)        (CG.f90:90-100, next precond)
//   MODE 2: X += alpha P                                          (replacement iteration)
//   MODE 3: (R,R), (R, M^-1 R)
// Updates of X, R and P touch only right-hand sides that are not done.
// Streams of the elementwise passes: every vector entry is read and written
// once per pass, so FP64 loads and stores bypass the caches (non-temporal).
// Measured on the 128^3 x 64-RHS CG: FP64 1460 -> 1409 ms; the same hint on
// the FP32 passes of the refined solve cost 4 %, so they stay plain.  Four row
// groups per wavefront step are loaded together (eight measured no faster).
constexpr int kMrowsUnroll = 4;
template <typename T>
constexpr bool mrows_nt() {
  return sizeof(T) == 8;
}
template <typename T>
__device__ __forceinline__ T ld_vec(const T* p) {
  if constexpr (mrows_nt<T>()) return __builtin_nontemporal_load(p);
  else return *p;
}
template <typename T>
__device__ __forceinline__ void st_vec(T* p, T v) {
  if constexpr (mrows_nt<T>()) __builtin_nontemporal_store(v, p);
  else *p = v;
}

template <typename T, int MODE>
__global__ void __launch_bounds__(kBlock) k_mrows(int n, int k, int ks, T* __restrict__ x, T* __restrict__ r,
                                                  T* __restrict__ p, const T* __restrict__ q,
                                                  const T* __restrict__ dinv, double* __restrict__ partials,
                                                  const MState* ms, const RhsState* __restrict__ rs) {
  if (ms->all_done) return;
  const Lanes L(k, ks);
  const int w = threadIdx.x / kWave;
  bool active = false;
  T alpha = (T)0, beta = (T)0;
  if (L.on) {
    active = !rs[L.j].done;
    alpha = (T)rs[L.j].alpha;
    beta = (T)rs[L.j].beta;
  }
  double part[2] = {0.0, 0.0};
  const long long groups = (n + L.rpw - 1) / L.rpw;
  long long u0, u1;
  unit_range(groups, &u0, &u1);
  constexpr int UR = kMrowsUnroll;  // row groups per wavefront step, loads issued together
  if (L.on) {
    for (long long g0 = u0 + (long long)w * UR; g0 < u1; g0 += (long long)kWpb * UR) {
      size_t e[UR];
      long long row[UR];
      bool ok[UR];
#pragma unroll
      for (int uu = 0; uu < UR; ++uu) {
        const long long i = (g0 + uu) * L.rpw + L.rsub;
        ok[uu] = g0 + uu < u1 && i < n;
        row[uu] = ok[uu] ? i : 0;
        e[uu] = (size_t)row[uu] * k + L.j;
      }
      if constexpr (MODE == 0) {
        T pv[UR], rv[UR], dv[UR];
#pragma unroll
        for (int uu = 0; uu < UR; ++uu) {
          pv[uu] = ld_vec(p + e[uu]);
          rv[uu] = ld_vec(r + e[uu]);
          dv[uu] = dinv ? dinv[row[uu]] : (T)1;
        }
#pragma unroll
        for (int uu = 0; uu < UR; ++uu) {
          const T zz = dinv ? dv[uu] * rv[uu] : rv[uu];  // diag_nn apply with B = 1 (diag_nn.f90:207-228)
          if (active && ok[uu]) st_vec(p + e[uu], fma(beta, pv[uu], zz));
        }
      } else if constexpr (MODE == 2) {
        T pv[UR], xv[UR];
#pragma unroll
        for (int uu = 0; uu < UR; ++uu) {
          pv[uu] = ld_vec(p + e[uu]);
          xv[uu] = ld_vec(x + e[uu]);
        }
#pragma unroll
        for (int uu = 0; uu < UR; ++uu)
          if (active && ok[uu]) st_vec(x + e[uu], fma(alpha, pv[uu], xv[uu]));
      } else {
        T rv[UR], dv[UR], xv[UR], pv[UR], qv[UR];
#pragma unroll
        for (int uu = 0; uu < UR; ++uu) {
          rv[uu] = ld_vec(r + e[uu]);
          dv[uu] = dinv ? dinv[row[uu]] : (T)1;
          if constexpr (MODE == 1) {
            xv[uu] = ld_vec(x + e[uu]);
            pv[uu] = ld_vec(p + e[uu]);
            qv[uu] = ld_vec(q + e[uu]);
          }
        }
#pragma unroll
        for (int uu = 0; uu < UR; ++uu) {
          if (!ok[uu]) continue;
          if constexpr (MODE == 1) {
            if (active) {
              st_vec(x + e[uu], fma(alpha, pv[uu], xv[uu]));
              rv[uu] = fma(-alpha, qv[uu], rv[uu]);
              st_vec(r + e[uu], rv[uu]);
            }
          }
          const T zz = dinv ? dv[uu] * rv[uu] : rv[uu];
          part[0] += (double)rv[uu] * (double)rv[uu];
          part[1] += (double)rv[uu] * (double)zz;
        }
      }
    }
  }
  if constexpr (MODE == 1 || MODE == 3) rhs_partials<2>(L.kk, part, partials);
}

// Threads of the one-workgroup finalize kernels.
constexpr int kFinal = 1024;

// Sums of the NV values over the ntiles per-workgroup partials for every
// right-hand side.  Wavefront w of the (kFinal-thread) workgroup sums a
// contiguous block of tiles with 8 interleaved accumulators; blocks are then
// combined in wavefront order.  Fixed order for a given ntiles.  Valid in
// every wavefront, lane j = right-hand side j.
template <int NV>
__device__ __forceinline__ void tile_sums(int ntiles, const double* __restrict__ partials, double (&out)[NV]) {
  constexpr int NW = kFinal / kWave;
  __shared__ double red[NV][NW][kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave;
  const int per = (ntiles + NW - 1) / NW;
  const int t0 = min(ntiles, w * per), t1 = min(ntiles, t0 + per);
  double a[NV][8];
#pragma unroll
  for (int q = 0; q < NV; ++q)
#pragma unroll
    for (int u = 0; u < 8; ++u) a[q][u] = 0.0;
  int t = t0;
  for (; t + 8 <= t1; t += 8)
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int q = 0; q < NV; ++q) a[q][u] += partials[((size_t)(t + u) * NV + q) * kWave + lane];
  for (; t < t1; ++t)
#pragma unroll
    for (int q = 0; q < NV; ++q) a[q][0] += partials[((size_t)t * NV + q) * kWave + lane];
#pragma unroll
  for (int q = 0; q < NV; ++q)
    red[q][w][lane] = ((a[q][0] + a[q][1]) + (a[q][2] + a[q][3])) + ((a[q][4] + a[q][5]) + (a[q][6] + a[q][7]));
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double v = red[q][0][lane];
    for (int u = 1; u < NW; ++u) v += red[q][u][lane];
    out[q] = v;
  }
}

// One workgroup; lane j of wave 0 owns right-hand side j.
//   STAGE 0: B2 = (R0, R0); zero right-hand sides are done (mat_converge.f90:38-48)
//   STAGE 1: (R,R), (R,Z): convergence test when `check` (mat_converge.f90:110-140),
//            then beta = rho / rho_prev (0 when `first`) and rho = (R, Z)
//   STAGE 2: alpha = rho / (P, Q)                                       (CG.f90:86-88)
template <int STAGE>
__global__ void __launch_bounds__(kFinal) k_mfinal(int k, int ntiles, const double* __restrict__ partials,
                                                   MState* ms, RhsState* rs, int it, int check, int first,
                                                   double tol, int max_iter) {
  if (ms->all_done) return;
  double v0, v1 = 0.0;
  if constexpr (STAGE == 1) {
    double v[2];
    tile_sums<2>(ntiles, partials, v);
    v0 = v[0];
    v1 = v[1];
  } else {
    double v[1];
    tile_sums<1>(ntiles, partials, v);
    v0 = v[0];
  }
  if (threadIdx.x >= kWave) return;
  const int j = threadIdx.x;
  int done = 1;
  if (j < k) {
    RhsState st = rs[j];
    if (STAGE == 0) {
      st.B2 = v0;
      if (v0 == 0.0) st.done = st.conv = st.zero = 1;
    } else if (!st.done) {
      if (STAGE == 2) {
        st.alpha = st.rho / v0;
      } else {
        const double rr = v0, rz = v1;
        bool stop = false;
        if (check) {
          st.resid = sqrt(rr / st.B2);
          st.iter = it;
          if (st.resid < tol) {
            st.done = st.conv = 1;
            stop = true;
          } else if (it >= max_iter) {
            st.done = 1;
            stop = true;
          }
        }
        if (!stop) {
          st.beta = first ? 0.0 : rz / st.rho;
          st.rho = rz;
        }
      }
    }
    rs[j] = st;
    done = st.done;
  }
  const bool all = __all(done);
  if (j == 0) ms->all_done = all ? 1 : 0;
}

__global__ void k_mstate_init(int k, MState* ms, RhsState* rs) {
  const int j = threadIdx.x;
  if (j < k) rs[j] = RhsState{};
  if (j == 0) ms->all_done = 0;
}

template <typename T>
__global__ void k_mfill(size_t m, T* __restrict__ v) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (size_t)gridDim.x * blockDim.x)
    v[e] = (T)0;
}

// X(:, j) = 0 for the right-hand sides whose bit is set in `mask`
// (zero right-hand sides, mat_converge.f90:40-48).
template <typename T>
__global__ void k_zero_cols(size_t m, int k, unsigned long long mask, T* __restrict__ x) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (size_t)gridDim.x * blockDim.x)
    if ((mask >> (e % k)) & 1ull) x[e] = (T)0;
}

// ---- refinement helpers --------------------------------------------------

// out[j] = sum of the per-workgroup partials of right-hand side j.
__global__ void __launch_bounds__(kFinal) k_lane_sum(int k, int ntiles, const double* __restrict__ partials,
                                                     double* __restrict__ out) {
  double v[1];
  tile_sums<1>(ntiles, partials, v);
  if (threadIdx.x < k) out[threadIdx.x] = v[0];
}

// r32 = (float) r on the right-hand sides whose bit is set in `mask`, else 0.
__global__ void k_round_masked(size_t m, int k, unsigned long long mask, const double* __restrict__ r,
                               float* __restrict__ r32) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (size_t)gridDim.x * blockDim.x)
    r32[e] = ((mask >> (e % k)) & 1ull) ? (float)r[e] : 0.0f;
}

// x += (double) c
__global__ void k_add_correction(size_t m, const float* __restrict__ c, double* __restrict__ x) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (size_t)gridDim.x * blockDim.x)
    x[e] += (double)c[e];
}

// Persistent grid of the elementwise kernels: about 4 workgroups per CU, a
// multiple of the XCD count.
int persistent_grid(long long units) {
  long long g = std::min<long long>(4LL * cu_count(), std::max<long long>(units, 1));
  return (int)((g + kXcd - 1) / kXcd * kXcd);
}

int lane_shift(int k) {
  int ks = 0;
  while ((1 << ks) < k) ++ks;
  return ks;
}

unsigned grid_for(size_t m) {
  size_t g = (m + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  return (unsigned)(g ? g : 1);
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// Launch configurations: steps whose gathers are in flight together, workgroups
// per CU, the chunk width in right-hand sides (0: all k in one chunk), the
// workgroup size and the kernel form (1: k_spmm_b).  Each workgroup walks one
// contiguous range of rows, so the rows of x an XCD gathers from at a time
// scale with its workgroups times the chunk width.  FP64 at 128^3 x 64: four
// 256-thread workgroups per CU fetched x about 4.5 times from beyond L2, one
// 1024-thread workgroup with 32-column chunks about twice.  Entries 0 and 1
// are the FP64 and FP32 defaults before autotuning (the fastest there).
// (A row form -- one wavefront per slice, rows taken R at a time with scalar
// loads of their column ids and values -- measured 1.19-1.82 ms against
// 1.12 ms for the best 256-thread configuration, FP64 at 128^3 x 64, and was dropped.)
struct SpmmCfg {
  int sb, wpc, kc, nt, lds;
};
constexpr SpmmCfg kSpmmCfg[] = {
    {2, 1, 32, 1024, 1}, {2, 2, 32, 1024, 1}, {4, 4, 32, 256, 1},  {8, 4, 32, 256, 1},
    {2, 4, 32, 256, 1},  {4, 4, 0, 256, 1},   {4, 8, 0, 256, 1},   {8, 4, 0, 256, 1},
    {2, 1, 0, 1024, 1},  {4, 4, 0, 256, 0},   {8, 4, 0, 256, 0},   {4, 8, 0, 256, 0},
    {2, 8, 0, 256, 0},   {2, 4, 32, 256, 0},  {2, 1, 32, 1024, 0}, {1, 1, 32, 1024, 0}};
constexpr int kNumSpmmCfg = sizeof(kSpmmCfg) / sizeof(kSpmmCfg[0]);

int spmm_cfg(const Matrix& M, int fp32) {
  const int c = fp32 ? M.mrhs_cfg32 : M.mrhs_cfg64;
  return (c >= 0 && c < kNumSpmmCfg) ? c : (fp32 ? 1 : 0);
}

// Persistent grid of the SpMM for configuration cfg.
int spmm_grid(const Matrix& M, int cfg) {
  const long long units = (long long)M.nslices * 8;
  long long g = std::min<long long>((long long)kSpmmCfg[cfg].wpc * cu_count(), std::max<long long>(units, 1));
  return (int)((g + kXcd - 1) / kXcd * kXcd);
}

template <typename T, int MODE, int KK, int SB, int NT, bool LDS>
void spmm_kk(const Matrix& M, int k, const T* val, const T* x, const T* b, T* y, double* parts, const MState* ms,
             int grid, hipStream_t s) {
  if constexpr (LDS)
    k_spmm_b<T, MODE, KK, SB, NT><<<grid, NT, 0, s>>>(M.n, k, M.nslices, M.slice_ptr.p, M.rowptr.p, M.col.p, val, x,
                                                      b, y, parts, ms);
  else
    k_spmm<T, MODE, KK, SB, NT><<<grid, NT, 0, s>>>(M.n, k, M.nslices, M.slice_ptr.p, M.rowptr.p, M.col.p, val, x, b,
                                                    y, parts, ms);
}

// SpMM launch: KK = max(8, 2^ks) lanes per row; steps per batch from the
// configuration, clamped to the steps a unit has (8 rows / rows per step) and,
// for 1024-thread workgroups, to 2 (1 with 64 lanes per row): the VGPR budget
// of 16 wavefronts per CU.
template <typename T, int MODE, int NT, bool LDS>
void spmm_nt(const Matrix& M, int k, int ks, int sb, const T* val, const T* x, const T* b, T* y, double* parts,
             const MState* ms, int grid, hipStream_t s) {
  constexpr bool BIG = NT > kBlock;
  switch (ks <= 3 ? 3 : ks) {
    case 3:
      spmm_kk<T, MODE, 8, 1, NT, LDS>(M, k, val, x, b, y, parts, ms, grid, s);
      break;
    case 4:
      if (sb >= 2) spmm_kk<T, MODE, 16, 2, NT, LDS>(M, k, val, x, b, y, parts, ms, grid, s);
      else spmm_kk<T, MODE, 16, 1, NT, LDS>(M, k, val, x, b, y, parts, ms, grid, s);
      break;
    case 5:
      if constexpr (!BIG)
        if (sb >= 4) return spmm_kk<T, MODE, 32, 4, NT, LDS>(M, k, val, x, b, y, parts, ms, grid, s);
      if (sb >= 2) spmm_kk<T, MODE, 32, 2, NT, LDS>(M, k, val, x, b, y, parts, ms, grid, s);
      else spmm_kk<T, MODE, 32, 1, NT, LDS>(M, k, val, x, b, y, parts, ms, grid, s);
      break;
    default:
      if constexpr (!BIG) {
        if (sb >= 8) return spmm_kk<T, MODE, 64, 8, NT, LDS>(M, k, val, x, b, y, parts, ms, grid, s);
        if (sb >= 4) return spmm_kk<T, MODE, 64, 4, NT, LDS>(M, k, val, x, b, y, parts, ms, grid, s);
        if (sb >= 2) return spmm_kk<T, MODE, 64, 2, NT, LDS>(M, k, val, x, b, y, parts, ms, grid, s);
      }
      spmm_kk<T, MODE, 64, 1, NT, LDS>(M, k, val, x, b, y, parts, ms, grid, s);
      break;
  }
}

template <typename T, int MODE>
void spmm(const Matrix& M, int k, int ks, int cfg, const T* val, const T* x, const T* b, T* y, double* parts,
          const MState* ms, int grid, hipStream_t s) {
  const SpmmCfg& C = kSpmmCfg[cfg];
  if (C.kc > 0) ks = std::min(ks, C.kc == 8 ? 3 : C.kc == 16 ? 4 : C.kc == 32 ? 5 : 6);  // lanes per row = chunk
  // the LDS form addresses x with 32-bit byte offsets and keeps 0xffffffff out of range
  const bool lds = C.lds && (size_t)M.n * k * sizeof(T) < 0xfff00000ull;
  if (C.nt == 1024) {
    if (lds) spmm_nt<T, MODE, 1024, true>(M, k, ks, C.sb, val, x, b, y, parts, ms, grid, s);
    else spmm_nt<T, MODE, 1024, false>(M, k, ks, C.sb, val, x, b, y, parts, ms, grid, s);
  } else {
    if (lds) spmm_nt<T, MODE, kBlock, true>(M, k, ks, C.sb, val, x, b, y, parts, ms, grid, s);
    else spmm_nt<T, MODE, kBlock, false>(M, k, ks, C.sb, val, x, b, y, parts, ms, grid, s);
  }
}

}  // namespace

template <typename T>
int multi_cg(Matrix& M, const MultiParams& P, int k, const T* b, T* x, Workspace& W, MultiResult* out,
             hipStream_t s) {
  if (M.b != 1 || M.vb) {
    set_error("multi-RHS CG takes scalar matrices (n_dof = 1), got n_dof = %d", M.vb ? -1 : M.b);
    return -2;
  }
  if (k < 1 || k > kMaxRhs) {
    set_error("multi-RHS CG takes 1..%d right-hand sides, got %d", kMaxRhs, k);
    return -2;
  }
  if (P.precond != 0 && P.precond != 1) {
    set_error("multi-RHS CG implements the NONE and DIAG preconditioners, got %d", P.precond);
    return -3;
  }
  if (P.max_iter < 1) {
    set_error("max_iter must be >= 1");
    return -2;
  }
  if (!is_device_pointer(b) || !is_device_pointer(x)) {
    set_error("multi-RHS CG takes device pointers for b and x");
    return -2;
  }
  Context* c = nullptr;
  MH_TRY(context(&c));
  out->iters.assign(k, 0);
  out->resid.assign(k, 0.0);
  out->status.assign(k, 2);
  out->loop_iters = 0;
  const int n = M.n;
  const size_t nk = (size_t)n * k, npk = (size_t)M.np * k;
  if (n == 0) return 0;
  if (sizeof(T) == 4) MH_TRY(matrix_make_fp32(&M, s));
  const T* val = sizeof(T) == 8 ? (const T*)M.val64.p : (const T*)M.val32.p;
  const T* dinv = nullptr;
  if (P.precond == 1) {
    MH_TRY(diag_setup<T>(M, s));
    dinv = sizeof(T) == 8 ? (const T*)M.D64.p : (const T*)M.D32.p;
  }
  const int ks = lane_shift(k);
  const int cfg = spmm_cfg(M, sizeof(T) == 4);
  const int sp_tiles = spmm_grid(M, cfg);
  const int rw_tiles = persistent_grid(((long long)n << ks) / kWave + 1);
  const size_t n_parts = (size_t)std::max(sp_tiles, rw_tiles) * 2 * kWave;
  const size_t off_r = 0, off_p = off_r + align256(sizeof(T) * nk);
  const size_t off_q = off_p + align256(sizeof(T) * npk), off_parts = off_q + align256(sizeof(T) * nk);
  const size_t off_ms = off_parts + align256(sizeof(double) * n_parts);
  const size_t off_rs = off_ms + align256(sizeof(MState));
  const size_t need = off_rs + align256(sizeof(RhsState) * kMaxRhs);
  MH_TRY(W.ensure(need));
  unsigned char* base = W.raw.p;
  T* r = (T*)(base + off_r);
  T* p = (T*)(base + off_p);
  T* q = (T*)(base + off_q);
  double* parts = (double*)(base + off_parts);
  MState* ms = (MState*)(base + off_ms);
  RhsState* rs = (RhsState*)(base + off_rs);

  k_mstate_init<<<1, kWave, 0, s>>>(k, ms, rs);
  if (P.init_x) k_mfill<T><<<grid_for(npk), kBlock, 0, s>>>(npk, x);
  k_mfill<T><<<grid_for(npk), kBlock, 0, s>>>(npk, p);
  // R = B - A X, B2 = (R, R)
  spmm<T, 2>(M, k, ks, cfg, val, x, b, r, parts, nullptr, sp_tiles, s);
  k_mfinal<0><<<1, kFinal, 0, s>>>(k, sp_tiles, parts, ms, rs, 0, 0, 0, P.tol, P.max_iter);
  // Z = M^-1 R, rho = (R, Z)
  k_mrows<T, 3><<<rw_tiles, kBlock, 0, s>>>(n, k, ks, x, r, p, q, dinv, parts, ms, rs);
  k_mfinal<1><<<1, kFinal, 0, s>>>(k, rw_tiles, parts, ms, rs, 0, 0, 1, P.tol, P.max_iter);
  MH_HIP(hipGetLastError());
  int* hflag = c->h_flags;
  const int chunk = std::max(1, P.chunk);
  int it = 1;
  while (it <= P.max_iter) {
    const int end = std::min(it + chunk - 1, P.max_iter);
    for (; it <= end; ++it) {
      k_mrows<T, 0><<<rw_tiles, kBlock, 0, s>>>(n, k, ks, x, r, p, q, dinv, parts, ms, rs);
      spmm<T, 1>(M, k, ks, cfg, val, p, nullptr, q, parts, ms, sp_tiles, s);
      k_mfinal<2><<<1, kFinal, 0, s>>>(k, sp_tiles, parts, ms, rs, it, 0, 0, P.tol, P.max_iter);
      if (it % kIterRR == 0) {
        k_mrows<T, 2><<<rw_tiles, kBlock, 0, s>>>(n, k, ks, x, r, p, q, dinv, parts, ms, rs);
        spmm<T, 2>(M, k, ks, cfg, val, x, b, r, parts, ms, sp_tiles, s);
        k_mrows<T, 3><<<rw_tiles, kBlock, 0, s>>>(n, k, ks, x, r, p, q, dinv, parts, ms, rs);
      } else {
        k_mrows<T, 1><<<rw_tiles, kBlock, 0, s>>>(n, k, ks, x, r, p, q, dinv, parts, ms, rs);
      }
      k_mfinal<1><<<1, kFinal, 0, s>>>(k, rw_tiles, parts, ms, rs, it, 1, 0, P.tol, P.max_iter);
    }
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(hflag, &ms->all_done, sizeof(int), hipMemcpyDeviceToHost, s));
    MH_HIP(hipStreamSynchronize(s));
    if (*hflag) break;
  }
  std::vector<RhsState> hs(k);
  MH_HIP(hipMemcpyAsync(hs.data(), rs, sizeof(RhsState) * k, hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  unsigned long long zero_mask = 0;
  for (int j = 0; j < k; ++j) {
    if (hs[j].zero) zero_mask |= 1ull << j;
    out->iters[j] = hs[j].zero ? 0 : hs[j].iter;
    out->resid[j] = hs[j].resid;
    out->status[j] = hs[j].zero ? 2 : (hs[j].conv ? 0 : 1);
    out->loop_iters = std::max(out->loop_iters, out->iters[j]);
  }
  if (zero_mask) {
    k_zero_cols<T><<<grid_for(nk), kBlock, 0, s>>>(nk, k, zero_mask, x);
    MH_HIP(hipGetLastError());
    MH_HIP(hipStreamSynchronize(s));
  }
  return 0;
}
template int multi_cg<double>(Matrix&, const MultiParams&, int, const double*, double*, Workspace&, MultiResult*,
                              hipStream_t);
template int multi_cg<float>(Matrix&, const MultiParams&, int, const float*, float*, Workspace&, MultiResult*,
                             hipStream_t);

int multi_refine(Matrix& M, const MultiParams& P, double inner_tol, int max_outer, int k, const double* b, double* x,
                 Workspace& Wi, Workspace& Wo, MultiRefineResult* out, hipStream_t s) {
  if (M.b != 1 || M.vb || k < 1 || k > kMaxRhs) {
    set_error("multi-RHS refinement takes scalar matrices and 1..%d right-hand sides", kMaxRhs);
    return -2;
  }
  if (P.precond != 0 && P.precond != 1) {
    set_error("multi-RHS CG implements the NONE and DIAG preconditioners, got %d", P.precond);
    return -3;
  }
  if (!is_device_pointer(b) || !is_device_pointer(x)) {
    set_error("multi-RHS refinement takes device pointers for b and x");
    return -2;
  }
  out->outer = 0;
  out->inner_iters.assign(k, 0);
  out->resid.assign(k, 0.0);
  out->status.assign(k, 2);
  const int n = M.n;
  if (n == 0) return 0;
  const size_t nk = (size_t)n * k, npk = (size_t)M.np * k;
  const int ks = lane_shift(k);
  const int cfg = spmm_cfg(M, 0);
  const int sp_tiles = spmm_grid(M, cfg);
  const size_t off_r32 = align256(sizeof(double) * nk);
  const size_t off_c32 = off_r32 + align256(sizeof(float) * nk);
  const size_t off_parts = off_c32 + align256(sizeof(float) * npk);
  const size_t off_rr = off_parts + align256(sizeof(double) * (size_t)sp_tiles * kWave);
  const size_t need = off_rr + align256(sizeof(double) * kMaxRhs);
  MH_TRY(Wo.ensure(need));
  unsigned char* base = Wo.raw.p;
  double* r = (double*)base;
  float* r32 = (float*)(base + off_r32);
  float* c32 = (float*)(base + off_c32);
  double* parts = (double*)(base + off_parts);
  double* rr_d = (double*)(base + off_rr);
  const unsigned g = grid_for(nk);
  if (P.init_x) k_mfill<double><<<grid_for(npk), kBlock, 0, s>>>(npk, x);
  MultiParams IP = P;
  IP.tol = inner_tol;
  IP.init_x = 1;
  std::vector<double> rr(k), rr0(k);
  unsigned long long zero_mask = 0;
  for (int outer = 0;; ++outer) {
    // FP64 true residual R = B - A X and (R, R) per right-hand side
    spmm<double, 2>(M, k, ks, cfg, M.val64.p, x, b, r, parts, nullptr, sp_tiles, s);
    k_lane_sum<<<1, kFinal, 0, s>>>(k, sp_tiles, parts, rr_d);
    MH_HIP(hipGetLastError());
    MH_HIP(hipMemcpyAsync(rr.data(), rr_d, sizeof(double) * k, hipMemcpyDeviceToHost, s));
    MH_HIP(hipStreamSynchronize(s));
    unsigned long long active = 0;
    for (int j = 0; j < k; ++j) {
      if (outer == 0) {
        rr0[j] = rr[j];
        if (rr[j] == 0.0) zero_mask |= 1ull << j;
      }
      if (zero_mask >> j & 1ull) continue;
      out->resid[j] = std::sqrt(rr[j] / rr0[j]);
      out->status[j] = out->resid[j] < P.tol ? 0 : 1;
      if (out->status[j] == 1) active |= 1ull << j;
    }
    if (outer == 0 && zero_mask) {
      k_zero_cols<double><<<g, kBlock, 0, s>>>(nk, k, zero_mask, x);
      MH_HIP(hipGetLastError());
    }
    if (active == 0 || outer >= max_outer) break;
    // FP32 correction A C = R on the active right-hand sides (others: zero RHS, C = 0)
    k_round_masked<<<g, kBlock, 0, s>>>(nk, k, active, r, r32);
    MH_HIP(hipGetLastError());
    MultiResult ir;
    MH_TRY(multi_cg<float>(M, IP, k, r32, c32, Wi, &ir, s));
    for (int j = 0; j < k; ++j)
      if (active >> j & 1ull) out->inner_iters[j] += ir.iters[j];
    k_add_correction<<<g, kBlock, 0, s>>>(nk, c32, x);
    MH_HIP(hipGetLastError());
    out->outer = outer + 1;
  }
  MH_HIP(hipStreamSynchronize(s));
  return 0;
}

template <typename T>
int multi_matvec(Matrix& M, int k, const T* x, T* y, hipStream_t s) {
  if (M.b != 1 || M.vb || k < 1 || k > kMaxRhs) {
    set_error("multi-RHS matvec takes scalar matrices and 1..%d right-hand sides", kMaxRhs);
    return -2;
  }
  if (!is_device_pointer(x) || !is_device_pointer(y)) {
    set_error("multi-RHS matvec takes device pointers");
    return -2;
  }
  if (M.n == 0) return 0;
  if (sizeof(T) == 4) MH_TRY(matrix_make_fp32(&M, s));
  const T* val = sizeof(T) == 8 ? (const T*)M.val64.p : (const T*)M.val32.p;
  const int cfg = spmm_cfg(M, sizeof(T) == 4);
  spmm<T, 0>(M, k, lane_shift(k), cfg, val, x, nullptr, y, nullptr, nullptr, spmm_grid(M, cfg), s);
  MH_HIP(hipGetLastError());
  return 0;
}
template int multi_matvec<double>(Matrix&, int, const double*, double*, hipStream_t);
template int multi_matvec<float>(Matrix&, int, const float*, float*, hipStream_t);

int multi_num_configs() { return kNumSpmmCfg; }

void multi_config_info(int cfg, int* steps_per_batch, int* wg_per_cu, int* chunk, int* threads, int* lds) {
  *steps_per_batch = kSpmmCfg[cfg].sb;
  *wg_per_cu = kSpmmCfg[cfg].wpc;
  if (chunk) *chunk = kSpmmCfg[cfg].kc;
  if (threads) *threads = kSpmmCfg[cfg].nt;
  if (lds) *lds = kSpmmCfg[cfg].lds;
}

int multi_autotune(Matrix& M, int k, int fp32, int reps, int* best, double* ms, hipStream_t s) {
  if (M.b != 1 || M.vb || k < 1 || k > kMaxRhs) {
    set_error("multi-RHS autotune takes scalar matrices and 1..%d right-hand sides", kMaxRhs);
    return -2;
  }
  if (M.n == 0) {
    *best = 0;
    return 0;
  }
  if (reps < 1) reps = 1;
  if (fp32) MH_TRY(matrix_make_fp32(&M, s));
  const int ks = lane_shift(k);
  const size_t npk = (size_t)M.np * k, es = fp32 ? sizeof(float) : sizeof(double);
  DevArray<unsigned char> xb, yb;
  DevArray<double> parts;
  MH_TRY(xb.alloc(npk * es));
  MH_TRY(yb.alloc(npk * es));
  long long gmax = 0;
  for (int c = 0; c < kNumSpmmCfg; ++c) gmax = std::max<long long>(gmax, spmm_grid(M, c));
  MH_TRY(parts.alloc((size_t)gmax * kWave));
  if (fp32) k_mfill<float><<<grid_for(npk), kBlock, 0, s>>>(npk, (float*)xb.p);
  else k_mfill<double><<<grid_for(npk), kBlock, 0, s>>>(npk, (double*)xb.p);
  MH_HIP(hipGetLastError());
  hipEvent_t e0, e1;
  MH_HIP(hipEventCreate(&e0));
  MH_HIP(hipEventCreate(&e1));
  double best_ms = 1e300;
  int best_c = 0, rc = 0;
  for (int c = 0; c < kNumSpmmCfg && !rc; ++c) {
    const int grid = spmm_grid(M, c);
    auto one = [&]() {
      if (fp32)
        spmm<float, 1>(M, k, ks, c, M.val32.p, (const float*)xb.p, nullptr, (float*)yb.p, parts.p, nullptr, grid, s);
      else
        spmm<double, 1>(M, k, ks, c, M.val64.p, (const double*)xb.p, nullptr, (double*)yb.p, parts.p, nullptr, grid,
                        s);
    };
    one();
    float t = 0.f;
    if (hipEventRecord(e0, s) != hipSuccess) rc = -1;
    for (int r = 0; r < reps && !rc; ++r) one();
    if (!rc && (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
                hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess)) {
      set_error("multi-RHS autotune: launch or timing failed");
      rc = -1;
    }
    const double avg = (double)t / reps;
    if (ms) ms[c] = avg;
    if (!rc && avg < best_ms) {
      best_ms = avg;
      best_c = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (fp32) M.mrhs_cfg32 = best_c;
  else M.mrhs_cfg64 = best_c;
  if (best) *best = best_c;
  return 0;
}

}  // namespace mh
