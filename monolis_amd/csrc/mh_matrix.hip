// Matrix ingress (monolis BCSR -> SELL-64), SpMV, BLAS-1 and block-Jacobi.
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <thread>
#include <vector>

#include "mh_internal.hpp"
#include "mh_launch.hpp"

namespace mh {

// ---------------------------------------------------------------------------
// Error state and context
// ---------------------------------------------------------------------------
static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}
const char* last_error() { return g_err; }
void clear_error() { g_err[0] = 0; }

static Context g_ctx;
static std::mutex g_ctx_mu;
static int g_cus = 256;

int cu_count() { return g_cus; }

int context(Context** out) {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  if (!g_ctx.ok) {
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev < 1) {
      set_error("no HIP device available (%s)", e == hipSuccess ? "count 0" : hipGetErrorString(e));
      return -1;
    }
    MH_HIP(hipGetDevice(&g_ctx.device));
    MH_HIP(hipDeviceGetAttribute(&g_cus, hipDeviceAttributeMultiprocessorCount, g_ctx.device));
    MH_HIP(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
    MH_HIP(hipHostMalloc((void**)&g_ctx.h_flags, sizeof(int) * kFlagSlots, hipHostMallocDefault));
    MH_HIP(hipHostMalloc((void**)&g_ctx.h_poll, sizeof(PollView) * kFlagSlots, hipHostMallocDefault));
    for (int i = 0; i < kFlagSlots; ++i) MH_HIP(hipEventCreateWithFlags(&g_ctx.ev[i], hipEventDisableTiming));
    for (int i = 0; i < kUploadThreads; ++i) MH_HIP(hipStreamCreateWithFlags(&g_ctx.up[i], hipStreamNonBlocking));
    g_ctx.ok = true;
  }
  *out = &g_ctx;
  return 0;
}

Context* context_if_ready() {
  std::lock_guard<std::mutex> lk(g_ctx_mu);
  return g_ctx.ok ? &g_ctx : nullptr;
}

template <typename T>
int DevArray<T>::alloc(size_t count) {
  release();
  if (count == 0) count = 1;
  MH_HIP(hipMalloc((void**)&p, count * sizeof(T)));
  n = count;
  return 0;
}
template <typename T>
void DevArray<T>::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  n = 0;
}
template struct DevArray<int>;
template struct DevArray<unsigned>;
template struct DevArray<unsigned short>;
template struct DevArray<double>;
template struct DevArray<float>;
template struct DevArray<unsigned char>;
template struct DevArray<unsigned long long>;
template struct DevArray<long long>;
template struct DevArray<zd>;

int is_device_pointer(const void* p) {
  if (!p) return 0;
  hipPointerAttribute_t a;
  hipError_t e = hipPointerGetAttributes(&a, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // plain host memory reports an error: clear it
    return 0;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Copies from pageable host memory pass through the runtime's staging
// buffers, which one host thread fills at about 30 GB/s.  Large ones are split
// over kUploadThreads threads, each copying its share on its own stream
// (2.1 GB: 71 ms -> 43-51 ms); the call then returns with the data in place.
// Device and pinned sources, and small copies, take one asynchronous copy on s.
constexpr size_t kUploadSplitBytes = (size_t)64 << 20;

static bool is_pageable(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // older runtimes report unregistered host memory as an error
    return true;
  }
  return a.type == hipMemoryTypeUnregistered;
}

int upload(void* dst, const void* src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return 0;
  if (bytes >= kUploadSplitBytes && is_pageable(src)) {
    Context* c = nullptr;
    MH_TRY(context(&c));
    MH_HIP(hipStreamSynchronize(s));  // earlier work on s may still use dst
    hipError_t err[kUploadThreads];
    constexpr size_t kAlign = (size_t)1 << 16;
    // share k: bytes [b0, b1), 64 KB-aligned starts, the last share takes the tail
    auto job = [=, &err](int k) {
      const size_t b0 = std::min(bytes, (bytes * k / kUploadThreads) / kAlign * kAlign);
      const size_t b1 =
          k + 1 == kUploadThreads ? bytes : std::min(bytes, (bytes * (k + 1) / kUploadThreads) / kAlign * kAlign);
      hipError_t e = hipSetDevice(c->device);
      if (e == hipSuccess && b1 > b0)
        e = hipMemcpyAsync(static_cast<char*>(dst) + b0, static_cast<const char*>(src) + b0, b1 - b0,
                           hipMemcpyHostToDevice, c->up[k]);
      if (e == hipSuccess) e = hipStreamSynchronize(c->up[k]);
      err[k] = e;
    };
    std::vector<std::thread> th;
    th.reserve(kUploadThreads);
    for (int k = 0; k < kUploadThreads; ++k) {
      try {
        th.emplace_back(job, k);
      } catch (...) {
        job(k);  // no thread available: this share is copied by the caller
      }
    }
    for (auto& t : th) t.join();
    for (int k = 0; k < kUploadThreads; ++k)
      if (err[k] != hipSuccess) {
        set_error("host-to-device copy failed: %s", hipGetErrorString(err[k]));
        return -1;
      }
    return 0;
  }
  MH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, s));
  return 0;
}

int Workspace::ensure(size_t need) {
  if (need <= bytes) return 0;
  MH_TRY(raw.alloc(need));
  bytes = need;
  return 0;
}

size_t partials_needed(long long n) {
  // one partial slot-group per 256-row workgroup (or per 4 slices) + margin
  return (size_t)((n + kBlock - 1) / kBlock + 8);
}

// ---------------------------------------------------------------------------
// SELL-64 construction
// ---------------------------------------------------------------------------
__global__ void k_slice_width(int n, int nslices, const int* __restrict__ rowptr, int* __restrict__ width) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int s = t / kWave;
  int len = 0;
  if (t < n) len = rowptr[t + 1] - rowptr[t];
  len = wave_max(len);
  if ((threadIdx.x & (kWave - 1)) == 0 && s < nslices) width[s] = len;
}

// Structural check of caller arrays before any kernel indexes with them:
// index[0] == 0, index non-decreasing within [0, nz], 1 <= item <= np.
__global__ void k_validate(int np, int nz, const int* __restrict__ rowptr, const int* __restrict__ item,
                           int* __restrict__ bad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const int a = rowptr[i], b = rowptr[i + 1];
  if (i == 0 && a != 0) atomicOr(bad, 1);
  if (i == np - 1 && b != nz) atomicOr(bad, 1);
  if (a < 0 || b < a || b > nz) {
    atomicOr(bad, 2);
    return;
  }
  for (int j = a; j < b; ++j) {
    const int c = item[j];
    if (c < 1 || c > np) {
      atomicOr(bad, 4);
      return;
    }
  }
}

// One thread per (slice, lane); writes every slot of that lane, padding included.
template <typename TS, typename TV>
__global__ void k_sell_fill(int n, int nslices, int b, const int* __restrict__ rowptr, const int* __restrict__ item,
                            const TS* __restrict__ A, const int* __restrict__ slice_ptr,
                            int* __restrict__ col, TV* __restrict__ val, int fill_col) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nslices * kWave) return;
  const int s = (int)(t / kWave);
  const int lane = (int)(t % kWave);
  const int row = (int)t;
  const int k0 = slice_ptr[s];
  const int w = slice_ptr[s + 1] - k0;
  const int bb = b * b;
  int j0 = 0, len = 0;
  if (row < n) {
    j0 = rowptr[row];
    len = rowptr[row + 1] - j0;
  }
  for (int k = 0; k < w; ++k) {
    const size_t slot = (size_t)(k0 + k);
    if (k < len) {
      if (fill_col) col[slot * kWave + lane] = item[j0 + k] - 1;
      for (int e = 0; e < bb; ++e) val[sell_vpos<TV>(slot, e, bb, lane)] = (TV)A[(size_t)(j0 + k) * bb + e];
    } else {
      if (fill_col) col[slot * kWave + lane] = row < n ? row : 0;
      for (int e = 0; e < bb; ++e) val[sell_vpos<TV>(slot, e, bb, lane)] = (TV)0;
    }
  }
}

// Largest |col - row| over the stored slots of the valid rows (atomicMax).
__global__ void k_col_span(int n, int nslices, const int* __restrict__ rowptr, const int* __restrict__ slice_ptr,
                           const int* __restrict__ col, int* __restrict__ span) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nslices * kWave) return;
  const int s = (int)(t / kWave), lane = (int)(t % kWave), row = (int)t;
  if (row >= n) return;
  const int k0 = slice_ptr[s];
  const int len = rowptr[row + 1] - rowptr[row];
  int m = 0;
  for (int k = 0; k < len; ++k) {
    const int d = col[((size_t)k0 + k) * kWave + lane] - row;
    m = max(m, d < 0 ? -d : d);
  }
  if (m > 0) atomicMax(span, m);
}

// 16-bit delta copy of the column ids: slot t of lane l goes to half (t & 1)
// of word (t / 2) * 64 + l (every slot, padding included; rows past n get 0).
__global__ void k_col16_fill(int n, int nslices, const int* __restrict__ slice_ptr, const int* __restrict__ col,
                             unsigned short* __restrict__ col16) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nslices * kWave) return;
  const int s = (int)(t / kWave), lane = (int)(t % kWave), row = (int)t;
  const int k0 = slice_ptr[s], w = slice_ptr[s + 1] - k0;
  for (int k = 0; k < w; ++k) {
    const size_t slot = (size_t)k0 + k;
    const int d = row < n ? col[slot * kWave + lane] - row : 0;
    col16[((slot >> 1) * kWave + lane) * 2 + (slot & 1)] = (unsigned short)(short)d;
  }
}

__global__ void k_to_fp32(size_t m, const double* __restrict__ a, float* __restrict__ b) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// FP32 copy of the SELL values; position i of the FP32 array holds the entry
// (slot, e, lane) that its own layout puts there (the FP64 and FP32 value
// layouts differ when only 8-byte values are stored in pairs).
__global__ void k_vals_to_fp32(size_t m, int bb, const double* __restrict__ a, float* __restrict__ b) {
  const size_t per = (size_t)bb * kWave;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x) {
    if constexpr (sell_paired<float>() == sell_paired<double>()) {
      b[i] = (float)a[i];
    } else {  // FP32 planes one per entry: i = slot*per + e*64 + lane
      const size_t slot = i / per;
      const int r = (int)(i % per);
      b[i] = (float)a[sell_vpos<double>(slot, r / kWave, bb, r % kWave)];
    }
  }
}

static int grid_1d(size_t m) {
  size_t g = (m + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  return (int)(g > 0 ? g : 1);
}

Matrix::~Matrix() {
  delete transpose;
  delete vb;
}

template <typename TV>
SellView<TV> view(const Matrix& M) {
  SellView<TV> v;
  v.nrows = M.n;
  v.nslices = M.nslices;
  v.slice_ptr = M.slice_ptr.p;
  v.rowptr = M.rowptr.p;
  v.col = M.col.p;
  v.col16 = M.col16.n ? (const unsigned*)M.col16.p : nullptr;
  if constexpr (sizeof(TV) == 16)
    v.val = (const TV*)M.valz.p;
  else if constexpr (sizeof(TV) == 8)
    v.val = (const TV*)M.val64.p;
  else
    v.val = (const TV*)M.val32.p;
  return v;
}
template SellView<double> view<double>(const Matrix&);
template SellView<float> view<float>(const Matrix&);
template SellView<zd> view<zd>(const Matrix&);

static int fill_values(Matrix* M, const double* A, hipStream_t s, int fill_col) {
  // complex values are two doubles each
  const size_t ndbl = (size_t)M->nz * M->b * M->b * (M->is_complex ? 2 : 1);
  const double* dA = A;
  DevArray<double> tmp;
  if (!is_device_pointer(A)) {
    MH_TRY(tmp.alloc(ndbl));
    MH_TRY(upload(tmp.p, A, ndbl * sizeof(double), s));
    dA = tmp.p;
  }
  const long long threads = (long long)M->nslices * kWave;
  const unsigned grid = (unsigned)((threads + kBlock - 1) / kBlock);
  if (M->is_complex)
    k_sell_fill<zd, zd><<<grid, kBlock, 0, s>>>(M->n, M->nslices, M->b, M->rowptr.p, M->csr_item.p,
                                               (const zd*)dA, M->slice_ptr.p, M->col.p, M->valz.p, fill_col);
  else
    k_sell_fill<double, double><<<grid, kBlock, 0, s>>>(M->n, M->nslices, M->b, M->rowptr.p, M->csr_item.p, dA,
                                                       M->slice_ptr.p, M->col.p, M->val64.p, fill_col);
  MH_HIP(hipGetLastError());
  if (tmp.p) MH_HIP(hipStreamSynchronize(s));  // tmp is released on return
  M->diag_ready = M->diag32_ready = M->diagz_ready = false;
  M->transpose_stale = true;
  if (M->val32_ready) {
    M->val32_ready = false;
    MH_TRY(matrix_make_fp32(M, s));
  }
  return 0;
}

// Builds the 16-bit column stream when every |col - row| <= 32767.
static int build_col16(Matrix* M, hipStream_t s) {
  M->col16.release();
  if (M->nslices == 0 || M->nslots == 0) return 0;
  DevArray<int> span;
  MH_TRY(span.alloc(1));
  MH_HIP(hipMemsetAsync(span.p, 0, sizeof(int), s));
  const long long threads = (long long)M->nslices * kWave;
  const unsigned grid = (unsigned)((threads + kBlock - 1) / kBlock);
  k_col_span<<<grid, kBlock, 0, s>>>(M->n, M->nslices, M->rowptr.p, M->slice_ptr.p, M->col.p, span.p);
  MH_HIP(hipGetLastError());
  int h = 0;
  MH_HIP(hipMemcpyAsync(&h, span.p, sizeof(int), hipMemcpyDeviceToHost, s));
  MH_HIP(hipStreamSynchronize(s));
  if (h > 32767) return 0;
  MH_TRY(M->col16.alloc(((M->nslots + 1) / 2) * kWave * 2));
  k_col16_fill<<<grid, kBlock, 0, s>>>(M->n, M->nslices, M->slice_ptr.p, M->col.p, M->col16.p);
  MH_HIP(hipGetLastError());
  return 0;
}

static int matrix_build(Matrix** out, int n, int np, int nz, int b, const int* index, const int* item,
                        const double* A, bool cplx, hipStream_t s) {
  *out = nullptr;
  if (n < 0 || np < n || nz < 0 || b < 1) {
    set_error("invalid matrix sizes n=%d np=%d nz=%d ndof=%d", n, np, nz, b);
    return -2;
  }
  if (b > kMaxB) {
    set_error("block size %d is larger than the supported maximum %d", b, kMaxB);
    return -2;
  }
  Matrix* M = new Matrix();
  M->n = n;
  M->np = np;
  M->nz = nz;
  M->b = b;
  M->is_complex = cplx;
  M->nslices = (n + kWave - 1) / kWave;
  // Untuned default (the flat entry points never tune): one wavefront per
  // slice, 3 slots per step.  Against the persistent 8-waves-per-CU grid it
  // takes 30.8 / 21.3 us instead of 41.7 / 33.3 (FP64 / FP32, 7-point 128^3
  // convection-diffusion) and 328 / 161 us instead of 351 / 171 (27-point
  // BCSR3 on 100^3 nodes, where it is the fastest FP32 variant).
  M->variant64 = M->variant32 = 4;
  int rc = 0;
  do {
    if ((rc = M->rowptr.alloc((size_t)np + 1))) break;
    if ((rc = upload(M->rowptr.p, index, sizeof(int) * ((size_t)np + 1), s))) break;
    if ((rc = M->csr_item.alloc((size_t)nz))) break;
    if ((rc = upload(M->csr_item.p, item, sizeof(int) * (size_t)nz, s))) break;
    // slice widths -> slice_ptr (exclusive scan on the host: nslices ints)
    DevArray<int> width;
    if ((rc = width.alloc((size_t)M->nslices + 1))) break;
    if (np > 0) {
      int bad = 0;
      if (hipMemsetAsync(width.p + M->nslices, 0, sizeof(int), s) != hipSuccess) {
        set_error("memset failed");
        rc = -1;
        break;
      }
      k_validate<<<(np + kBlock - 1) / kBlock, kBlock, 0, s>>>(np, nz, M->rowptr.p, M->csr_item.p,
                                                               width.p + M->nslices);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(&bad, width.p + M->nslices, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        set_error("CSR validation kernel failed");
        rc = -1;
        break;
      }
      if (bad) {
        set_error("invalid BCSR arrays (%s%s%s)", (bad & 1) ? "index[0] != 0 or index[NP] != nz; " : "",
                  (bad & 2) ? "index not non-decreasing within [0, nz]; " : "",
                  (bad & 4) ? "item outside 1..NP" : "");
        rc = -2;
        break;
      }
    }
    if (M->nslices > 0) {
      const long long threads = (long long)M->nslices * kWave;
      k_slice_width<<<(unsigned)((threads + kBlock - 1) / kBlock), kBlock, 0, s>>>(n, M->nslices, M->rowptr.p,
                                                                                 width.p);
      if (hipGetLastError() != hipSuccess) {
        set_error("slice width kernel launch failed");
        rc = -1;
        break;
      }
    }
    std::vector<int> hw((size_t)M->nslices + 1, 0), hp((size_t)M->nslices + 1, 0);
    if (M->nslices > 0) {
      if (hipMemcpyAsync(hw.data(), width.p, sizeof(int) * M->nslices, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess) {
        set_error("slice width copy failed");
        rc = -1;
        break;
      }
    }
    size_t acc = 0;
    for (int i = 0; i < M->nslices; ++i) {
      hp[i] = (int)acc;
      acc += (size_t)hw[i];
    }
    hp[M->nslices] = (int)acc;
    if (acc * kWave * (size_t)b * b >= ((size_t)1 << 40)) {
      set_error("matrix too large");
      rc = -2;
      break;
    }
    M->nslots = acc;
    if ((rc = M->slice_ptr.alloc((size_t)M->nslices + 1))) break;
    if ((rc = upload(M->slice_ptr.p, hp.data(), sizeof(int) * ((size_t)M->nslices + 1), s))) break;
    if ((rc = M->col.alloc(acc * kWave))) break;
    if (cplx) {
      if ((rc = M->valz.alloc(acc * kWave * (size_t)b * b))) break;
    } else {
      if ((rc = M->val64.alloc(acc * kWave * (size_t)b * b))) break;
    }
    if ((rc = fill_values(M, A, s, 1))) break;
    if ((rc = build_col16(M, s))) break;
    if (hipStreamSynchronize(s) != hipSuccess) {
      set_error("matrix build failed");
      rc = -1;
      break;
    }
  } while (0);
  if (rc) {
    delete M;
    return rc;
  }
  *out = M;
  return 0;
}

int matrix_create(Matrix** out, int n, int np, int nz, int b, const int* index, const int* item,
                  const double* A, hipStream_t s) {
  return matrix_build(out, n, np, nz, b, index, item, A, false, s);
}

int matrix_create_complex(Matrix** out, int n, int np, int nz, int b, const int* index, const int* item,
                          const double* A, hipStream_t s) {
  return matrix_build(out, n, np, nz, b, index, item, A, true, s);
}

int matrix_set_values(Matrix* M, const double* A, hipStream_t s) { return fill_values(M, A, s, 0); }

int matrix_make_fp32(Matrix* M, hipStream_t s) {
  if (M->is_complex) {
    set_error("complex matrices have no FP32 copy");
    return -2;
  }
  if (M->val32_ready) return 0;
  if (M->val32.n < M->val64.n) MH_TRY(M->val32.alloc(M->val64.n));
  k_vals_to_fp32<<<grid_1d(M->val64.n), kBlock, 0, s>>>(M->val64.n, M->b * M->b, M->val64.p, M->val32.p);
  MH_HIP(hipGetLastError());
  M->val32_ready = true;
  return 0;
}

// ---------------------------------------------------------------------------
// SpMV ops
// ---------------------------------------------------------------------------
template <typename T, int B>
struct OpSpmv {
  static constexpr int NV = 0;
  const T* __restrict__ x;
  T* __restrict__ y;
  __device__ bool skip() const { return false; }
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

template <typename T, int B>
struct OpResid {
  static constexpr int NV = 1;
  const T* __restrict__ x;
  const T* __restrict__ b;
  T* __restrict__ r;
  double* out;
  __device__ bool skip() const { return false; }
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
      v[k] = v[k] - acc[k];  // monolis_residual_main_R: R = 1*B + (-1)*(A X)   (matvec.f90:901)
      if (valid) part[0] += (double)v[k] * (double)v[k];
    }
    io.store(r, v);
  }
  __device__ void finalize(const double (&tot)[1]) const {
    if (out) *out = tot[0];
  }
};

template <typename T>
int spmv_variant(const Matrix& M, const T* x, T* y, int variant, hipStream_t s) {
  if (sizeof(T) == 4 && !M.val32_ready) {
    set_error("fp32 values not built");
    return -2;
  }
  if (variant < 0 || variant >= kNumVariants) {
    set_error("SpMV variant %d out of range", variant);
    return -2;
  }
  MH_DISPATCH_B(M.b, {
    OpSpmv<T, B> op{x, y};
    MH_TRY((launch_sell<T, B>(M, op, nullptr, variant, s)));
  });
  return 0;
}
template int spmv_variant<double>(const Matrix&, const double*, double*, int, hipStream_t);
template int spmv_variant<float>(const Matrix&, const float*, float*, int, hipStream_t);

template <typename T>
int spmv(const Matrix& M, const T* x, T* y, hipStream_t s) {
  return spmv_variant<T>(M, x, y, sizeof(T) == 8 ? M.variant64 : M.variant32, s);
}
template int spmv<double>(const Matrix&, const double*, double*, hipStream_t);
template int spmv<float>(const Matrix&, const float*, float*, hipStream_t);

template <typename T>
int residual(const Matrix& M, const T* x, const T* b, T* r, double* partials, double* rr_out, hipStream_t s) {
  MH_DISPATCH_B(M.b, {
    OpResid<T, B> op{x, b, r, rr_out};
    MH_TRY((launch_sell<T, B>(M, op, partials, sizeof(T) == 8 ? M.variant64 : M.variant32, s)));
  });
  return 0;
}
template int residual<double>(const Matrix&, const double*, const double*, double*, double*, double*,
                              hipStream_t);
template int residual<float>(const Matrix&, const float*, const float*, float*, double*, double*, hipStream_t);

// ---------------------------------------------------------------------------
// BLAS-1: deterministic dot, AXPBY
// ---------------------------------------------------------------------------
template <typename T>
struct OpDot {
  using VT = T;
  static constexpr int BS = 1;
  static constexpr int NV = 1;
  const T* __restrict__ x;
  const T* __restrict__ y;
  double* out;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ void row(int i, bool valid, double (&part)[1], const WaveRows<T, 1>&) const {
    if (valid) part[0] += (double)x[i] * (double)y[i];
  }
  __device__ void finalize(const double (&tot)[1]) const { *out = tot[0]; }
};

template <typename T>
int dot(long long m, const T* x, const T* y, double* out, double* partials, hipStream_t s) {
  if (m > 0x7fffffffLL) {
    set_error("vector too long");
    return -2;
  }
  OpDot<T> op{x, y, out};
  return launch_rows((int)m, op, partials, s);
}
template int dot<double>(long long, const double*, const double*, double*, double*, hipStream_t);
template int dot<float>(long long, const float*, const float*, double*, double*, hipStream_t);

template <typename T>
__global__ void k_axpby(long long m, T alpha, const T* __restrict__ x, T beta, const T* __restrict__ y,
                        T* __restrict__ z) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (long long)gridDim.x * blockDim.x)
    z[i] = alpha * x[i] + beta * y[i];
}

template <typename T>
int axpby(long long m, double alpha, const T* x, double beta, const T* y, T* z, hipStream_t s) {
  if (m <= 0) return 0;
  k_axpby<T><<<grid_1d((size_t)m), kBlock, 0, s>>>(m, (T)alpha, x, (T)beta, y, z);
  MH_HIP(hipGetLastError());
  return 0;
}
template int axpby<double>(long long, double, const double*, double, const double*, double*, hipStream_t);
template int axpby<float>(long long, double, const float*, double, const float*, float*, hipStream_t);

// ---------------------------------------------------------------------------
// Block-Jacobi setup: diagonal block extraction + no-pivot LU with reciprocal
// pivots, in place in structure-of-arrays D (entry e of row i at D[e*n + i]).
// b == 3 follows diag_33.f90:40-96 (every row is factored, a missing diagonal
// block becomes the identity); other b follow diag_nn.f90:45-80 (only rows
// that hold a diagonal block are factored, the others stay zero).
// ---------------------------------------------------------------------------
__global__ void k_diag_setup(int n, int b, const int* __restrict__ slice_ptr, const int* __restrict__ rowptr,
                             const int* __restrict__ col, const double* __restrict__ val, double* __restrict__ D) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int s = i / kWave, lane = i % kWave;
  const int bb = b * b;
  const size_t N = (size_t)n;
  const int k0 = slice_ptr[s];
  const int len = rowptr[i + 1] - rowptr[i];
  int found = 0;
  for (int e = 0; e < bb; ++e) D[(size_t)e * N + i] = 0.0;
  for (int k = 0; k < len; ++k) {
    const size_t slot = (size_t)(k0 + k);
    if (col[slot * kWave + lane] == i) {
      found = 1;
      for (int e = 0; e < bb; ++e) D[(size_t)e * N + i] = val[sell_vpos<double>(slot, e, bb, lane)];
    }
  }
  if (!found && b != 3) return;
#define LU(r, c) D[(size_t)((r) * b + (c)) * N + i]
  for (int k = 0; k < b; ++k) {
    if (LU(k, k) == 0.0) LU(k, k) = 1.0;
    LU(k, k) = 1.0 / LU(k, k);
    for (int l = k + 1; l < b; ++l) {
      LU(l, k) = LU(l, k) * LU(k, k);
      for (int j = k + 1; j < b; ++j) LU(l, j) = LU(l, j) - LU(l, k) * LU(k, j);
    }
  }
#undef LU
}

template <typename T>
int diag_setup(Matrix& M, hipStream_t s) {
  if (M.vb) {  // diag_V factors are built with the values
    if (sizeof(T) == 4) MH_TRY(vdiag_make_fp32(M, s));
    return 0;
  }
  if (!M.diag_ready) {
    const size_t need = (size_t)M.n * M.b * M.b;
    if (M.D64.n < need) MH_TRY(M.D64.alloc(need));
    if (M.n > 0) {
      k_diag_setup<<<(M.n + kBlock - 1) / kBlock, kBlock, 0, s>>>(M.n, M.b, M.slice_ptr.p, M.rowptr.p, M.col.p,
                                                                   M.val64.p, M.D64.p);
      MH_HIP(hipGetLastError());
    }
    M.diag_ready = true;
    M.diag32_ready = false;
  }
  if (sizeof(T) == 4 && !M.diag32_ready) {
    const size_t need = (size_t)M.n * M.b * M.b;
    if (M.D32.n < need) MH_TRY(M.D32.alloc(need));
    if (need > 0) {
      k_to_fp32<<<grid_1d(need), kBlock, 0, s>>>(need, M.D64.p, M.D32.p);
      MH_HIP(hipGetLastError());
    }
    M.diag32_ready = true;
  }
  return 0;
}
template int diag_setup<double>(Matrix&, hipStream_t);
template int diag_setup<float>(Matrix&, hipStream_t);

template <typename T, int B>
struct OpDiagApply {
  using VT = T;
  static constexpr int BS = B;
  static constexpr int NV = 0;
  const T* __restrict__ D;
  const T* __restrict__ x;
  T* __restrict__ y;
  int n;
  __device__ bool skip() const { return false; }
  __device__ void begin() {}
  __device__ void row(int i, bool valid, double (&)[1], const WaveRows<T, B>& io) const {
    T v[B];
    io.load(x, v);
    if (valid) diag_apply<T, T, B>(D, n, i, v);
    io.store(y, v);
  }
  __device__ void finalize(const double (&)[1]) const {}
};

template <typename T>
int diag_apply(const Matrix& M, const T* x, T* y, hipStream_t s) {
  if (M.vb) return vdiag_apply<T>(M, x, y, s);
  if (M.n == 0) return 0;
  const T* D = nullptr;
  if constexpr (sizeof(T) == 8) D = (const T*)M.D64.p;
  else D = (const T*)M.D32.p;
  MH_DISPATCH_B(M.b, {
    OpDiagApply<T, B> op{D, x, y, M.n};
    MH_TRY(launch_rows(M.n, op, nullptr, s));
  });
  return 0;
}
template int diag_apply<double>(const Matrix&, const double*, double*, hipStream_t);
template int diag_apply<float>(const Matrix&, const float*, float*, hipStream_t);

}  // namespace mh
