// Magnitude-sorted inner product, the breakdown fallback of monolis's BiCGSTAB
// (monolis_global_sorted_inner_product_main_R, src/linalg/inner_product.f90:262-331):
// products x_i*y_i are ordered by |x_i*y_i| ascending and summed left to right.
// On the GPU: a product kernel, a radix sort of (|p| bit pattern, p) pairs
// (non-negative doubles order like their unsigned bit patterns) and a single
// lane that adds the sorted values in order.  Only runs when a dot is exactly 0.
#include <hipcub/hipcub.hpp>

#include "mh_internal.hpp"

namespace mh {

template <typename T>
__global__ void k_abs_prod(long long n, const T* __restrict__ x, const T* __restrict__ y,
                           unsigned long long* __restrict__ key, double* __restrict__ val) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    const double p = (double)x[i] * (double)y[i];
    val[i] = p;
    key[i] = (unsigned long long)__double_as_longlong(fabs(p));
  }
}

__global__ void k_seq_sum(long long n, const double* __restrict__ v, double* __restrict__ out) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    double s = 0.0;
    for (long long i = 0; i < n; ++i) s += v[i];
    *out = s;
  }
}

template <typename T>
int sorted_dot(long long n, const T* x, const T* y, double* out, hipStream_t s) {
  if (n <= 0) {
    MH_HIP(hipMemsetAsync(out, 0, sizeof(double), s));
    return 0;
  }
  if (n > 0x7fffffffLL) {
    set_error("sorted dot: vector too long");
    return -2;
  }
  DevArray<unsigned long long> k0, k1;
  DevArray<double> v0, v1;
  MH_TRY(k0.alloc((size_t)n));
  MH_TRY(k1.alloc((size_t)n));
  MH_TRY(v0.alloc((size_t)n));
  MH_TRY(v1.alloc((size_t)n));
  long long g = (n + 255) / 256;
  if (g > 65536) g = 65536;
  k_abs_prod<T><<<(unsigned)g, 256, 0, s>>>(n, x, y, k0.p, v0.p);
  MH_HIP(hipGetLastError());
  size_t tmp_bytes = 0;
  MH_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, k0.p, k1.p, v0.p, v1.p, (int)n, 0, 64, s));
  DevArray<unsigned char> tmp;
  MH_TRY(tmp.alloc(tmp_bytes));
  MH_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, tmp_bytes, k0.p, k1.p, v0.p, v1.p, (int)n, 0, 64, s));
  k_seq_sum<<<1, 64, 0, s>>>(n, v1.p, out);
  MH_HIP(hipGetLastError());
  MH_HIP(hipStreamSynchronize(s));  // temporaries are released on return
  return 0;
}
template int sorted_dot<double>(long long, const double*, const double*, double*, hipStream_t);
template int sorted_dot<float>(long long, const float*, const float*, double*, hipStream_t);

}  // namespace mh
