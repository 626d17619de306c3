// Precision conversion of column-major blocks (src/convert.cu:324-360) and
// strided 2-D copies. HBM-bound streaming kernels: each thread converts 4
// consecutive rows of one column, so a wavefront reads 256 consecutive
// elements of a column; the y grid dimension walks the columns.
#include "common.hpp"

namespace remifa {
namespace {

struct bf16s {
  uint16_t v;
};
struct e4m3s {
  uint8_t v;
};

__device__ __forceinline__ double ld_wide(const float* p) { return (double)*p; }
__device__ __forceinline__ double ld_wide(const double* p) { return *p; }
__device__ __forceinline__ double ld_wide(const half_t* p) { return (double)(float)*p; }
__device__ __forceinline__ double ld_wide(const bf16s* p) { return (double)bf16_bits_to_f32(p->v); }
__device__ __forceinline__ double ld_wide(const e4m3s* p) { return (double)e4m3_bits_to_f32(p->v); }

// Narrowing from float sources goes through float (exact), from double
// sources straight from the double value: one rounding either way.
template <typename S>
struct is_double_src {
  static constexpr bool value = false;
};
template <>
struct is_double_src<double> {
  static constexpr bool value = true;
};

template <typename S>
__device__ __forceinline__ void st_narrow(float* p, const S* s) {
  if constexpr (is_double_src<S>::value)
    *p = (float)*s;
  else
    *p = (float)ld_wide(s);
}
template <typename S>
__device__ __forceinline__ void st_narrow(double* p, const S* s) {
  *p = ld_wide(s);
}
template <typename S>
__device__ __forceinline__ void st_narrow(half_t* p, const S* s) {
  if constexpr (is_double_src<S>::value)
    *p = (half_t)*s;  // fptrunc f64 -> f16: single rounding
  else
    *p = (half_t)(float)ld_wide(s);
}
template <typename S>
__device__ __forceinline__ void st_narrow(bf16s* p, const S* s) {
  p->v = f32_to_bf16_bits((float)ld_wide(s));
}
template <typename S>
__device__ __forceinline__ void st_narrow(e4m3s* p, const S* s) {
  p->v = f32_to_e4m3_bits((float)ld_wide(s));
}

template <typename S, typename D>
__global__ void __launch_bounds__(256) convert_kernel(int64_t m, int64_t n, const S* __restrict__ a,
                                                      int64_t lda, D* __restrict__ b, int64_t ldb) {
  const int64_t i0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
  if (i0 >= m) return;
  for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
    const S* ac = a + j * lda;
    D* bc = b + j * ldb;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int64_t i = i0 + r;
      if (i < m) st_narrow(bc + i, ac + i);
    }
  }
}

template <typename S, typename D>
int launch_typed(hipStream_t s, int64_t m, int64_t n, const void* a, int64_t lda, void* b,
                 int64_t ldb) {
  dim3 grid((unsigned)ceil_div(m, 1024), (unsigned)(n < 65535 ? n : 65535));
  convert_kernel<S, D>
      <<<grid, 256, 0, s>>>(m, n, (const S*)a, lda, (D*)b, ldb);
  RF_HIP(hipGetLastError());
  return 0;
}

template <typename S>
int launch_src(hipStream_t s, int dst_t, int64_t m, int64_t n, const void* a, int64_t lda,
               void* b, int64_t ldb) {
  switch (dst_t) {
    case REMIFA_R_32F: return launch_typed<S, float>(s, m, n, a, lda, b, ldb);
    case REMIFA_R_16F: return launch_typed<S, half_t>(s, m, n, a, lda, b, ldb);
    case REMIFA_R_64F: return launch_typed<S, double>(s, m, n, a, lda, b, ldb);
    case REMIFA_R_16BF: return launch_typed<S, bf16s>(s, m, n, a, lda, b, ldb);
    case REMIFA_R_8F_E4M3: return launch_typed<S, e4m3s>(s, m, n, a, lda, b, ldb);
  }
  return -3;
}

// strided copy of an m x n block of elements of `elt` bytes (2, 4 or 8)
template <typename T>
__global__ void __launch_bounds__(256) copy2d_kernel(int64_t m, int64_t n, const T* __restrict__ a,
                                                     int64_t lda, T* __restrict__ b, int64_t ldb) {
  const int64_t i0 = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
  if (i0 >= m) return;
  for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int64_t i = i0 + r;
      if (i < m) b[i + j * ldb] = a[i + j * lda];
    }
  }
}

}  // namespace

int convert_launch(remifa_handle_t h, int src_t, int dst_t, int64_t m, int64_t n, const void* a,
                   int64_t lda, void* b, int64_t ldb) {
  if (m == 0 || n == 0) return 0;
  hipStream_t s = h->stream;
  double bytes = (double)m * n * (dtype_size(src_t) + dtype_size(dst_t));
  timer_begin(h, KC_CONVERT, bytes);
  int st;
  switch (src_t) {
    case REMIFA_R_32F: st = launch_src<float>(s, dst_t, m, n, a, lda, b, ldb); break;
    case REMIFA_R_16F: st = launch_src<half_t>(s, dst_t, m, n, a, lda, b, ldb); break;
    case REMIFA_R_64F: st = launch_src<double>(s, dst_t, m, n, a, lda, b, ldb); break;
    case REMIFA_R_16BF: st = launch_src<bf16s>(s, dst_t, m, n, a, lda, b, ldb); break;
    case REMIFA_R_8F_E4M3: st = launch_src<e4m3s>(s, dst_t, m, n, a, lda, b, ldb); break;
    default: st = -2;
  }
  timer_end(h);
  return st;
}

int copy2d_launch(remifa_handle_t h, size_t elt, int64_t m, int64_t n, const void* a, int64_t lda,
                  void* b, int64_t ldb) {
  if (m == 0 || n == 0) return 0;
  dim3 grid((unsigned)ceil_div(m, 1024), (unsigned)(n < 65535 ? n : 65535));
  timer_begin(h, KC_OTHER, (double)m * n * elt * 2);
  switch (elt) {
    case 2:
      copy2d_kernel<uint16_t><<<grid, 256, 0, h->stream>>>(m, n, (const uint16_t*)a, lda,
                                                          (uint16_t*)b, ldb);
      break;
    case 4:
      copy2d_kernel<float><<<grid, 256, 0, h->stream>>>(m, n, (const float*)a, lda, (float*)b, ldb);
      break;
    case 8:
      copy2d_kernel<double><<<grid, 256, 0, h->stream>>>(m, n, (const double*)a, lda, (double*)b,
                                                        ldb);
      break;
    default: timer_end(h); return -2;
  }
  timer_end(h);
  RF_HIP(hipGetLastError());
  return 0;
}

}  // namespace remifa
