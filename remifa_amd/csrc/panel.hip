// lu_nopiv_panel<T, ib> (src/lu_nopiv.cu:330-400): unpivoted factorization
// of an ib-wide block column / block row.
//
// One 64-lane wavefront per workgroup. Every workgroup factors the n x n
// diagonal tile redundantly in LDS from the snapshot d (as every reference
// thread block does, :123-170), then
//   blockIdx.y == 0: 64 rows of the block column, one row per lane:
//                    l_ik <- l_ik * (1/u_kk); l_ij -= l_ik u_kj   (j > k)
//   blockIdx.y == 1: 64 columns of the block row, one column per lane:
//                    u_ij -= l_ik u_kj                           (i > k)
// Workgroup (0,0) stores the factored diagonal tile. The arithmetic per
// element is the reference's: correctly rounded reciprocal, multiply, then a
// fused multiply-subtract, column k by column k.
#include "common.hpp"

namespace remifa {
namespace {

template <typename T>
struct Arith;
template <>
struct Arith<float> {
  static __device__ __forceinline__ float rcp(float x) { return 1.0f / x; }
  static __device__ __forceinline__ float mul(float a, float b) { return a * b; }
  static __device__ __forceinline__ float fnma(float a, float b, float c) {
    return __builtin_fmaf(-a, b, c);
  }
};
template <>
struct Arith<double> {
  static __device__ __forceinline__ double rcp(double x) { return 1.0 / x; }
  static __device__ __forceinline__ double mul(double a, double b) { return a * b; }
  static __device__ __forceinline__ double fnma(double a, double b, double c) {
    return __builtin_fma(-a, b, c);
  }
};
template <>
struct Arith<half_t> {
  // reciprocal: fp32 correctly rounded, then rounded to fp16
  static __device__ __forceinline__ half_t rcp(half_t x) { return (half_t)(1.0f / (float)x); }
  static __device__ __forceinline__ half_t mul(half_t a, half_t b) { return a * b; }
  static __device__ __forceinline__ half_t fnma(half_t a, half_t b, half_t c) {
    return __builtin_fmaf16(-a, b, c);
  }
};

template <typename T, int IB>
__global__ void __launch_bounds__(64) panel_kernel(int64_t m, int64_t n, const T* __restrict__ d,
                                                   int64_t ldd, T* __restrict__ l, int64_t ldl,
                                                   T* __restrict__ u, int64_t ldu, int* stat) {
  using A = Arith<T>;
  __shared__ T D[IB][IB + 1];  // D[i][j]
  __shared__ T R[IB];
  const int t = threadIdx.x;
  const int by = blockIdx.y;
  if (by == 1 && m <= n) return;

  // load the diagonal tile snapshot: lane t < n owns row t
  if (t < n) {
#pragma unroll
    for (int j = 0; j < IB; ++j)
      if (j < n) D[t][j] = d[t + j * ldd];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < IB; ++k) {
    if (k < n) {
      T r = A::rcp(D[k][k]);
      if (t == 0) R[k] = r;
      if (t > k && t < n) D[t][k] = A::mul(D[t][k], r);
      __syncthreads();
      if (t > k && t < n) {
        T ltk = D[t][k];
#pragma unroll
        for (int j = 0; j < IB; ++j)
          if (j > k && j < n) D[t][j] = A::fnma(ltk, D[k][j], D[t][j]);
      }
      __syncthreads();
    }
  }

  const int64_t idx = n + (int64_t)blockIdx.x * 64 + t;  // row (L) or column (U)
  if (by == 0) {
    if (blockIdx.x == 0 && t < n) {
#pragma unroll
      for (int j = 0; j < IB; ++j)
        if (j < n) l[t + j * ldl] = D[t][j];
    }
    if (idx < m) {
      T v[IB];
#pragma unroll
      for (int j = 0; j < IB; ++j) v[j] = (j < n) ? l[idx + j * ldl] : (T)0;
#pragma unroll
      for (int k = 0; k < IB; ++k) {
        if (k < n) {
          T lik = A::mul(v[k], R[k]);
          v[k] = lik;
#pragma unroll
          for (int j = 0; j < IB; ++j)
            if (j > k && j < n) v[j] = A::fnma(lik, D[k][j], v[j]);
        }
      }
#pragma unroll
      for (int j = 0; j < IB; ++j)
        if (j < n) l[idx + j * ldl] = v[j];
    }
    if (blockIdx.x == 0 && t == 0 && stat) stat[0] = (int)n;  // src/lu_nopiv.cu:173-174
  } else if (idx < m) {
    T v[IB];
#pragma unroll
    for (int i = 0; i < IB; ++i) v[i] = (i < n) ? u[i + idx * ldu] : (T)0;
#pragma unroll
    for (int k = 0; k < IB; ++k) {
      if (k < n) {
        T ukj = v[k];
#pragma unroll
        for (int i = 0; i < IB; ++i)
          if (i > k && i < n) v[i] = A::fnma(D[i][k], ukj, v[i]);
      }
    }
#pragma unroll
    for (int i = 0; i < IB; ++i)
      if (i < n) u[i + idx * ldu] = v[i];
  }
}

template <typename T, int IB>
int launch_panel(hipStream_t s, int64_t m, int64_t n, const void* d, int64_t ldd, void* l,
                 int64_t ldl, void* u, int64_t ldu, int* stat) {
  int64_t rest = m - n;
  unsigned gx = (unsigned)(rest > 0 ? ceil_div(rest, 64) : 1);
  dim3 grid(gx, 2);
  panel_kernel<T, IB><<<grid, 64, 0, s>>>(m, n, (const T*)d, ldd, (T*)l, ldl, (T*)u, ldu, stat);
  RF_HIP(hipGetLastError());
  return 0;
}

template <typename T>
int dispatch_ib(hipStream_t s, int ib, int64_t m, int64_t n, const void* d, int64_t ldd, void* l,
                int64_t ldl, void* u, int64_t ldu, int* stat) {
  if (ib <= 8) return launch_panel<T, 8>(s, m, n, d, ldd, l, ldl, u, ldu, stat);
  if (ib <= 16) return launch_panel<T, 16>(s, m, n, d, ldd, l, ldl, u, ldu, stat);
  return launch_panel<T, 32>(s, m, n, d, ldd, l, ldl, u, ldu, stat);
}

}  // namespace

int panel_launch(remifa_handle_t h, int t, int ib, int64_t m, int64_t n, const void* d,
                 int64_t ldd, void* l, int64_t ldl, void* u, int64_t ldu, int* stat) {
  if (n == 0) return 0;
  size_t es = dtype_size(t);
  // algorithmic bytes: the block column and block row, read and written once
  double bytes = 2.0 * ((double)m * n + (double)n * (double)(m > n ? m - n : 0)) * es;
  timer_begin(h, KC_PANEL, bytes);
  int st;
  switch (t) {
    case REMIFA_R_32F: st = dispatch_ib<float>(h->stream, ib, m, n, d, ldd, l, ldl, u, ldu, stat); break;
    case REMIFA_R_16F: st = dispatch_ib<half_t>(h->stream, ib, m, n, d, ldd, l, ldl, u, ldu, stat); break;
    case REMIFA_R_64F: st = dispatch_ib<double>(h->stream, ib, m, n, d, ldd, l, ldl, u, ldu, stat); break;
    default: st = -2;
  }
  timer_end(h);
  return st;
}

}  // namespace remifa
