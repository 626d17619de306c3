// Equilibration on the GPU: the power-of-two row and column scalings of
// LAPACK's sgeequb and their application as slaqge does, which the
// reference harness runs on the host before factoring (--scale,
// tests/testing_lu_gpu.hxx:288-311, through src/wrappers.cxx:209-250).
//
// geequb: r_i = 1 / 2^trunc(log2 max_j |a_ij|), c_j = 1 / 2^trunc(log2
// max_i |a_ij| r_i) (each clamped to [SMLNUM, BIGNUM] before inverting),
// ROWCND / COLCND = min / max of the unclamped powers, AMAX = the largest row
// power; a zero row i (column j) stops with INFO = i (m + j), 1-based.
// laqge: with THRESH = 0.1 and SMALL = 2^-103 (SLAMCH('S') / SLAMCH('P')):
// no scaling, columns only, rows only or both (a_ij = (c_j r_i) a_ij);
// powers of two multiply exactly, so the scaled matrix is exact barring
// overflow or underflow.
//
// HBM-bound: the row maxima are one read of A (lanes on consecutive rows,
// column chunks in the grid, maxima combined by integer atomicMax on the
// bit patterns of non-negative floats, which orders like the values), the
// column maxima one more read (a workgroup per column), the scaling one read
// and one write. trunc(log2 x) is taken exactly from the exponent.
#include <algorithm>
#include <cfloat>

#include "common.hpp"

namespace remifa {
namespace {

constexpr int EQ_COLS = 256;  // columns per workgroup of the row-maxima pass

__device__ __forceinline__ float pow2_trunc_log2(float x) {
  int e;
  const float f = frexpf(x, &e);  // x = f 2^e, f in [0.5, 1)
  const int k = (x >= 1.0f || f == 0.5f) ? e - 1 : e;
  return ldexpf(1.0f, k);
}

__global__ void fill_u32_kernel(unsigned* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// rmax[i] = max_j |a_ij| (as float bits)
__global__ void __launch_bounds__(256) row_absmax_kernel(int64_t m, int64_t n,
                                                         const float* __restrict__ a, int64_t lda,
                                                         unsigned* __restrict__ rmax) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const int64_t j0 = (int64_t)blockIdx.y * EQ_COLS;
  const int64_t j1 = std::min<int64_t>(n, j0 + EQ_COLS);
  float mx = 0.0f;
  for (int64_t j = j0; j < j1; ++j) mx = fmaxf(mx, fabsf(a[i + j * lda]));
  atomicMax(rmax + i, __float_as_uint(mx));
}

// one workgroup: powers of the row maxima, their min / max, INFO, then r
// (out[0] = rowcnd, out[1] = amax, info[0] = first zero row + 1 or 0)
__global__ void __launch_bounds__(256) row_scale_kernel(int64_t m, const unsigned* __restrict__ rmax,
                                                        float* __restrict__ r,
                                                        float* __restrict__ out,
                                                        int* __restrict__ info) {
  __shared__ float smin[256], smax[256];
  __shared__ int szero[256];
  const int t = threadIdx.x;
  float mn = FLT_MAX, mx = 0.0f;
  int z = INT32_MAX;
  for (int64_t i = t; i < m; i += 256) {
    float v = __uint_as_float(rmax[i]);
    if (v > 0.0f) v = pow2_trunc_log2(v);
    r[i] = v;
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
    if (v == 0.0f) z = std::min<int>(z, (int)i);
  }
  smin[t] = mn;
  smax[t] = mx;
  szero[t] = z;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      smin[t] = fminf(smin[t], smin[t + w]);
      smax[t] = fmaxf(smax[t], smax[t + w]);
      szero[t] = std::min(szero[t], szero[t + w]);
    }
    __syncthreads();
  }
  const float smlnum = FLT_MIN, bignum = 1.0f / FLT_MIN;
  const float rcmin = fminf(smin[0], bignum), rcmax = smax[0];
  if (t == 0) {
    out[1] = rcmax;
    info[0] = szero[0] == INT32_MAX ? 0 : szero[0] + 1;
    out[0] = fmaxf(rcmin, smlnum) / fminf(rcmax, bignum);
  }
  if (szero[0] != INT32_MAX) return;  // uniform
  for (int64_t i = t; i < m; i += 256) r[i] = 1.0f / fminf(fmaxf(r[i], smlnum), bignum);
}

// c[j] = 2^trunc(log2 max_i |a_ij| r_i) (one workgroup per column)
__global__ void __launch_bounds__(256) col_absmax_kernel(int64_t m, const float* __restrict__ a,
                                                         int64_t lda, const float* __restrict__ r,
                                                         const int* __restrict__ info,
                                                         float* __restrict__ c) {
  __shared__ float red[256];
  if (info[0] != 0) return;  // a zero row: LAPACK returns before the columns
  const int64_t j = blockIdx.x;
  float mx = 0.0f;
  for (int64_t i = threadIdx.x; i < m; i += 256) mx = fmaxf(mx, fabsf(a[i + j * lda]) * r[i]);
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) c[j] = red[0] > 0.0f ? pow2_trunc_log2(red[0]) : 0.0f;
}

// one workgroup: min / max of the column powers, INFO, then c
// (out[2] = colcnd; info[0] = m + first zero column + 1 if no zero row)
__global__ void __launch_bounds__(256) col_scale_kernel(int64_t m, int64_t n, float* __restrict__ c,
                                                        float* __restrict__ out,
                                                        int* __restrict__ info) {
  __shared__ float smin[256], smax[256];
  __shared__ int szero[256];
  if (info[0] != 0) return;
  const int t = threadIdx.x;
  float mn = FLT_MAX, mx = 0.0f;
  int z = INT32_MAX;
  for (int64_t j = t; j < n; j += 256) {
    const float v = c[j];
    mn = fminf(mn, v);
    mx = fmaxf(mx, v);
    if (v == 0.0f) z = std::min<int>(z, (int)j);
  }
  smin[t] = mn;
  smax[t] = mx;
  szero[t] = z;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      smin[t] = fminf(smin[t], smin[t + w]);
      smax[t] = fmaxf(smax[t], smax[t + w]);
      szero[t] = std::min(szero[t], szero[t + w]);
    }
    __syncthreads();
  }
  const float smlnum = FLT_MIN, bignum = 1.0f / FLT_MIN;
  if (szero[0] != INT32_MAX) {
    if (t == 0) info[0] = (int)(m + szero[0] + 1);
    return;
  }
  if (t == 0) out[2] = fmaxf(fminf(smin[0], bignum), smlnum) / fminf(smax[0], bignum);
  for (int64_t j = t; j < n; j += 256) c[j] = 1.0f / fminf(fmaxf(c[j], smlnum), bignum);
}

// a_ij <- s a_ij for EQUED 'R' (r_i), 'C' (c_j), 'B' ((c_j r_i) a_ij)
// (one-dimensional grid: column j = blockIdx.x / rb, rows of block
// blockIdx.x % rb, rb = ceil(m / 256))
template <int MODE>
__global__ void __launch_bounds__(256) scale_kernel(int64_t m, int64_t rb, float* __restrict__ a,
                                                    int64_t lda, const float* __restrict__ r,
                                                    const float* __restrict__ c) {
  const int64_t j = blockIdx.x / rb;
  const int64_t i = (blockIdx.x % rb) * 256 + threadIdx.x;
  if (i >= m) return;
  float* p = a + i + j * lda;
  if constexpr (MODE == 0) *p = r[i] * *p;
  if constexpr (MODE == 1) *p = c[j] * *p;
  if constexpr (MODE == 2) *p = (c[j] * r[i]) * *p;
}

}  // namespace
}  // namespace remifa

using namespace remifa;

extern "C" {

int remifa_geequb(remifa_handle_t h, int64_t m, int64_t n, const float* d_a, int64_t lda,
                  float* d_r, float* d_c, float* rowcnd, float* colcnd, float* amax, int* info) {
  if (!h) return -1;
  if (m < 0 || m > INT32_MAX) return -2;
  if (n < 0 || n > INT32_MAX) return -3;
  if (!d_a && m > 0 && n > 0) return -4;
  if (lda < std::max<int64_t>(m, 1)) return -5;
  if (!d_r && m > 0) return -6;
  if (!d_c && n > 0) return -7;
  if (!rowcnd) return -8;
  if (!colcnd) return -9;
  if (!amax) return -10;
  if (!info) return -11;
  *info = 0;
  if (m == 0 || n == 0) {
    *rowcnd = 1.0f;
    *colcnd = 1.0f;
    *amax = 0.0f;
    return 0;
  }
  // workspace: row maxima (m), results (rowcnd, amax, colcnd), info
  RF_CHECK(ensure_workspace(h, sizeof(unsigned) * (size_t)m + 256));
  unsigned* rmax = (unsigned*)h->ws;
  char* tail = (char*)h->ws + ((sizeof(unsigned) * (size_t)m + 127) & ~(size_t)127);
  float* out = (float*)tail;
  int* dinfo = (int*)(tail + 64);
  hipStream_t s = h->stream;
  fill_u32_kernel<<<(unsigned)ceil_div(m, 256), 256, 0, s>>>(rmax, m);
  row_absmax_kernel<<<dim3((unsigned)ceil_div(m, 256), (unsigned)ceil_div(n, EQ_COLS)), 256, 0,
                      s>>>(m, n, d_a, lda, rmax);
  row_scale_kernel<<<1, 256, 0, s>>>(m, rmax, d_r, out, dinfo);
  col_absmax_kernel<<<(unsigned)n, 256, 0, s>>>(m, d_a, lda, d_r, dinfo, d_c);
  col_scale_kernel<<<1, 256, 0, s>>>(m, n, d_c, out, dinfo);
  RF_HIP(hipGetLastError());
  float ho[3] = {0.0f, 0.0f, 1.0f};
  int hi = 0;
  RF_HIP(hipMemcpyAsync(ho, out, sizeof(ho), hipMemcpyDeviceToHost, s));
  RF_HIP(hipMemcpyAsync(&hi, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  RF_HIP(hipStreamSynchronize(s));
  *info = hi;
  *amax = ho[1];
  if (hi > 0 && hi <= m) return 0;  // zero row: LAPACK leaves rowcnd / colcnd unset
  *rowcnd = ho[0];
  if (hi > m) return 0;  // zero column: colcnd unset
  *colcnd = ho[2];
  return 0;
}

int remifa_laqge(remifa_handle_t h, int64_t m, int64_t n, float* d_a, int64_t lda,
                 const float* d_r, const float* d_c, float rowcnd, float colcnd, float amax,
                 char* equed) {
  if (!h) return -1;
  if (m < 0 || n < 0) return m < 0 ? -2 : -3;
  if (n > INT32_MAX) return -3;
  if (!d_a && m > 0 && n > 0) return -4;
  if (lda < std::max<int64_t>(m, 1)) return -5;
  if (!equed) return -11;
  if (m == 0 || n == 0) {
    *equed = 'N';
    return 0;
  }
  const float thresh = 0.1f;
  const float small = ldexpf(1.0f, -103), large = 1.0f / small;  // SLAMCH('S') / SLAMCH('P')
  int mode;
  if (rowcnd >= thresh && amax >= small && amax <= large) {
    if (colcnd >= thresh) {
      *equed = 'N';
      return 0;
    }
    mode = 1;
    *equed = 'C';
  } else if (colcnd >= thresh) {
    mode = 0;
    *equed = 'R';
  } else {
    mode = 2;
    *equed = 'B';
  }
  if ((mode != 1 && !d_r)) return -6;
  if ((mode != 0 && !d_c)) return -7;
  const int64_t rb = ceil_div(m, 256);
  if (rb * n > INT32_MAX) return -2;
  const unsigned g = (unsigned)(rb * n);
  hipStream_t s = h->stream;
  timer_begin(h, KC_CONVERT, 8.0 * (double)m * (double)n, 3);
  if (mode == 0) scale_kernel<0><<<g, 256, 0, s>>>(m, rb, d_a, lda, d_r, d_c);
  if (mode == 1) scale_kernel<1><<<g, 256, 0, s>>>(m, rb, d_a, lda, d_r, d_c);
  if (mode == 2) scale_kernel<2><<<g, 256, 0, s>>>(m, rb, d_a, lda, d_r, d_c);
  timer_end(h);
  RF_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"
