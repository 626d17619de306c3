// The reference's error measures (tests/errors.hxx) on the GPU, in FP64.
//
//  * normwise backward error ||b - A x||_inf / (||A||_1 ||x||_inf + ||b||_inf)
//    (unsym_backward_error, tests/errors.hxx:186-249), worst over the columns;
//  * componentwise backward error max_i |b - A x|_i / (|A||x| + |L||U||x|)_i
//    and the Oettli-Prager form max_i |b - A x|_i / (|A||x|)_i
//    (unsym_compwise_error, tests/errors.hxx:22-182; the reference takes one
//    right-hand side, here the worst over the columns);
//  * the factor error max_ij |A - L U|_ij / (|L||U|)_ij
//    (factor_backward_error, tests/errors.hxx:251-346), with the largest
//    |A - L U|_ij and ||A - L U||_F / ||A||_F beside it.
//
// A is fp32, the factors fp32 or fp16 (unit L strictly below the diagonal,
// U on and above it, in place), x and b FP64.
//
// Residual pass: one read of A per 16 right-hand sides gives r = b - A x and
// |A||x| (per-chunk partials summed in chunk order, so the result does not
// depend on scheduling). |L||U||x| is two triangular matrix-vector products
// with the absolute factors, |U||x| first, then t + |L| t, each a read of
// half the factors. The factor error is O(n^3): every 64 x 64 output tile
// sums sum_k L_ik U_kj and sum_k |L_ik||U_kj| over k <= min(i, j) in FP64
// (VALU fma with absolute-value source modifiers), L and U tiles staged in
// LDS; per-tile results are reduced in tile order.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <vector>

#include "common.hpp"

namespace remifa {
namespace {

constexpr int ER_ROWS = 256;  // rows per workgroup (one per lane)
constexpr int ER_COLS = 512;  // columns per chunk
constexpr int ER_RHS = 16;    // right-hand sides per pass

template <typename LT>
__device__ __forceinline__ double ldd(const LT* p) {
  return (double)(float)*p;
}

// part[(chunk * nr + k) * n + i] = sum over the chunk's columns j of a_ij x_jk
// (sums of products, ps) and |a_ij| |x_jk| (pa); MODE 0: A itself, MODE 1:
// |U| (j >= i) of the factors, MODE 2: |L| strictly below the diagonal.
// NR = 1: four interleaved accumulators combined in a fixed order (several
// loads in flight per lane); NR = 16: one accumulator per right-hand side.
template <typename LT, int MODE, int NR>
__global__ void __launch_bounds__(256) absmv_partial_kernel(int64_t n, int nr,
                                                            const LT* __restrict__ a, int64_t lda,
                                                            const double* __restrict__ x,
                                                            int64_t ldx, double* __restrict__ ps,
                                                            double* __restrict__ pa) {
  __shared__ double xs[ER_COLS][NR];
  const int64_t i = (int64_t)blockIdx.x * ER_ROWS + threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.y * ER_COLS;
  for (int e = threadIdx.x; e < ER_COLS * NR; e += 256) {
    const int c = e % ER_COLS, k = e / ER_COLS;
    xs[c][k] = (j0 + c < n && k < nr) ? x[j0 + c + k * ldx] : 0.0;
  }
  __syncthreads();
  if (i >= n) return;
  const int jn = (int)(n - j0 < ER_COLS ? n - j0 : ER_COLS);
  // columns of the chunk this row takes: all (A), j >= i (U), j < i (L)
  int jb = 0, je = jn;
  if (MODE == 1) jb = (int)std::max<int64_t>(0, std::min<int64_t>(jn, i - j0));
  if (MODE == 2) je = (int)std::max<int64_t>(0, std::min<int64_t>(jn, i - j0));
  const LT* ai = a + i + j0 * lda;
  if constexpr (NR == 1) {
    double s[4] = {0.0, 0.0, 0.0, 0.0}, sa[4] = {0.0, 0.0, 0.0, 0.0};
    int j = jb;
    for (; j + 8 <= je; j += 8) {
      double v[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) v[q] = ldd(ai + (int64_t)(j + q) * lda);
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        if (MODE == 0) s[q & 3] = __builtin_fma(v[q], xs[j + q][0], s[q & 3]);
        sa[q & 3] = __builtin_fma(fabs(v[q]), fabs(xs[j + q][0]), sa[q & 3]);
      }
    }
    for (; j < je; ++j) {
      const double v = ldd(ai + (int64_t)j * lda);
      if (MODE == 0) s[j & 3] = __builtin_fma(v, xs[j][0], s[j & 3]);
      sa[j & 3] = __builtin_fma(fabs(v), fabs(xs[j][0]), sa[j & 3]);
    }
    const int64_t o = (int64_t)blockIdx.y * n + i;
    if (MODE == 0) ps[o] = (s[0] + s[1]) + (s[2] + s[3]);
    pa[o] = (sa[0] + sa[1]) + (sa[2] + sa[3]);
  } else {
    double s[NR], sa[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) s[k] = sa[k] = 0.0;
    for (int j = jb; j < je; ++j) {
      const double v = ldd(ai + (int64_t)j * lda);
#pragma unroll
      for (int k = 0; k < NR; ++k) {
        if (MODE == 0) s[k] = __builtin_fma(v, xs[j][k], s[k]);
        sa[k] = __builtin_fma(fabs(v), fabs(xs[j][k]), sa[k]);
      }
    }
    for (int k = 0; k < nr; ++k) {
      const int64_t o = ((int64_t)blockIdx.y * nr + k) * n + i;
      if (MODE == 0) ps[o] = s[k];
      pa[o] = sa[k];
    }
  }
}

// per row: r = b - sum(ps), ax = sum(pa) (MODE 0), or t = sum(pa) (+ base)
__global__ void __launch_bounds__(256) sum_chunks_kernel(int64_t n, int nr, int nchunk,
                                                         const double* __restrict__ ps,
                                                         const double* __restrict__ pa,
                                                         const double* __restrict__ b,
                                                         int64_t ldb, double* __restrict__ r,
                                                         double* __restrict__ out,
                                                         const double* __restrict__ base) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= n) return;
  double s = 0.0, sa = 0.0;
  for (int c = 0; c < nchunk; ++c) {
    const int64_t o = ((int64_t)c * nr + k) * n + i;
    if (ps) s += ps[o];
    sa += pa[o];
  }
  if (r) r[i + k * n] = b[i + k * ldb] - s;
  out[i + k * n] = base ? base[i + k * n] + sa : sa;
}

// per-workgroup maxima of one column k: |r|, |x|, |b|, r / (ax + lux),
// r / ax (the reference's NaN handling: std::max keeps the running value
// when the new one is NaN, except for ||r|| which propagates NaN)
__global__ void __launch_bounds__(256) row_measures_kernel(int64_t n, const double* __restrict__ r,
                                                           const double* __restrict__ ax,
                                                           const double* __restrict__ lux,
                                                           const double* __restrict__ x,
                                                           int64_t ldx, const double* __restrict__ b,
                                                           int64_t ldb, double* __restrict__ out) {
  __shared__ double red[5][256];
  const int k = blockIdx.y;
  double mr = 0.0, mx = 0.0, mb = 0.0, cw = 0.0, cwa = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double ri = fabs(r[i + k * n]);
    mr = (ri > mr || ri != ri) ? ri : mr;  // NaN sticks
    mx = fmax(mx, fabs(x[i + k * ldx]));
    mb = fmax(mb, fabs(b[i + k * ldb]));
    const double a1 = ax[i + k * n];
    const double d1 = 1.0 / a1;
    if (d1 > 0.0) {
      const double v = ri * d1;
      cwa = v > cwa ? v : cwa;
    }
    double rr = ri;
    if (lux) {
      const double d2 = 1.0 / (a1 + lux[i + k * n]);
      if (d2 > 0.0) rr = rr * d2;
      cw = rr > cw ? rr : cw;
    }
  }
  red[0][threadIdx.x] = mr;
  red[1][threadIdx.x] = mx;
  red[2][threadIdx.x] = mb;
  red[3][threadIdx.x] = cw;
  red[4][threadIdx.x] = cwa;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      const double o = red[0][threadIdx.x + w], s = red[0][threadIdx.x];
      red[0][threadIdx.x] = (o > s || o != o) ? o : s;
      for (int q = 1; q < 5; ++q) {
        const double oq = red[q][threadIdx.x + w], sq = red[q][threadIdx.x];
        red[q][threadIdx.x] = oq > sq ? oq : sq;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x < 5)
    out[((int64_t)k * gridDim.x + blockIdx.x) * 5 + threadIdx.x] = red[threadIdx.x][0];
}

// column sums of |A| (one workgroup per column), FP64
__global__ void __launch_bounds__(256) abs_colsum_kernel(int64_t n, const float* __restrict__ a,
                                                         int64_t lda, double* __restrict__ cs) {
  __shared__ double red[256];
  const int64_t j = blockIdx.x;
  double s = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += fabs((double)a[i + j * lda]);
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) cs[j] = red[0];
}

// ---------------------------------------------------------------- factor error
constexpr int FT = 64;  // output tile edge
constexpr int FK = 16;  // k per LDS stage

// Per 64 x 64 tile (i0, j0): with the effective factors L_ik (1 on the
// diagonal, 0 above) and U_kj (0 below the diagonal), s = sum_k L_ik U_kj
// and d = sum_k |L_ik| |U_kj| over k < min(i0, j0) + 64; the tile's
// max |a - s| / d (d > 0; |a - s| where d = 0), max |a - s|, sum (a - s)^2,
// sum a^2 go to res[tile * 4 ..].
template <typename LT>
__global__ void __launch_bounds__(256) factor_error_kernel(int64_t n, const float* __restrict__ a,
                                                           int64_t lda, const LT* __restrict__ lu,
                                                           int64_t ldlu, int tiles_m,
                                                           double* __restrict__ res) {
  __shared__ double Ls[FK][FT + 1];  // Ls[k][i]
  __shared__ double Us[FK][FT + 1];  // Us[k][j]
  __shared__ double red[4][256];
  const int t = threadIdx.x;
  const int tile = blockIdx.x;
  const int64_t i0 = (int64_t)(tile % tiles_m) * FT, j0 = (int64_t)(tile / tiles_m) * FT;
  const int ti = (t % 16) * 4, tj = (t / 16) * 4;  // this lane's 4 x 4 block
  double s[4][4], d[4][4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) s[p][q] = d[p][q] = 0.0;
  const int64_t kend = std::min<int64_t>(n, std::min(i0, j0) + FT);
  for (int64_t k0 = 0; k0 < kend; k0 += FK) {
    // stage: 64 x 16 of L (rows i0.., cols k0..) and 16 x 64 of U
    for (int e = t; e < FT * FK; e += 256) {
      const int ii = e % FT, kk = e / FT;
      const int64_t gi = i0 + ii, gk = k0 + kk;
      double lv = 0.0;
      if (gi < n && gk < n) lv = gk < gi ? ldd(lu + gi + gk * ldlu) : (gk == gi ? 1.0 : 0.0);
      Ls[kk][ii] = lv;
      const int jj = e / FK, k2 = e % FK;
      const int64_t gj = j0 + jj, gk2 = k0 + k2;
      double uv = 0.0;
      if (gj < n && gk2 < n && gk2 <= gj) uv = ldd(lu + gk2 + gj * ldlu);
      Us[k2][jj] = uv;
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < FK; ++kk) {
      double lv[4], uv[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) lv[p] = Ls[kk][ti + p];
#pragma unroll
      for (int q = 0; q < 4; ++q) uv[q] = Us[kk][tj + q];
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          s[p][q] = __builtin_fma(lv[p], uv[q], s[p][q]);
          d[p][q] = __builtin_fma(fabs(lv[p]), fabs(uv[q]), d[p][q]);
        }
    }
    __syncthreads();
  }
  double emax = 0.0, eabs = 0.0, e2 = 0.0, a2 = 0.0;
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t gi = i0 + ti + p, gj = j0 + tj + q;
      if (gi < n && gj < n) {
        const double av = (double)a[gi + gj * lda];
        const double diff = av - s[p][q];
        double e = fabs(diff);
        eabs = fmax(eabs, e);
        if (d[p][q] > 0.0) e /= d[p][q];
        emax = fmax(emax, e);
        e2 = __builtin_fma(diff, diff, e2);
        a2 = __builtin_fma(av, av, a2);
      }
    }
  red[0][t] = emax;
  red[1][t] = eabs;
  red[2][t] = e2;
  red[3][t] = a2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      red[0][t] = fmax(red[0][t], red[0][t + w]);
      red[1][t] = fmax(red[1][t], red[1][t + w]);
      red[2][t] += red[2][t + w];
      red[3][t] += red[3][t + w];
    }
    __syncthreads();
  }
  if (t < 4) res[(int64_t)tile * 4 + t] = red[t][0];
}

// res[tile][4] -> out[0..3] in tile order (one workgroup)
__global__ void __launch_bounds__(256) factor_error_reduce_kernel(int64_t tiles,
                                                                  const double* __restrict__ res,
                                                                  double* __restrict__ out) {
  __shared__ double red[4][256];
  const int t = threadIdx.x;
  double m0 = 0.0, m1 = 0.0, s2 = 0.0, s3 = 0.0;
  // contiguous ranges per lane, then a fixed tree: deterministic
  const int64_t per = (tiles + 255) / 256;
  for (int64_t e = t * per; e < std::min<int64_t>(tiles, (t + 1) * per); ++e) {
    m0 = fmax(m0, res[e * 4]);
    m1 = fmax(m1, res[e * 4 + 1]);
    s2 += res[e * 4 + 2];
    s3 += res[e * 4 + 3];
  }
  red[0][t] = m0;
  red[1][t] = m1;
  red[2][t] = s2;
  red[3][t] = s3;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) {
      red[0][t] = fmax(red[0][t], red[0][t + w]);
      red[1][t] = fmax(red[1][t], red[1][t + w]);
      red[2][t] += red[2][t + w];
      red[3][t] += red[3][t + w];
    }
    __syncthreads();
  }
  if (t < 4) out[t] = red[t][0];
}

// tests/common.hxx:107-174 (device_conv_compwise_error): the converted copy
// widened to FP64, e_ij = |a_ij - a~_ij|, divided by |a_ij| where that is
// nonzero; res[0] = max e (relative), res[1] = max |a - a~| (absolute), both
// as the bit patterns of non-negative doubles (which order like the values),
// combined by atomicMax. One-dimensional grid: column j = blockIdx.x / rb.
template <typename CT>
__global__ void __launch_bounds__(256) conv_error_kernel(int64_t m, int64_t rb,
                                                         const float* __restrict__ a, int64_t lda,
                                                         const CT* __restrict__ ac, int64_t ldc,
                                                         unsigned long long* __restrict__ res) {
  __shared__ double red[2][256];
  const int64_t j = blockIdx.x / rb;
  const int64_t i = (blockIdx.x % rb) * 256 + threadIdx.x;
  double er = 0.0, ea = 0.0;
  if (i < m) {
    const double av = (double)a[i + j * lda];
    double cv;
    if constexpr (sizeof(CT) == 2 && std::is_same<CT, uint16_t>::value)
      cv = (double)bf16_bits_to_f32(ac[i + j * ldc]);
    else
      cv = (double)(float)ac[i + j * ldc];
    ea = fabs(av - cv);
    er = fabs(av) > 0.0 ? ea / fabs(av) : ea;
  }
  red[0][threadIdx.x] = er;
  red[1][threadIdx.x] = ea;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      red[0][threadIdx.x] = fmax(red[0][threadIdx.x], red[0][threadIdx.x + w]);
      red[1][threadIdx.x] = fmax(red[1][threadIdx.x], red[1][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 2)
    atomicMax(res + threadIdx.x, (unsigned long long)__double_as_longlong(red[threadIdx.x][0]));
}

__global__ void zero_u64x2_kernel(unsigned long long* p) {
  if (threadIdx.x < 2) p[threadIdx.x] = 0ull;
}

template <typename LT, int MODE>
void absmv_launch(hipStream_t s, int64_t n, int nr, const void* a, int64_t ld, const double* x,
                  int64_t ldx, double* ps, double* pa) {
  const dim3 g((unsigned)ceil_div(n, ER_ROWS), (unsigned)ceil_div(n, ER_COLS));
  if (nr == 1)
    absmv_partial_kernel<LT, MODE, 1><<<g, 256, 0, s>>>(n, nr, (const LT*)a, ld, x, ldx, ps, pa);
  else
    absmv_partial_kernel<LT, MODE, ER_RHS>
        <<<g, 256, 0, s>>>(n, nr, (const LT*)a, ld, x, ldx, ps, pa);
}

}  // namespace
}  // namespace remifa

using namespace remifa;

extern "C" {

int remifa_error_measures(remifa_handle_t h, int lu_t, int64_t n, int64_t nrhs, const float* d_a,
                          int64_t lda, const void* d_lu, int64_t ldlu, const double* d_b,
                          int64_t ldb, const double* d_x, int64_t ldx, double* out) {
  if (!h) return -1;
  if (lu_t != REMIFA_R_32F && lu_t != REMIFA_R_16F) return -2;
  if (n < 0) return -3;
  if (nrhs < 0) return -4;
  if (!d_a && n > 0) return -5;
  if (lda < std::max<int64_t>(n, 1)) return -6;
  if (d_lu && ldlu < std::max<int64_t>(n, 1)) return -8;
  if (!d_b && n > 0 && nrhs > 0) return -9;
  if (ldb < std::max<int64_t>(n, 1)) return -10;
  if (!d_x && n > 0 && nrhs > 0) return -11;
  if (ldx < std::max<int64_t>(n, 1)) return -12;
  if (!out) return -13;
  out[0] = out[1] = out[2] = 0.0;
  if (!d_lu) out[1] = __builtin_nan("");
  if (n == 0 || nrhs == 0) return 0;
  // workspace: partials (2 x chunks x 16 x n) | r, ax, t, lux (4 x 16 x n) |
  // column sums (n) | row-measure blocks
  const int64_t nchunk = ceil_div(n, ER_COLS);
  const int rows_wg = (int)std::min<int64_t>(ceil_div(n, 256), 64);
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_part = al(sizeof(double) * (size_t)nchunk * ER_RHS * (size_t)n);
  const size_t b_vec = al(sizeof(double) * ER_RHS * (size_t)n);
  const size_t b_cs = al(sizeof(double) * (size_t)n);
  const size_t b_blk = al(sizeof(double) * 5 * ER_RHS * rows_wg);
  RF_CHECK(ensure_workspace(h, 2 * b_part + 4 * b_vec + b_cs + b_blk));
  char* p = (char*)h->ws;
  double* ps = (double*)p;
  double* pa = (double*)(p + b_part);
  double* r = (double*)(p + 2 * b_part);
  double* ax = (double*)(p + 2 * b_part + b_vec);
  double* tt = (double*)(p + 2 * b_part + 2 * b_vec);
  double* lux = (double*)(p + 2 * b_part + 3 * b_vec);
  double* cs = (double*)(p + 2 * b_part + 4 * b_vec);
  double* blk = (double*)(p + 2 * b_part + 4 * b_vec + b_cs);
  hipStream_t s = h->stream;
  abs_colsum_kernel<<<(unsigned)n, 256, 0, s>>>(n, d_a, lda, cs);
  std::vector<double> hcs((size_t)n);
  RF_HIP(hipMemcpyAsync(hcs.data(), cs, sizeof(double) * n, hipMemcpyDeviceToHost, s));
  RF_HIP(hipStreamSynchronize(s));
  double anorm = 0.0;  // ||A||_1 (tests/errors.hxx:203-213)
  for (double v : hcs) anorm = std::max(anorm, v);
  std::vector<double> hb((size_t)5 * ER_RHS * rows_wg);
  double nw = 0.0, cw = 0.0, cwa = 0.0;
  bool nw_nan = false;
  for (int64_t k0 = 0; k0 < nrhs; k0 += ER_RHS) {
    const int nr = (int)std::min<int64_t>(ER_RHS, nrhs - k0);
    const double* xk = d_x + k0 * ldx;
    const double* bk = d_b + k0 * ldb;
    const dim3 gv((unsigned)ceil_div(n, 256), (unsigned)nr);
    absmv_launch<float, 0>(s, n, nr, d_a, lda, xk, ldx, ps, pa);
    sum_chunks_kernel<<<gv, 256, 0, s>>>(n, nr, (int)nchunk, ps, pa, bk, ldb, r, ax, nullptr);
    if (d_lu) {
      // t = |U| |x|, then lux = t + |L| t
      if (lu_t == REMIFA_R_16F)
        absmv_launch<half_t, 1>(s, n, nr, d_lu, ldlu, xk, ldx, nullptr, pa);
      else
        absmv_launch<float, 1>(s, n, nr, d_lu, ldlu, xk, ldx, nullptr, pa);
      sum_chunks_kernel<<<gv, 256, 0, s>>>(n, nr, (int)nchunk, nullptr, pa, nullptr, 0, nullptr,
                                           tt, nullptr);
      if (lu_t == REMIFA_R_16F)
        absmv_launch<half_t, 2>(s, n, nr, d_lu, ldlu, tt, n, nullptr, pa);
      else
        absmv_launch<float, 2>(s, n, nr, d_lu, ldlu, tt, n, nullptr, pa);
      sum_chunks_kernel<<<gv, 256, 0, s>>>(n, nr, (int)nchunk, nullptr, pa, nullptr, 0, nullptr,
                                           lux, tt);
    }
    row_measures_kernel<<<dim3((unsigned)rows_wg, (unsigned)nr), 256, 0, s>>>(
        n, r, ax, d_lu ? lux : nullptr, xk, ldx, bk, ldb, blk);
    RF_HIP(hipGetLastError());
    RF_HIP(hipMemcpyAsync(hb.data(), blk, sizeof(double) * 5 * nr * rows_wg,
                          hipMemcpyDeviceToHost, s));
    RF_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < nr; ++k) {
      double mr = 0.0, mx = 0.0, mb = 0.0, ck = 0.0, cak = 0.0;
      bool rnan = false;
      for (int w = 0; w < rows_wg; ++w) {
        const double* q = &hb[((size_t)k * rows_wg + w) * 5];
        if (q[0] != q[0]) rnan = true;
        mr = std::max(mr, q[0]);
        mx = std::max(mx, q[1]);
        mb = std::max(mb, q[2]);
        ck = std::max(ck, q[3]);
        cak = std::max(cak, q[4]);
      }
      const double e = rnan ? __builtin_nan("") : mr / (anorm * mx + mb);
      if (e != e) nw_nan = true;
      nw = std::max(nw, e);
      cw = std::max(cw, ck);
      cwa = std::max(cwa, cak);
    }
  }
  out[0] = nw_nan ? __builtin_nan("") : nw;
  if (d_lu) out[1] = cw;
  out[2] = cwa;
  return 0;
}

int remifa_factor_error(remifa_handle_t h, int lu_t, int64_t n, const float* d_a, int64_t lda,
                        const void* d_lu, int64_t ldlu, double* out) {
  if (!h) return -1;
  if (lu_t != REMIFA_R_32F && lu_t != REMIFA_R_16F) return -2;
  if (n < 0) return -3;
  if (!d_a && n > 0) return -4;
  if (lda < std::max<int64_t>(n, 1)) return -5;
  if (!d_lu && n > 0) return -6;
  if (ldlu < std::max<int64_t>(n, 1)) return -7;
  if (!out) return -8;
  out[0] = out[1] = out[2] = out[3] = 0.0;
  if (n == 0) return 0;
  const int64_t tm = ceil_div(n, FT);
  const int64_t tiles = tm * tm;
  if (tiles > INT32_MAX) return -3;
  RF_CHECK(ensure_workspace(h, sizeof(double) * (size_t)(tiles * 4 + 4)));
  double* res = (double*)h->ws;
  double* fin = res + tiles * 4;
  hipStream_t s = h->stream;
  timer_begin(h, KC_OTHER, 4.0 / 3.0 * (double)n * n * n, 3);
  if (lu_t == REMIFA_R_16F)
    factor_error_kernel<half_t><<<(unsigned)tiles, 256, 0, s>>>(n, d_a, lda, (const half_t*)d_lu,
                                                                ldlu, (int)tm, res);
  else
    factor_error_kernel<float><<<(unsigned)tiles, 256, 0, s>>>(n, d_a, lda, (const float*)d_lu,
                                                               ldlu, (int)tm, res);
  timer_end(h);
  factor_error_reduce_kernel<<<1, 256, 0, s>>>(tiles, res, fin);
  RF_HIP(hipGetLastError());
  double hf[4];
  RF_HIP(hipMemcpyAsync(hf, fin, sizeof(hf), hipMemcpyDeviceToHost, s));
  RF_HIP(hipStreamSynchronize(s));
  out[0] = hf[0];
  out[1] = hf[1];
  out[2] = hf[3] > 0.0 ? std::sqrt(hf[2] / hf[3]) : 0.0;
  out[3] = std::sqrt(hf[2]);
  return 0;
}

}  // extern "C"

extern "C" int remifa_conv_error(remifa_handle_t h, int conv_t, int64_t m, int64_t n,
                                 const float* d_a, int64_t lda, const void* d_conv, int64_t ldc,
                                 double* out) {
  if (!h) return -1;
  if (conv_t != REMIFA_R_16F && conv_t != REMIFA_R_16BF && conv_t != REMIFA_R_32F) return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (!d_a && m > 0 && n > 0) return -5;
  if (lda < std::max<int64_t>(m, 1)) return -6;
  if (!d_conv && m > 0 && n > 0) return -7;
  if (ldc < std::max<int64_t>(m, 1)) return -8;
  if (!out) return -9;
  out[0] = out[1] = 0.0;
  if (m == 0 || n == 0) return 0;
  const int64_t rb = ceil_div(m, 256);
  if (rb * n > INT32_MAX) return -4;
  RF_CHECK(ensure_workspace(h, 2 * sizeof(unsigned long long)));
  unsigned long long* res = (unsigned long long*)h->ws;
  hipStream_t s = h->stream;
  zero_u64x2_kernel<<<1, 64, 0, s>>>(res);
  const unsigned g = (unsigned)(rb * n);
  if (conv_t == REMIFA_R_16F)
    conv_error_kernel<half_t><<<g, 256, 0, s>>>(m, rb, d_a, lda, (const half_t*)d_conv, ldc, res);
  else if (conv_t == REMIFA_R_16BF)
    conv_error_kernel<uint16_t><<<g, 256, 0, s>>>(m, rb, d_a, lda, (const uint16_t*)d_conv, ldc,
                                                  res);
  else
    conv_error_kernel<float><<<g, 256, 0, s>>>(m, rb, d_a, lda, (const float*)d_conv, ldc, res);
  RF_HIP(hipGetLastError());
  unsigned long long hr[2];
  RF_HIP(hipMemcpyAsync(hr, res, sizeof(hr), hipMemcpyDeviceToHost, s));
  RF_HIP(hipStreamSynchronize(s));
  out[0] = __builtin_bit_cast(double, hr[0]);
  out[1] = __builtin_bit_cast(double, hr[1]);
  return 0;
}
