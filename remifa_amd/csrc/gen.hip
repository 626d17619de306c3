// Seeded matrix generators on the GPU.
//
// The reference draws a_ij ~ uniform(-1,1) row by row from libstdc++'s
// std::default_random_engine (minstd_rand0: x <- 16807 x mod 2^31-1) with
// std::uniform_real_distribution<float> (tests/genmat.hxx:68-102). minstd is
// a multiplicative LCG, so the state of draw k is seed * 16807^k mod (2^31-1):
// every lane jumps straight to its first draw and then steps along its
// columns, which reproduces the sequential stream bit for bit. Lanes take
// consecutive rows so the column-major stores coalesce.
#include "common.hpp"

namespace remifa {
namespace {

constexpr uint64_t kA = 16807ULL;
constexpr uint64_t kM = 2147483647ULL;  // 2^31 - 1

__device__ __forceinline__ uint64_t mod_m(uint64_t x) {  // x < 2^62
  x = (x & kM) + (x >> 31);
  x = (x & kM) + (x >> 31);
  return x >= kM ? x - kM : x;
}
__device__ __forceinline__ uint64_t mulmod(uint64_t a, uint64_t b) { return mod_m(a * b); }

__device__ uint64_t jump(uint64_t state, uint64_t k) {
  uint64_t r = 1, p = kA;
  while (k) {
    if (k & 1) r = mulmod(r, p);
    p = mulmod(p, p);
    k >>= 1;
  }
  return mulmod(state, r);
}

// uniform_real_distribution<float>(-1,1) of one engine output x:
// generate_canonical<float,24> = float(x-1) / 2^31 (clamped below 1), then
// ret * (b - a) + a.
__device__ __forceinline__ float uniform_m11(uint64_t x) {
  float ret = (float)(uint32_t)(x - 1) * 4.656612873077392578125e-10f;  // * 2^-31, exact
  if (ret >= 1.0f) ret = 0.999999940395355224609375f;                  // nextafter(1, 0)
  return ret * 2.0f + (-1.0f);
}

constexpr int kChunk = 128;  // columns per lane

__global__ void __launch_bounds__(256) gen_kernel(int64_t m, int64_t n, uint64_t s0,
                                                  float* __restrict__ a, int64_t lda,
                                                  double* __restrict__ rowsum) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const int64_t j0 = (int64_t)blockIdx.y * kChunk;
  if (j0 >= n) return;
  const int64_t j1 = j0 + kChunk < n ? j0 + kChunk : n;
  uint64_t x = jump(s0, (uint64_t)(i * n + j0));
  double s = 0.0;
  for (int64_t j = j0; j < j1; ++j) {
    x = mulmod(x, kA);
    float v = uniform_m11(x);
    a[i + j * lda] = v;
    if (j != i) s += fabs((double)v);
  }
  // every |a_ij| is a multiple of 2^-30 below 1: the FP64 sum is exact in
  // any order, so the atomic accumulation is deterministic
  if (rowsum) atomicAdd(rowsum + i, s);
}

// uniform_real_distribution<float>(lo, hi) of one engine output x
__device__ __forceinline__ float uniform_ab(uint64_t x, float lo, float hi) {
  float ret = (float)(uint32_t)(x - 1) * 4.656612873077392578125e-10f;
  if (ret >= 1.0f) ret = 0.999999940395355224609375f;
  return ret * (hi - lo) + lo;
}

// gen_randmat (tests/common.hxx:298-319): draws column by column, so entry
// (i, j) is draw j m + i; a lane takes row i and steps along its columns by
// A^m, the state m draws later.
__global__ void __launch_bounds__(256) gen_colmajor_kernel(int64_t m, int64_t n, uint64_t s0,
                                                           uint64_t am, float* __restrict__ a,
                                                           int64_t lda) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= m) return;
  const int64_t j0 = (int64_t)blockIdx.y * kChunk;
  if (j0 >= n) return;
  const int64_t j1 = j0 + kChunk < n ? j0 + kChunk : n;
  uint64_t x = jump(s0, (uint64_t)(j0 * m + i + 1));  // state of draw j0 m + i
  for (int64_t j = j0; j < j1; ++j) {
    a[i + j * lda] = uniform_m11(x);
    x = mulmod(x, am);
  }
}

// gen_arrowhead (tests/common.hxx:391-455, k = 256) for a square n x n
// matrix: the draw that each entry keeps, computed in closed form from the
// reference's loop order (two draws per visited pair, later writes win):
//  * diagonal blocks c (rows / columns [ck, min(n, (c+1)k))), visited j then
//    i: entry (r, col) with col > r keeps the first draw of iteration (j = col,
//    i = r), with col < r the second draw of iteration (j = r, i = col);
//  * then the arrow, K0 = max(n - k, 0): for every column j, rows i in
//    [max(K0, j), n): entries with r >= K0, col <= r keep the first draw of
//    (j = col, i = r), entries with col >= K0, r < col the second draw of
//    (j = r, i = col); the arrow overwrites the last diagonal block;
//  * the diagonal is 2k.
constexpr int64_t kArrow = 256;

__device__ __forceinline__ int64_t arrow_iter(int64_t n, int64_t K0, int64_t j, int64_t i) {
  // iterations of the arrow loop before (j, i): columns j' < j contribute
  // n - max(K0, j') each
  const int64_t kk = n - K0;
  int64_t S;
  if (j <= K0)
    S = j * kk;
  else
    S = K0 * kk + (j - K0) * n - (j - K0) * (K0 + j - 1) / 2;
  return S + (i - (j > K0 ? j : K0));
}

// (one-dimensional grid: column c = blockIdx.x / rb, rows of block
// blockIdx.x % rb, rb = ceil(n / 256))
__global__ void __launch_bounds__(256) gen_arrowhead_kernel(int64_t n, int64_t rb, float lo,
                                                            float hi, float* __restrict__ a,
                                                            int64_t lda) {
  const int64_t c = blockIdx.x / rb;
  const int64_t r = (blockIdx.x % rb) * 256 + threadIdx.x;
  if (r >= n) return;
  const int64_t K0 = n - kArrow > 0 ? n - kArrow : 0;
  float v = 0.0f;
  int64_t draw = -1;  // 0-based index of the kept draw
  if (r == c) {
    v = (float)(2 * kArrow);
  } else if (r >= K0 || c >= K0) {
    const int64_t nb = (n + kArrow - 1) / kArrow;  // diagonal blocks
    const int64_t last = n - (nb - 1) * kArrow;    // order of the last block
    const int64_t base = 2 * kArrow * kArrow * (nb - 1) + 2 * last * last;
    if (r > c)
      draw = base + 2 * arrow_iter(n, K0, c, r);
    else
      draw = base + 2 * arrow_iter(n, K0, r, c) + 1;
  } else if (r / kArrow == c / kArrow) {
    const int64_t b = r / kArrow, b0 = b * kArrow;
    const int64_t bs = (n - b0 < kArrow ? n - b0 : kArrow);
    const int64_t base = 2 * kArrow * kArrow * b;
    if (c > r)
      draw = base + 2 * ((c - b0) * bs + (r - b0));
    else
      draw = base + 2 * ((r - b0) * bs + (c - b0)) + 1;
  }
  if (draw >= 0) v = uniform_ab(jump(1, (uint64_t)draw + 1), lo, hi);
  a[r + c * lda] = v;
}

__global__ void set_diag_kernel(int64_t n, const double* __restrict__ rowsum, double alpha,
                                float* __restrict__ a, int64_t lda) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  a[i + i * lda] = alpha > 0.0 ? (float)(alpha * rowsum[i]) : (float)n;
}

__global__ void zero_f64_kernel(int64_t n, double* p) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

uint64_t seed_state(uint64_t seed) {
  uint64_t x = seed % kM;
  return x == 0 ? 1 : x;
}

}  // namespace
}  // namespace remifa

using namespace remifa;

extern "C" int remifa_gen_diagdom(remifa_handle_t h, int64_t n, uint64_t seed, double alpha,
                                  float* d_a, int64_t lda) {
  if (!h) return -1;
  if (n < 0) return -2;
  if (!d_a && n > 0) return -5;
  if (lda < (n > 1 ? n : 1)) return -6;
  if (n == 0) return 0;
  // gen_diagdom itself uses the default-constructed engine (seed 1)
  uint64_t s0 = alpha > 0.0 ? seed_state(seed) : 1;
  double* rowsum = nullptr;
  if (alpha > 0.0) {
    RF_CHECK(ensure_workspace(h, sizeof(double) * (size_t)n));
    rowsum = (double*)h->ws;
    zero_f64_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, h->stream>>>(n, rowsum);
  }
  dim3 grid((unsigned)ceil_div(n, 256), (unsigned)ceil_div(n, kChunk));
  gen_kernel<<<grid, 256, 0, h->stream>>>(n, n, s0, d_a, lda, rowsum);
  set_diag_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, h->stream>>>(n, rowsum, alpha, d_a, lda);
  RF_HIP(hipGetLastError());
  RF_HIP(hipStreamSynchronize(h->stream));
  return 0;
}

extern "C" int remifa_gen_randmat(remifa_handle_t h, int64_t m, int64_t n, uint64_t seed,
                                  float* d_a, int64_t lda) {
  if (!h) return -1;
  if (m < 0) return -2;
  if (n < 0) return -3;
  if (!d_a && m > 0 && n > 0) return -5;
  if (lda < (m > 1 ? m : 1)) return -6;
  if (m == 0 || n == 0) return 0;
  // A^m mod M, the step between the draws of one row in adjacent columns
  uint64_t am = 1, p = kA, e = (uint64_t)m;
  while (e) {
    if (e & 1) am = (am * p) % kM;
    p = (p * p) % kM;
    e >>= 1;
  }
  dim3 grid((unsigned)ceil_div(m, 256), (unsigned)ceil_div(n, kChunk));
  gen_colmajor_kernel<<<grid, 256, 0, h->stream>>>(m, n, seed_state(seed), am, d_a, lda);
  RF_HIP(hipGetLastError());
  return 0;
}

extern "C" int remifa_gen_arrowhead(remifa_handle_t h, int64_t n, float lo, float hi, float* d_a,
                                    int64_t lda) {
  if (!h) return -1;
  if (n < 0 || n > INT32_MAX) return -2;
  if (!(lo <= hi)) return -4;
  if (!d_a && n > 0) return -5;
  if (lda < (n > 1 ? n : 1)) return -6;
  if (n == 0) return 0;
  const int64_t rb = ceil_div(n, 256);
  if (rb * n > INT32_MAX) return -2;
  gen_arrowhead_kernel<<<(unsigned)(rb * n), 256, 0, h->stream>>>(n, rb, lo, hi, d_a, lda);
  RF_HIP(hipGetLastError());
  return 0;
}

extern "C" int remifa_gen_uniform(remifa_handle_t h, int64_t m, int64_t n, uint64_t seed,
                                  float* d_b, int64_t ldb) {
  if (!h) return -1;
  if (m < 0) return -2;
  if (n < 0) return -3;
  if (!d_b && m > 0 && n > 0) return -5;
  if (ldb < (m > 1 ? m : 1)) return -6;
  if (m == 0 || n == 0) return 0;
  dim3 grid((unsigned)ceil_div(m, 256), (unsigned)ceil_div(n, kChunk));
  gen_kernel<<<grid, 256, 0, h->stream>>>(m, n, seed_state(seed), d_b, ldb, nullptr);
  RF_HIP(hipGetLastError());
  return 0;
}
