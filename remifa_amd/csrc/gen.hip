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
