// Triangular solves with the stored no-pivot factors, the FP64 residual and
// iterative refinement.
//
// TRSM (forward L, backward U and the transposed pair for A^T): one launch
// per triangle. Workgroups take 64-row block rows in dependency order from an
// atomic ticket, so every block a workgroup waits for belongs to a workgroup
// that is already running (no deadlock whatever the residency). Block row i
// accumulates B_i - sum_j T_ij X_j, waiting on a per-block flag (agent-scope
// release/acquire) for each X_j it needs, then solves its diagonal block in
// LDS and publishes X_i. Every wait is bounded: on timeout the kernel records
// an error and exits, so the grid always drains. The factors are read once
// (HBM-bound); the critical path is one diagonal-block solve per block row.
//
// Residual R = B - A X in FP64: 2-D grid over (256-row block, 1024-column
// chunk), fp64 fma per element, per-chunk partials summed in chunk order by a
// second kernel (deterministic).
#include <algorithm>
#include <vector>

#include "common.hpp"

namespace remifa {
namespace {

constexpr int NBS = 64;  // block size of the triangular solves
constexpr int RC = 16;   // right-hand sides per launch
constexpr long long kSpinLimit = 1LL << 28;

template <typename LT>
__device__ __forceinline__ float ldf(const LT* p) {
  return (float)*p;
}

// element T[r][c] of the effective triangle
template <typename LT, bool TRANS>
__device__ __forceinline__ float tel(const LT* lu, int64_t ld, int64_t r, int64_t c) {
  return TRANS ? ldf(lu + c + r * ld) : ldf(lu + r + c * ld);
}

template <typename LT, bool LOWER, bool UNIT, bool TRANS>
__global__ void __launch_bounds__(256) trsm_kernel(int64_t n, int nrhs, const LT* __restrict__ lu,
                                                   int64_t ld, float* __restrict__ b, int64_t ldb,
                                                   int* __restrict__ sync, int nblk) {
  // sync[0] = ticket counter, sync[1] = error flag, sync[2 + i] = block i done
  __shared__ float Ts[NBS][NBS + 1];
  __shared__ float Xs[NBS][RC];
  __shared__ int s_ticket;
  const int t = threadIdx.x;
  const int row = t % NBS, kg = t / NBS;  // rhs kg*4 .. kg*4+3
  if (t == 0)
    s_ticket = __hip_atomic_fetch_add(&sync[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int tk = s_ticket;
  const int bi = LOWER ? tk : nblk - 1 - tk;
  const int64_t r0 = (int64_t)bi * NBS;
  const int64_t gr = r0 + row;
  float acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int k = kg * 4 + q;
    acc[q] = (gr < n && k < nrhs) ? b[gr + k * ldb] : 0.0f;
  }
  auto load_T = [&](int bj) {
    const int64_t c0 = (int64_t)bj * NBS;
    // 64 x 64 block: threads walk the contiguous direction of the storage
    for (int e = t; e < NBS * NBS; e += 256) {
      int a = e % NBS, bq = e / NBS;
      int rr = TRANS ? bq : a, cc = TRANS ? a : bq;
      int64_t R = r0 + rr, Cc = c0 + cc;
      // rows/columns past n are padded with the identity so that the
      // substitution leaves them at zero (no 0/0 in a ragged last block)
      Ts[rr][cc] = (R < n && Cc < n) ? tel<LT, TRANS>(lu, ld, R, Cc) : (R == Cc ? 1.0f : 0.0f);
    }
  };
  // off-diagonal blocks, in dependency order
  for (int s = 0; s < (LOWER ? bi : nblk - 1 - bi); ++s) {
    const int bj = LOWER ? s : nblk - 1 - s;
    load_T(bj);
    if (t == 0) {
      long long spins = 0;
      while (__hip_atomic_load(&sync[2 + bj], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0) {
        if (++spins > kSpinLimit ||
            __hip_atomic_load(&sync[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
          __hip_atomic_store(&sync[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __syncthreads();
    // X_j was published with agent-coherent stores before its flag; read it
    // with agent-coherent loads (they bypass the non-coherent caches, so no
    // cache invalidation is needed)
    {
      const int64_t c0 = (int64_t)bj * NBS;
      for (int e = t; e < NBS * RC; e += 256) {
        int c = e % NBS, k = e / NBS;
        int64_t C = c0 + c;
        Xs[c][k] = (C < n && k < nrhs)
                       ? __hip_atomic_load(&b[C + k * ldb], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                       : 0.0f;
      }
    }
    __syncthreads();
#pragma unroll 8
    for (int c = 0; c < NBS; ++c) {
      float tv = Ts[row][c];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = __builtin_fmaf(-tv, Xs[c][kg * 4 + q], acc[q]);
    }
    __syncthreads();
  }
  // diagonal block: substitution in LDS
  load_T(bi);
  __syncthreads();
  for (int s = 0; s < NBS; ++s) {
    const int c = LOWER ? s : NBS - 1 - s;
    if (row == c) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float x = acc[q];
        if (!UNIT) x = x / Ts[c][c];
        acc[q] = x;
        Xs[c][kg * 4 + q] = x;
      }
    }
    __syncthreads();
    if (LOWER ? row > c : row < c) {
      float tv = Ts[row][c];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = __builtin_fmaf(-tv, Xs[c][kg * 4 + q], acc[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    int k = kg * 4 + q;
    if (gr < n && k < nrhs)
      __hip_atomic_store(&b[gr + k * ldb], acc[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __syncthreads();
  if (t == 0) __hip_atomic_store(&sync[2 + bi], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void fill_int_kernel(int* p, int n, int v) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- residual
constexpr int RES_ROWS = 256;
constexpr int RES_COLS = 512;
constexpr int RES_RHS = 16;

__global__ void __launch_bounds__(256) resid_partial_kernel(int64_t n, int nrhs,
                                                            const float* __restrict__ a,
                                                            int64_t lda,
                                                            const double* __restrict__ x,
                                                            int64_t ldx, double* __restrict__ part) {
  // part[chunk][rhs][n]
  __shared__ double xs[RES_COLS][RES_RHS];
  const int64_t i = (int64_t)blockIdx.x * RES_ROWS + threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.y * RES_COLS;
  const int64_t j1 = std::min<int64_t>(j0 + RES_COLS, n);
  for (int e = threadIdx.x; e < RES_COLS * RES_RHS; e += 256) {
    int c = e % RES_COLS, k = e / RES_COLS;
    xs[c][k] = (j0 + c < n && k < nrhs) ? x[j0 + c + k * ldx] : 0.0;
  }
  __syncthreads();
  if (i >= n) return;
  double s[RES_RHS];
#pragma unroll
  for (int k = 0; k < RES_RHS; ++k) s[k] = 0.0;
  if (nrhs == 1) {
    for (int64_t j = j0; j < j1; ++j) s[0] = __builtin_fma((double)a[i + j * lda], xs[j - j0][0], s[0]);
  } else {
    for (int64_t j = j0; j < j1; ++j) {
      double av = (double)a[i + j * lda];
#pragma unroll
      for (int k = 0; k < RES_RHS; ++k) s[k] = __builtin_fma(av, xs[j - j0][k], s[k]);
    }
  }
  for (int k = 0; k < nrhs; ++k) part[((int64_t)blockIdx.y * nrhs + k) * n + i] = s[k];
}

__global__ void __launch_bounds__(256) resid_reduce_kernel(int64_t n, int nrhs, int nchunk,
                                                           const double* __restrict__ part,
                                                           const double* __restrict__ b,
                                                           int64_t ldb, double* __restrict__ r,
                                                           int64_t ldr) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[((int64_t)c * nrhs + k) * n + i];
  r[i + k * ldr] = b[i + k * ldb] - s;
}

// ---------------------------------------------------------------- IR helpers
__global__ void d2f_kernel(int64_t n, int nrhs, const double* __restrict__ src, int64_t lds,
                           float* __restrict__ dst, int64_t ldd) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i + blockIdx.y * ldd] = (float)src[i + blockIdx.y * lds];
}
__global__ void axpy_f2d_kernel(int64_t n, const float* __restrict__ w, int64_t ldw,
                                double* __restrict__ x, int64_t ldx, int assign) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    double v = (double)w[i + blockIdx.y * ldw];
    double* p = x + i + blockIdx.y * ldx;
    *p = assign ? v : *p + v;
  }
}

// max_j sum_i |a_ij| (one-norm) and per-column inf norms, FP64
__global__ void colsum_kernel(int64_t n, const float* __restrict__ a, int64_t lda,
                              double* __restrict__ cs) {
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

// out[k] = max_i |v_ik| (NaN propagates); one workgroup per column
__global__ void colmax_kernel(int64_t n, const double* __restrict__ v, int64_t ldv,
                              double* __restrict__ out) {
  __shared__ double red[256];
  const int k = blockIdx.x;
  double m = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) {
    double x = fabs(v[i + k * ldv]);
    m = (x > m || x != x) ? x : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      double o = red[threadIdx.x + w];
      double s = red[threadIdx.x];
      red[threadIdx.x] = (o > s || o != o) ? o : s;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = red[0];
}

__global__ void maxreduce_kernel(int64_t n, const double* __restrict__ v, double* __restrict__ out) {
  __shared__ double red[256];
  double m = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) m = fmax(m, v[i]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

template <typename LT, bool LOWER, bool UNIT, bool TRANS>
int launch_trsm(remifa_handle_t h, int64_t n, int nrhs, const void* lu, int64_t ld, float* b,
                int64_t ldb, int* sync) {
  const int nblk = (int)ceil_div(n, NBS);
  fill_int_kernel<<<(unsigned)ceil_div(nblk + 2, 256), 256, 0, h->stream>>>(sync, nblk + 2, 0);
  trsm_kernel<LT, LOWER, UNIT, TRANS>
      <<<nblk, 256, 0, h->stream>>>(n, nrhs, (const LT*)lu, ld, b, ldb, sync, nblk);
  RF_HIP(hipGetLastError());
  return 0;
}

template <typename LT>
int getrs_typed(remifa_handle_t h, int trans, int64_t n, int64_t nrhs, const void* lu,
                int64_t ld, float* b, int64_t ldb, int* sync) {
  for (int64_t k0 = 0; k0 < nrhs; k0 += RC) {
    int nr = (int)std::min<int64_t>(RC, nrhs - k0);
    float* bk = b + k0 * ldb;
    if (!trans) {
      RF_CHECK((launch_trsm<LT, true, true, false>(h, n, nr, lu, ld, bk, ldb, sync)));
      RF_CHECK((launch_trsm<LT, false, false, false>(h, n, nr, lu, ld, bk, ldb, sync)));
    } else {
      // A^T = U^T L^T: U^T (lower, non-unit) then L^T (upper, unit)
      RF_CHECK((launch_trsm<LT, true, false, true>(h, n, nr, lu, ld, bk, ldb, sync)));
      RF_CHECK((launch_trsm<LT, false, true, true>(h, n, nr, lu, ld, bk, ldb, sync)));
    }
  }
  return 0;
}

}  // namespace

// sync words for one solve: ticket, error flag, one flag per block
static size_t trsm_sync_bytes(int64_t n) { return sizeof(int) * (size_t)(ceil_div(n, NBS) + 2); }

int getrs_launch(remifa_handle_t h, int lu_t, int trans, int64_t n, int64_t nrhs, const void* lu,
                 int64_t ld, float* b, int64_t ldb, int* sync) {
  if (n == 0 || nrhs == 0) return 0;
  timer_begin(h, KC_TRSM, 2.0 * (double)n * (double)n * (double)nrhs);
  int st = lu_t == REMIFA_R_16F ? getrs_typed<half_t>(h, trans, n, nrhs, lu, ld, b, ldb, sync)
                                : getrs_typed<float>(h, trans, n, nrhs, lu, ld, b, ldb, sync);
  timer_end(h);
  return st;
}

static size_t resid_part_bytes(int64_t n, int64_t nrhs) {
  return sizeof(double) * (size_t)ceil_div(n, RES_COLS) * (size_t)nrhs * (size_t)n;
}

int residual_launch(remifa_handle_t h, int64_t n, int64_t nrhs, const float* a, int64_t lda,
                    const double* x, int64_t ldx, const double* b, int64_t ldb, double* r,
                    int64_t ldr, double* part) {
  if (n == 0 || nrhs == 0) return 0;
  timer_begin(h, KC_RESID, 2.0 * (double)n * (double)n * (double)nrhs);
  for (int64_t k0 = 0; k0 < nrhs; k0 += RES_RHS) {
    int nr = (int)std::min<int64_t>(RES_RHS, nrhs - k0);
    int nchunk = (int)ceil_div(n, RES_COLS);
    dim3 g1((unsigned)ceil_div(n, RES_ROWS), (unsigned)nchunk);
    resid_partial_kernel<<<g1, 256, 0, h->stream>>>(n, nr, a, lda, x + k0 * ldx, ldx, part);
    dim3 g2((unsigned)ceil_div(n, 256), (unsigned)nr);
    resid_reduce_kernel<<<g2, 256, 0, h->stream>>>(n, nr, nchunk, part, b + k0 * ldb, ldb,
                                                   r + k0 * ldr, ldr);
  }
  timer_end(h);
  RF_HIP(hipGetLastError());
  return 0;
}

}  // namespace remifa

using namespace remifa;

extern "C" {

int remifa_getrs_nopiv(remifa_handle_t h, int lu_t, int trans, int64_t n, int64_t nrhs,
                       const void* d_lu, int64_t ldlu, float* d_b, int64_t ldb) {
  if (!h) return -1;
  if (lu_t != REMIFA_R_32F && lu_t != REMIFA_R_16F) return -2;
  if (n < 0) return -4;
  if (nrhs < 0) return -5;
  if (!d_lu && n > 0) return -6;
  if (ldlu < std::max<int64_t>(n, 1)) return -7;
  if (!d_b && n > 0 && nrhs > 0) return -8;
  if (ldb < std::max<int64_t>(n, 1)) return -9;
  if (n == 0 || nrhs == 0) return 0;
  RF_CHECK(ensure_workspace(h, trsm_sync_bytes(n)));
  int* sync = (int*)h->ws;
  RF_CHECK(getrs_launch(h, lu_t, trans, n, nrhs, d_lu, ldlu, d_b, ldb, sync));
  int err = 0;
  RF_HIP(hipMemcpyAsync(&err, sync + 1, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  RF_HIP(hipStreamSynchronize(h->stream));
  if (err) {
    set_last_error("triangular solve: dependency wait timed out");
    return REMIFA_ERR_HIP;
  }
  return 0;
}

int remifa_residual_f64(remifa_handle_t h, int64_t n, int64_t nrhs, const float* d_a, int64_t lda,
                        const double* d_x, int64_t ldx, const double* d_b, int64_t ldb,
                        double* d_r, int64_t ldr) {
  if (!h) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (!d_a && n > 0) return -4;
  if (lda < std::max<int64_t>(n, 1)) return -5;
  if (!d_x && n > 0 && nrhs > 0) return -6;
  if (ldx < std::max<int64_t>(n, 1)) return -7;
  if (!d_b && n > 0 && nrhs > 0) return -8;
  if (ldb < std::max<int64_t>(n, 1)) return -9;
  if (!d_r && n > 0 && nrhs > 0) return -10;
  if (ldr < std::max<int64_t>(n, 1)) return -11;
  if (n == 0 || nrhs == 0) return 0;
  RF_CHECK(ensure_workspace(h, resid_part_bytes(n, std::min<int64_t>(nrhs, RES_RHS))));
  return residual_launch(h, n, nrhs, d_a, lda, d_x, ldx, d_b, ldb, d_r, ldr, (double*)h->ws);
}

int remifa_ir_solve(remifa_handle_t h, int lu_t, int64_t n, int64_t nrhs, const float* d_a,
                    int64_t lda, const void* d_lu, int64_t ldlu, const double* d_b, int64_t ldb,
                    double* d_x, int64_t ldx, double tol, int maxit, int* iters, double* bwd_err,
                    double* history) {
  if (!h) return -1;
  if (lu_t != REMIFA_R_32F && lu_t != REMIFA_R_16F) return -2;
  if (n < 0) return -3;
  if (nrhs < 0) return -4;
  if (!d_a && n > 0) return -5;
  if (lda < std::max<int64_t>(n, 1)) return -6;
  if (!d_lu && n > 0) return -7;
  if (ldlu < std::max<int64_t>(n, 1)) return -8;
  if (!d_b && n > 0 && nrhs > 0) return -9;
  if (ldb < std::max<int64_t>(n, 1)) return -10;
  if (!d_x && n > 0 && nrhs > 0) return -11;
  if (ldx < std::max<int64_t>(n, 1)) return -12;
  if (maxit < 0) return -14;
  if (iters) *iters = -1;
  if (n == 0 || nrhs == 0) {
    if (iters) *iters = 0;
    if (bwd_err) *bwd_err = 0.0;
    return 0;
  }
  // workspace: sync words | fp32 correction (n x nrhs) | residual (n x nrhs)
  // | residual partials | column sums (n) | norms (2 nrhs + 1)
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t b_sync = al(trsm_sync_bytes(n));
  const size_t b_w = al(sizeof(float) * n * nrhs);
  const size_t b_r = al(sizeof(double) * n * nrhs);
  const size_t b_part = al(resid_part_bytes(n, std::min<int64_t>(nrhs, RES_RHS)));
  const size_t b_cs = al(sizeof(double) * n);
  const size_t b_nrm = al(sizeof(double) * (2 * nrhs + 2));
  RF_CHECK(ensure_workspace(h, b_sync + b_w + b_r + b_part + b_cs + b_nrm));
  char* p = (char*)h->ws;
  int* sync = (int*)p;
  float* w = (float*)(p + b_sync);
  double* r = (double*)(p + b_sync + b_w);
  double* part = (double*)(p + b_sync + b_w + b_r);
  double* cs = (double*)(p + b_sync + b_w + b_r + b_part);
  double* nrm = (double*)(p + b_sync + b_w + b_r + b_part + b_cs);
  hipStream_t s = h->stream;
  const dim3 gv((unsigned)ceil_div(n, 256), (unsigned)nrhs);

  // ||A||_1 and ||b||_inf per column (fixed for the whole solve)
  colsum_kernel<<<(unsigned)n, 256, 0, s>>>(n, d_a, lda, cs);
  maxreduce_kernel<<<1, 256, 0, s>>>(n, cs, nrm + 2 * nrhs);
  colmax_kernel<<<(unsigned)nrhs, 256, 0, s>>>(n, d_b, ldb, nrm + nrhs);
  // x0 = (LU)^-1 b in fp32
  d2f_kernel<<<gv, 256, 0, s>>>(n, (int)nrhs, d_b, ldb, w, n);
  RF_CHECK(getrs_launch(h, lu_t, 0, n, nrhs, d_lu, ldlu, w, n, sync));
  axpy_f2d_kernel<<<gv, 256, 0, s>>>(n, w, n, d_x, ldx, 1);
  RF_HIP(hipGetLastError());

  std::vector<double> hn(2 * nrhs + 2);
  int it = 0, done = -1;
  double worst = 0.0;
  for (;; ++it) {
    RF_CHECK(residual_launch(h, n, nrhs, d_a, lda, d_x, ldx, d_b, ldb, r, n, part));
    colmax_kernel<<<(unsigned)nrhs, 256, 0, s>>>(n, r, n, nrm);  // ||r||_inf
    colmax_kernel<<<(unsigned)nrhs, 256, 0, s>>>(n, d_x, ldx, part);  // ||x||_inf (part is free now)
    RF_HIP(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * (2 * nrhs + 1), hipMemcpyDeviceToHost, s));
    std::vector<double> xn(nrhs);
    RF_HIP(hipMemcpyAsync(xn.data(), part, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
    int serr = 0;
    RF_HIP(hipMemcpyAsync(&serr, sync + 1, sizeof(int), hipMemcpyDeviceToHost, s));
    RF_HIP(hipStreamSynchronize(s));
    if (serr) {
      set_last_error("triangular solve: dependency wait timed out");
      return REMIFA_ERR_HIP;
    }
    // tests/errors.hxx:237: ||r||_inf / (||A||_1 ||x||_inf + ||b||_inf)
    worst = 0.0;
    const double anorm = hn[2 * nrhs];
    for (int64_t k = 0; k < nrhs; ++k) {
      double e = hn[k] / (anorm * xn[k] + hn[nrhs + k]);
      if (e != e) {
        worst = e;
        break;
      }
      worst = std::max(worst, e);
    }
    if (history) history[it] = worst;
    if (worst <= tol) {
      done = it;
      break;
    }
    if (it >= maxit || worst != worst) break;
    d2f_kernel<<<gv, 256, 0, s>>>(n, (int)nrhs, r, n, w, n);
    RF_CHECK(getrs_launch(h, lu_t, 0, n, nrhs, d_lu, ldlu, w, n, sync));
    axpy_f2d_kernel<<<gv, 256, 0, s>>>(n, w, n, d_x, ldx, 0);
    RF_HIP(hipGetLastError());
  }
  if (iters) *iters = done;
  if (bwd_err) *bwd_err = worst;
  return 0;
}

}  // extern "C"
