// Triangular solves with the stored no-pivot factors, the FP64 residual and
// iterative refinement.
//
// TRSM (forward L, backward U and the transposed pair for A^T): the 128 x 128
// diagonal blocks are inverted first, all in parallel (diag_inv_kernel), so
// the sequential chain of the substitution is one small matrix-vector product
// per block row instead of 128 dependent steps. Then one launch per triangle:
// workgroups take 128-row block rows in dependency order from an atomic
// ticket, so every block a workgroup waits for belongs to a workgroup that is
// already running (no deadlock whatever the residency). Block row i
// accumulates B_i - sum_j T_ij X_j, multiplies by its inverted diagonal block
// (the last dependency is folded into it ahead of time: X_i = (T_ii)^-1
// (B_i - sum of the others) - ((T_ii)^-1 T_i,i-1) X_i-1, so one product and
// one reduction follow the last wait instead of two of each) and publishes X_i as tagged 64-bit words (value, 1) written with
// agent-scope atomic stores into a tag buffer cleared before the launch: a
// reader retries until every word it needs carries the tag, so a value and
// its "published" mark arrive together (one memory round trip per hop
// instead of store, wait, counter, poll, load). Every wait is bounded: on
// timeout the kernel records an error and exits, so the grid always drains.
// The factors are read once (HBM-bound).
//
// Residual R = B - A X in FP64: 2-D grid over (256-row block, 512-column
// chunk), fp64 fma per element, per-chunk partials summed in chunk order by a
// second kernel (deterministic).
#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"

namespace remifa {
namespace {

constexpr int NBS = 128;     // block size of the triangular solves
constexpr int TT = 4 * NBS;  // threads of a triangular-solve workgroup
constexpr int QC = NBS / 4;  // columns of a block per lane (column quarter)
constexpr int RC = 16;   // right-hand sides per launch
constexpr long long kSpinLimit = 1LL << 22;  // retries of a wait (each a memory round trip)

template <typename LT>
__device__ __forceinline__ float ldf(const LT* p) {
  return (float)*p;
}

// element T[r][c] of the effective triangle
template <typename LT, bool TRANS>
__device__ __forceinline__ float tel(const LT* lu, int64_t ld, int64_t r, int64_t c) {
  return TRANS ? ldf(lu + c + r * ld) : ldf(lu + r + c * ld);
}

// Inverses of the diagonal blocks: blockIdx.y = 0 for the unit lower L_bb,
// 1 for the upper U_bb. inv[b][r][c] (128 x 128 row-major per block); rows
// and columns past n are the identity. The block is split 2 x 2 into 64 x 64
// quarters: lanes 0-63 invert the first diagonal quarter and lanes 64-127 the
// second, lane c computing column c by substitution with the block in LDS;
// the off-diagonal quarter of the inverse is then -D22^-1 D21 D11^-1 (lower)
// or -D11^-1 D12 D22^-1 (upper), two 64 x 64 products in LDS.
template <typename LT>
__global__ void __launch_bounds__(NBS) diag_inv_kernel(int64_t n, const LT* __restrict__ lu,
                                                      int64_t ld, float* __restrict__ inv_l,
                                                      float* __restrict__ inv_u) {
  constexpr int H = NBS / 2;
  __shared__ float D[NBS][NBS + 1];
  __shared__ float V[NBS][NBS + 1];  // the inverse
  const int t = threadIdx.x;
  const int64_t r0 = (int64_t)blockIdx.x * NBS;
  const bool upper = blockIdx.y == 1;
  for (int e = t; e < NBS * NBS; e += NBS) {
    const int r = e % NBS, c = e / NBS;
    const int64_t R = r0 + r, C = r0 + c;
    D[r][c] = (R < n && C < n) ? ldf(lu + R + C * ld) : (r == c ? 1.0f : 0.0f);
  }
  __syncthreads();
  // diagonal quarters (o is uniform over a wavefront: D reads broadcast)
  const int o = (t / H) * H, c = t % H;
  {
    float x[H];
    if (!upper) {
#pragma unroll
      for (int r = 0; r < H; ++r) {
        float s = (r == c) ? 1.0f : 0.0f;
#pragma unroll
        for (int k = 0; k < r; ++k) s = __builtin_fmaf(-D[o + r][o + k], x[k], s);
        x[r] = s;
      }
    } else {
#pragma unroll
      for (int r = H - 1; r >= 0; --r) {
        float s = (r == c) ? 1.0f : 0.0f;
#pragma unroll
        for (int k = r + 1; k < H; ++k) s = __builtin_fmaf(-D[o + r][o + k], x[k], s);
        x[r] = s / D[o + r][o + r];
      }
    }
#pragma unroll
    for (int r = 0; r < H; ++r) {
      V[o + r][o + c] = x[r];
      V[H - o + r][o + c] = 0.0f;  // the zero off-diagonal quarter (overwritten below if needed)
    }
  }
  __syncthreads();
  // lane (half i0, column j): 32 rows of a 64 x 64 product
  const int j = t % H, i0 = (t / H) * (H / 2);
  float acc[H / 2];
  // step 1: P = D21 V11 (lower) or D12 V22 (upper)
#pragma unroll
  for (int i = 0; i < H / 2; ++i) acc[i] = 0.0f;
  for (int k = 0; k < H; ++k) {
    const float vk = upper ? V[H + k][H + j] : V[k][j];
#pragma unroll
    for (int i = 0; i < H / 2; ++i)
      acc[i] = __builtin_fmaf(upper ? D[i0 + i][H + k] : D[H + i0 + i][k], vk, acc[i]);
  }
  __syncthreads();
  // P goes to the D quarter it was computed from (no longer needed)
#pragma unroll
  for (int i = 0; i < H / 2; ++i) {
    if (upper)
      D[i0 + i][H + j] = acc[i];
    else
      D[H + i0 + i][j] = acc[i];
  }
  __syncthreads();
  // step 2: V21 = -V22 P (lower) or V12 = -V11 P (upper)
#pragma unroll
  for (int i = 0; i < H / 2; ++i) acc[i] = 0.0f;
  for (int k = 0; k < H; ++k) {
    const float pk = upper ? D[k][H + j] : D[H + k][j];
#pragma unroll
    for (int i = 0; i < H / 2; ++i)
      acc[i] = __builtin_fmaf(upper ? V[i0 + i][k] : V[H + i0 + i][H + k], pk, acc[i]);
  }
#pragma unroll
  for (int i = 0; i < H / 2; ++i) {
    if (upper)
      V[i0 + i][H + j] = -acc[i];
    else
      V[H + i0 + i][j] = -acc[i];
  }
  __syncthreads();
  float* out = (upper ? inv_u : inv_l) + (int64_t)blockIdx.x * NBS * NBS;
  for (int r = 0; r < NBS; ++r) out[r * NBS + t] = V[r][t];
}

// One block row of T X = B per workgroup (see the file comment). Lane
// (row, cq) owns row `row` of the block and the column quarter cq (32
// columns) of every 128 x 128 block it multiplies; the four partial sums are
// added in LDS in a fixed order. NR: right-hand sides per launch (1 or 16).
// sync[0] = ticket, sync[1] = error flag; tags[r + k n] = (value, 1) once
// entry r of solution column k is published (zeroed before the launch).
template <typename LT, bool LOWER, bool TRANS, int NR>
__global__ void __launch_bounds__(TT) trsm_kernel(int64_t n, int nrhs, const LT* __restrict__ lu,
                                                   int64_t ld, float* __restrict__ b, int64_t ldb,
                                                   const float* __restrict__ inv,
                                                   int* __restrict__ sync,
                                                   unsigned long long* __restrict__ tags,
                                                   int nblk) {
  constexpr int CH = NR == 1 ? 2 : 1;  // solution blocks staged in LDS at a time
  constexpr int NP = NR == 1 ? 1 : NR + 1;
  // Xs and Red share their LDS with the near block of T, which is only read
  // before the first solution block is staged (a barrier between)
  constexpr int XS_N = CH * NBS * NR, RED_N = 4 * NBS * NP, TS_N = NBS * NBS;
  constexpr int U_N = XS_N + RED_N > TS_N ? XS_N + RED_N : TS_N;
  __shared__ __attribute__((aligned(16))) float ush[U_N];
  float(*const Xs)[NR] = reinterpret_cast<float(*)[NR]>(ush);
  float(*const Red)[NBS][NP] = reinterpret_cast<float(*)[NBS][NP]>(ush + XS_N);
  float(*const Ts)[NBS] = reinterpret_cast<float(*)[NBS]>(ush);
  __shared__ int s_ticket;
  const int t = threadIdx.x;
  const int row = t % NBS, cq = t / NBS;
  if (t == 0)
    s_ticket = __hip_atomic_fetch_add(&sync[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  const int tk = s_ticket;
  const int bi = LOWER ? tk : nblk - 1 - tk;
  const int64_t gr = (int64_t)bi * NBS + row;

  // this lane's entries of the inverted diagonal block (transposed for the
  // transposed triangles: (T_ii)^-1 = ((D_ii)^-1)^T)
  float dv[QC];
  {
    const float* ib = inv + (int64_t)bi * NBS * NBS;
#pragma unroll
    for (int c = 0; c < QC; ++c) {
      const int col = cq * QC + c;
      dv[c] = TRANS ? ib[col * NBS + row] : ib[row * NBS + col];
    }
  }
  float acc[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) acc[k] = 0.0f;

  // The last dependency (block tk-1 in dependency order, the "near" block)
  // is folded into the diagonal block's inverse ahead of time, M =
  // (T_ii)^-1 T_i,near, so that once X_near is published only X_i = C - M
  // X_near remains (one product and one reduction), with C = (T_ii)^-1 (B_i
  // - sum over the other blocks) formed while waiting.
  constexpr bool FOLD = true;
  const bool fold = FOLD && tk > 0;
  const int nsteps = fold ? tk - 1 : tk;  // blocks summed before the wait for the near one
  // rows of the stored factors start on 16 bytes (16-byte vector reads of
  // the transposed triangle's rows)
  const bool vec_rows =
      TRANS && ((uintptr_t)lu % 16 == 0) && (ld % (16 / (int64_t)sizeof(LT)) == 0);
  const int64_t grc = gr < n ? gr : n - 1;  // in-bounds row for the loads
  // one right-hand side: this lane's entry of b (lanes cq = 0), loaded now
  // so that no memory round trip follows the last dependency wait (only
  // this workgroup writes these rows of b). With 16 right-hand sides the
  // extra registers would spill; those are loaded after the waits.
  const float bpf = NR == 1 ? b[grc] : 0.0f;
  float mv[QC];
  // (T_ii)^-1 in LDS (rows padded against bank conflicts) for M; with one
  // right-hand side it stays there for the diagonal product, which frees
  // the lane's registers for the loop
  __shared__ float Ds[NBS][NBS + 1];
  if constexpr (FOLD) {
    if (NR == 1 || fold) {
#pragma unroll
      for (int c = 0; c < QC; ++c) Ds[row][cq * QC + c] = dv[c];
    }
    if (fold) {
      // M: this lane's row and column quarter, from (T_ii)^-1 and T_i,near
      // in LDS
      const int bjn = LOWER ? tk - 1 : nblk - 1 - (tk - 1);
      const int64_t c0 = (int64_t)bjn * NBS + cq * QC;
#pragma unroll
      for (int c = 0; c < QC; ++c) {
        const int64_t C = c0 + c < n ? c0 + c : n - 1;
        const float v = tel<LT, TRANS>(lu, ld, grc, C);
        Ts[row][cq * QC + c] = (gr < n && c0 + c < n) ? v : 0.0f;
      }
      __syncthreads();
#pragma unroll
      for (int c = 0; c < QC; ++c) mv[c] = 0.0f;
#pragma unroll 2
      for (int k = 0; k < NBS; ++k) {
        const float d = Ds[row][k];
#pragma unroll
        for (int c = 0; c < QC; ++c) mv[c] = __builtin_fmaf(d, Ts[k][cq * QC + c], mv[c]);
      }
      __syncthreads();  // (Ts is Xs and Red from here on)
    }
  }
  // the solution blocks s0 .. s0 + cnt - 1 (dependency order) into Xs:
  // agent-scope atomic loads of the tagged words, each wavefront retrying
  // its own until every one carries the published tag (a wavefront vote),
  // then one barrier; bounded, and a wavefront that gives up raises the
  // error flag that ends the others' waits
  auto stage_blocks = [&](int s0, int cnt) {
    for (long long tries = 0;; ++tries) {
      int bad = 0;
      for (int e = t; e < cnt * NBS * NR; e += TT) {
        const int rr = e % (cnt * NBS), k = e / (cnt * NBS);
        const int q = rr / NBS, c = rr % NBS;
        const int bj = LOWER ? s0 + q : nblk - 1 - (s0 + q);
        const int64_t C = (int64_t)bj * NBS + c;
        float v = 0.0f;
        if (C < n && k < nrhs) {
          const unsigned long long w = __hip_atomic_load(&tags[C + (int64_t)k * n],
                                                         __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
          bad |= (unsigned)(w >> 32) != 1u;
          v = __uint_as_float((unsigned)w);
        }
        Xs[rr][k] = v;
      }
      if (__ballot(bad) == 0) break;
      if (tries > kSpinLimit ||
          ((tries & 63) == 63 &&
           __hip_atomic_load(&sync[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
        // give up (the host reports the error and discards the results)
        if ((t & 63) == 0)
          __hip_atomic_store(&sync[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __syncthreads();
  };
  for (int s0 = 0; s0 < nsteps; s0 += CH) {
    const int cnt = nsteps - s0 < CH ? nsteps - s0 : CH;
    // this lane's entries of the chunk's blocks of T: issued before the wait
    // (they do not depend on it), with clamped addresses and no branches so
    // the loads are all in flight together
    float tv[CH][QC];
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      const int qq = q < cnt ? q : cnt - 1;
      const int bj = LOWER ? s0 + qq : nblk - 1 - (s0 + qq);
      const int64_t c0 = (int64_t)bj * NBS + cq * QC;
      if (TRANS && vec_rows && gr < n && c0 + QC <= n) {
        // the transposed triangle: the lane's 32 entries are contiguous in
        // its row of the stored factors, read as 16-byte vectors
        const LT* pr = lu + c0 + grc * ld;
        if constexpr (sizeof(LT) == 2) {
          typedef _Float16 h8 __attribute__((ext_vector_type(8)));
#pragma unroll
          for (int u = 0; u < QC / 8; ++u) {
            const h8 v = *(const h8*)(pr + 8 * u);
#pragma unroll
            for (int j = 0; j < 8; ++j) tv[q][8 * u + j] = (float)v[j];
          }
        } else {
#pragma unroll
          for (int u = 0; u < QC / 4; ++u) {
            const float4 v = *(const float4*)(pr + 4 * u);
            tv[q][4 * u] = v.x, tv[q][4 * u + 1] = v.y, tv[q][4 * u + 2] = v.z,
                      tv[q][4 * u + 3] = v.w;
          }
        }
        continue;
      }
#pragma unroll
      for (int c = 0; c < QC; ++c) {
        const int64_t C = c0 + c < n ? c0 + c : n - 1;
        const float v = tel<LT, TRANS>(lu, ld, grc, C);
        tv[q][c] = (gr < n && c0 + c < n) ? v : 0.0f;
      }
    }
    stage_blocks(s0, cnt);
#pragma unroll
    for (int q = 0; q < CH; ++q) {
      if (q < cnt) {
#pragma unroll
        for (int c = 0; c < QC; ++c) {
#pragma unroll
          for (int k = 0; k < NR; ++k)
            acc[k] = __builtin_fmaf(tv[q][c], Xs[q * NBS + cq * QC + c][k], acc[k]);
        }
      }
    }
    __syncthreads();
  }

  // right-hand side of the diagonal block: B_i - sum_j T_ij X_j (over the
  // blocks summed so far)
#pragma unroll
  for (int k = 0; k < NR; ++k) Red[cq][row][k] = acc[k];
  __syncthreads();
  for (int k = cq; k < NR; k += 4) {
    const float sum = ((Red[0][row][k] + Red[1][row][k]) + Red[2][row][k]) + Red[3][row][k];
    const float bv = NR == 1 ? bpf : b[grc + (k < nrhs ? k : 0) * ldb];
    Xs[row][k] = (gr < n && k < nrhs) ? bv - sum : 0.0f;
  }
  __syncthreads();
  // X_i = (T_ii)^-1 (B_i - ...)  (with the fold: C)
  float y[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) y[k] = 0.0f;
#pragma unroll
  for (int c = 0; c < QC; ++c) {
    const float d = NR == 1 ? Ds[row][cq * QC + c] : dv[c];
#pragma unroll
    for (int k = 0; k < NR; ++k) y[k] = __builtin_fmaf(d, Xs[cq * QC + c][k], y[k]);
  }
#pragma unroll
  for (int k = 0; k < NR; ++k) Red[cq][row][k] = y[k];
  __syncthreads();
  if constexpr (FOLD) {
    if (fold) {
      // C (right-hand sides k = cq + 4 j in lane (row, cq)), then the near
      // block: X_i = C - M X_near
      constexpr int NJ = (NR + 3) / 4;
      float cv[NJ];
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int k = cq + 4 * j;
        cv[j] = k < NR ? ((Red[0][row][k] + Red[1][row][k]) + Red[2][row][k]) + Red[3][row][k]
                       : 0.0f;
      }
      __syncthreads();  // (Red and Xs are rewritten below)
      stage_blocks(tk - 1, 1);
      float z[NR];
#pragma unroll
      for (int k = 0; k < NR; ++k) z[k] = 0.0f;
#pragma unroll
      for (int c = 0; c < QC; ++c) {
#pragma unroll
        for (int k = 0; k < NR; ++k) z[k] = __builtin_fmaf(mv[c], Xs[cq * QC + c][k], z[k]);
      }
#pragma unroll
      for (int k = 0; k < NR; ++k) Red[cq][row][k] = z[k];
      __syncthreads();
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const int k = cq + 4 * j;
        if (k < NR && k < nrhs && gr < n) {
          const float xv =
              cv[j] - (((Red[0][row][k] + Red[1][row][k]) + Red[2][row][k]) + Red[3][row][k]);
          b[gr + k * ldb] = xv;  // the caller's copy
          __hip_atomic_store(&tags[gr + (int64_t)k * n],
                             (1ull << 32) | (unsigned long long)__float_as_uint(xv),
                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      return;
    }
  }
  for (int k = cq; k < NR; k += 4) {
    const float xv = ((Red[0][row][k] + Red[1][row][k]) + Red[2][row][k]) + Red[3][row][k];
    if (gr < n && k < nrhs) {
      b[gr + k * ldb] = xv;  // the caller's copy
      __hip_atomic_store(&tags[gr + (int64_t)k * n],
                         (1ull << 32) | (unsigned long long)__float_as_uint(xv), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

__global__ void fill_int_kernel(int* p, int n, int v) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------- residual
constexpr int RES_ROWS = 256;  // rows per workgroup (one per lane)
constexpr int RES_COLS = 512;  // columns per chunk
constexpr int RES_RHS = 16;    // right-hand sides per launch

// part[chunk][rhs][n] = sum over the chunk's columns j of a_ij x_jk, in FP64.
// NR = 1: four interleaved accumulators (j mod 4) combined in a fixed order,
// so several loads are in flight per lane; NR = 16: one accumulator per
// right-hand side (16 FMAs per element of A).
template <int NR>
__global__ void __launch_bounds__(256) resid_partial_kernel(int64_t n, int nrhs,
                                                            const float* __restrict__ a,
                                                            int64_t lda,
                                                            const double* __restrict__ x,
                                                            int64_t ldx, double* __restrict__ part) {
  __shared__ double xs[RES_COLS][NR];
  const int64_t i = (int64_t)blockIdx.x * RES_ROWS + threadIdx.x;
  const int64_t j0 = (int64_t)blockIdx.y * RES_COLS;
  for (int e = threadIdx.x; e < RES_COLS * NR; e += 256) {
    int c = e % RES_COLS, k = e / RES_COLS;
    xs[c][k] = (j0 + c < n && k < nrhs) ? x[j0 + c + k * ldx] : 0.0;
  }
  __syncthreads();
  if (i >= n) return;
  const int jn = (int)(n - j0 < RES_COLS ? n - j0 : RES_COLS);
  const float* ai = a + i + j0 * lda;
  if constexpr (NR == 1) {
    double s[4] = {0.0, 0.0, 0.0, 0.0};
    int j = 0;
    for (; j + 8 <= jn; j += 8) {
      float av[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) av[q] = ai[(int64_t)(j + q) * lda];
#pragma unroll
      for (int q = 0; q < 8; ++q) s[q & 3] = __builtin_fma((double)av[q], xs[j + q][0], s[q & 3]);
    }
    for (; j < jn; ++j) s[j & 3] = __builtin_fma((double)ai[(int64_t)j * lda], xs[j][0], s[j & 3]);
    part[(int64_t)blockIdx.y * n + i] = (s[0] + s[1]) + (s[2] + s[3]);
  } else {
    double s[NR];
#pragma unroll
    for (int k = 0; k < NR; ++k) s[k] = 0.0;
    int j = 0;
    for (; j + 4 <= jn; j += 4) {
      float av[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) av[q] = ai[(int64_t)(j + q) * lda];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int k = 0; k < NR; ++k) s[k] = __builtin_fma((double)av[q], xs[j + q][k], s[k]);
    }
    for (; j < jn; ++j) {
      const double av = (double)ai[(int64_t)j * lda];
#pragma unroll
      for (int k = 0; k < NR; ++k) s[k] = __builtin_fma(av, xs[j][k], s[k]);
    }
    for (int k = 0; k < nrhs; ++k) part[((int64_t)blockIdx.y * nrhs + k) * n + i] = s[k];
  }
}

// part[chunk][rhs][n] for R = B - A^T X: entry i is column i of A dotted with
// x over the chunk's rows. 16 lanes per column (64-byte contiguous reads of
// the column), 16 columns per workgroup; the 16 lane sums of a column are
// added in lane order (deterministic).
template <int NR>
__global__ void __launch_bounds__(256) resid_t_partial_kernel(int64_t n, int nrhs,
                                                              const float* __restrict__ a,
                                                              int64_t lda,
                                                              const double* __restrict__ x,
                                                              int64_t ldx,
                                                              double* __restrict__ part) {
  constexpr int ROWS = NR == 1 ? 2048 : 256;  // rows per chunk
  __shared__ double xs[ROWS][NR];
  __shared__ double red[16][16][NR];
  const int t = threadIdx.x;
  const int c = t / 16, q = t % 16;
  const int64_t i = (int64_t)blockIdx.x * 16 + c;  // column of A = entry of r
  const int64_t j0 = (int64_t)blockIdx.y * ROWS;
  for (int e = t; e < ROWS * NR; e += 256) {
    const int jj = e % ROWS, k = e / ROWS;
    xs[jj][k] = (j0 + jj < n && k < nrhs) ? x[j0 + jj + k * ldx] : 0.0;
  }
  __syncthreads();
  double s[NR];
#pragma unroll
  for (int k = 0; k < NR; ++k) s[k] = 0.0;
  if (i < n) {
    const int jn = (int)(n - j0 < ROWS ? n - j0 : ROWS);
    const float* ac = a + j0 + i * lda;
    for (int jj = q; jj < jn; jj += 16) {
      const double av = (double)ac[jj];
#pragma unroll
      for (int k = 0; k < NR; ++k) s[k] = __builtin_fma(av, xs[jj][k], s[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < NR; ++k) red[c][q][k] = s[k];
  __syncthreads();
  if (q == 0 && i < n) {
    for (int k = 0; k < nrhs; ++k) {
      double v = 0.0;
      for (int r = 0; r < 16; ++r) v += red[c][r][k];
      part[((int64_t)blockIdx.y * nrhs + k) * n + i] = v;
    }
  }
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

// rs[i] = sum_j |a_ij| (one-norm of A^T), one lane per row
__global__ void rowsum_kernel(int64_t n, const float* __restrict__ a, int64_t lda,
                              double* __restrict__ rs) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int64_t j = 0; j < n; ++j) s += fabs((double)a[i + j * lda]);
  rs[i] = s;
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

__global__ void zero_u64_kernel(unsigned long long* p, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

template <typename LT, bool LOWER, bool TRANS>
int launch_trsm(remifa_handle_t h, int64_t n, int nrhs, const void* lu, int64_t ld, float* b,
                int64_t ldb, const float* inv, int* sync, unsigned long long* tags) {
  const int nblk = (int)ceil_div(n, NBS);
  fill_int_kernel<<<1, 256, 0, h->stream>>>(sync, 1, 0);  // the ticket
  const int64_t nt = n * (int64_t)nrhs;
  zero_u64_kernel<<<(unsigned)ceil_div(nt, 256), 256, 0, h->stream>>>(tags, nt);
  if (nrhs == 1)
    trsm_kernel<LT, LOWER, TRANS, 1>
        <<<nblk, TT, 0, h->stream>>>(n, nrhs, (const LT*)lu, ld, b, ldb, inv, sync, tags, nblk);
  else
    trsm_kernel<LT, LOWER, TRANS, RC>
        <<<nblk, TT, 0, h->stream>>>(n, nrhs, (const LT*)lu, ld, b, ldb, inv, sync, tags, nblk);
  RF_HIP(hipGetLastError());
  return 0;
}

template <typename LT>
int getrs_typed(remifa_handle_t h, int trans, int64_t n, int64_t nrhs, const void* lu,
                int64_t ld, float* b, int64_t ldb, const float* inv_l, const float* inv_u,
                int* sync, unsigned long long* tags) {
  fill_int_kernel<<<1, 256, 0, h->stream>>>(sync, 2, 0);  // ticket and error flag
  for (int64_t k0 = 0; k0 < nrhs; k0 += RC) {
    int nr = (int)std::min<int64_t>(RC, nrhs - k0);
    float* bk = b + k0 * ldb;
    if (!trans) {
      RF_CHECK((launch_trsm<LT, true, false>(h, n, nr, lu, ld, bk, ldb, inv_l, sync, tags)));
      RF_CHECK((launch_trsm<LT, false, false>(h, n, nr, lu, ld, bk, ldb, inv_u, sync, tags)));
    } else {
      // A^T = U^T L^T: U^T (lower) then L^T (upper, unit)
      RF_CHECK((launch_trsm<LT, true, true>(h, n, nr, lu, ld, bk, ldb, inv_u, sync, tags)));
      RF_CHECK((launch_trsm<LT, false, true>(h, n, nr, lu, ld, bk, ldb, inv_l, sync, tags)));
    }
  }
  return 0;
}

}  // namespace

// solve workspace: sync words (ticket, error flag, published count), then
// the inverted diagonal blocks of L and of U
struct TrsmWs {
  int* sync;
  float* inv_l;
  float* inv_u;
  unsigned long long* tags;  // n x RC published-solution words
};

static size_t trsm_ws_bytes(int64_t n) {
  return 256 + 2 * sizeof(float) * NBS * NBS * (size_t)ceil_div(n, NBS) +
         sizeof(unsigned long long) * (size_t)n * RC;
}

static TrsmWs trsm_ws(void* p, int64_t n) {
  char* c = (char*)p;
  const size_t inv = sizeof(float) * NBS * NBS * (size_t)ceil_div(n, NBS);
  return TrsmWs{(int*)c, (float*)(c + 256), (float*)(c + 256 + inv),
                (unsigned long long*)(c + 256 + 2 * inv)};
}

// invert the diagonal blocks of the stored factors (once per factorization)
static int trsm_invert_launch(remifa_handle_t h, int lu_t, int64_t n, const void* lu, int64_t ld,
                              const TrsmWs& w) {
  const int nblk = (int)ceil_div(n, NBS);
  timer_begin(h, KC_TRSM, (double)n * NBS * (lu_t == REMIFA_R_16F ? 2 : 4) * 2, 1);
  if (lu_t == REMIFA_R_16F)
    diag_inv_kernel<half_t><<<dim3(nblk, 2), NBS, 0, h->stream>>>(n, (const half_t*)lu, ld,
                                                                  w.inv_l, w.inv_u);
  else
    diag_inv_kernel<float><<<dim3(nblk, 2), NBS, 0, h->stream>>>(n, (const float*)lu, ld, w.inv_l,
                                                                 w.inv_u);
  RF_HIP(hipGetLastError());
  timer_end(h);
  return 0;
}

// (L U) X = B or (L U)^T X = B in place in b; the diagonal blocks must have
// been inverted into w
int getrs_launch(remifa_handle_t h, int lu_t, int trans, int64_t n, int64_t nrhs, const void* lu,
                 int64_t ld, float* b, int64_t ldb, const TrsmWs& w) {
  if (n == 0 || nrhs == 0) return 0;
  // bytes: both triangles read once per 16 right-hand sides
  const double es = lu_t == REMIFA_R_16F ? 2.0 : 4.0;
  timer_begin(h, KC_TRSM, (double)n * (double)n * es * (double)ceil_div(nrhs, RC), 0);
  int st = lu_t == REMIFA_R_16F
               ? getrs_typed<half_t>(h, trans, n, nrhs, lu, ld, b, ldb, w.inv_l, w.inv_u, w.sync,
                                     w.tags)
               : getrs_typed<float>(h, trans, n, nrhs, lu, ld, b, ldb, w.inv_l, w.inv_u, w.sync,
                                    w.tags);
  timer_end(h);
  return st;
}

// chunks of the residual reduction (rows of A for R = B - A X, rows of A
// read down the columns for R = B - A^T X)
static int64_t resid_chunks(int64_t n, int trans, int nr) {
  return trans ? ceil_div(n, nr == 1 ? 2048 : 256) : ceil_div(n, RES_COLS);
}

static size_t resid_part_bytes(int64_t n, int64_t nrhs) {
  // the largest of the two layouts
  const int64_t ch = std::max(resid_chunks(n, 0, (int)nrhs), resid_chunks(n, 1, (int)nrhs));
  return sizeof(double) * (size_t)ch * (size_t)nrhs * (size_t)n;
}

int residual_launch(remifa_handle_t h, int64_t n, int64_t nrhs, const float* a, int64_t lda,
                    const double* x, int64_t ldx, const double* b, int64_t ldb, double* r,
                    int64_t ldr, double* part, int trans = 0) {
  if (n == 0 || nrhs == 0) return 0;
  // bytes: A once per 16 right-hand sides
  timer_begin(h, KC_RESID, 4.0 * (double)n * (double)n * (double)ceil_div(nrhs, RES_RHS), trans);
  for (int64_t k0 = 0; k0 < nrhs; k0 += RES_RHS) {
    int nr = (int)std::min<int64_t>(RES_RHS, nrhs - k0);
    int nchunk = (int)resid_chunks(n, trans, nr);
    if (trans) {
      dim3 g1((unsigned)ceil_div(n, 16), (unsigned)nchunk);
      if (nr == 1)
        resid_t_partial_kernel<1><<<g1, 256, 0, h->stream>>>(n, nr, a, lda, x + k0 * ldx, ldx,
                                                             part);
      else
        resid_t_partial_kernel<RES_RHS><<<g1, 256, 0, h->stream>>>(n, nr, a, lda, x + k0 * ldx,
                                                                   ldx, part);
    } else {
      dim3 g1((unsigned)ceil_div(n, RES_ROWS), (unsigned)nchunk);
      if (nr == 1)
        resid_partial_kernel<1><<<g1, 256, 0, h->stream>>>(n, nr, a, lda, x + k0 * ldx, ldx, part);
      else
        resid_partial_kernel<RES_RHS><<<g1, 256, 0, h->stream>>>(n, nr, a, lda, x + k0 * ldx, ldx,
                                                                 part);
    }
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
  RF_CHECK(ensure_workspace(h, trsm_ws_bytes(n)));
  const TrsmWs w = trsm_ws(h->ws, n);
  RF_CHECK(trsm_invert_launch(h, lu_t, n, d_lu, ldlu, w));
  RF_CHECK(getrs_launch(h, lu_t, trans, n, nrhs, d_lu, ldlu, d_b, ldb, w));
  int err = 0;
  RF_HIP(hipMemcpyAsync(&err, w.sync + 1, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  RF_HIP(hipStreamSynchronize(h->stream));
  if (err) {
    set_last_error("triangular solve: dependency wait timed out");
    return REMIFA_ERR_HIP;
  }
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------- no-allocation solve
// Scratch of remifa_getrs_nopiv_noalloc: a file-scope device array sized for
// the largest order it takes (sync words, the inverted 128 x 128 diagonal
// blocks of L and U, the n x 16 tagged solution words: trsm_ws_bytes(n)).
// Every GPU holds its own copy for the life of the process. Host code never
// uses the array's name as an address: its device address comes from
// hipGetSymbolAddress, once per device (noalloc_scratch).
namespace remifa {
namespace {
constexpr int64_t kNoAllocMaxN = 32768;
constexpr size_t kNoAllocBytes = 256 + 2 * sizeof(float) * NBS * NBS * (kNoAllocMaxN / NBS) +
                                 sizeof(unsigned long long) * kNoAllocMaxN * RC;
static_assert(kNoAllocMaxN % NBS == 0, "whole diagonal blocks");
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// device address of this device's scratch array (looked up on the first call
// for the device, which must be current)
int noalloc_scratch(int device, void** ptr) {
  *ptr = nullptr;
  if (device < 0 || device >= kMaxDevices) {
    set_last_error("device index out of range of the scratch table");
    return REMIFA_ERR_UNSUPPORTED;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    RF_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch[device] = p;
  }
  *ptr = g_scratch[device];
  return 0;
}

size_t noalloc_scratch_bytes() { return kNoAllocBytes; }
}  // namespace remifa

extern "C" {

int remifa_getrs_nopiv_noalloc(remifa_handle_t h, int lu_t, int trans, int64_t n, int64_t nrhs,
                               const void* d_lu, int64_t ldlu, float* d_b, int64_t ldb) {
  if (!h) return -1;
  if (lu_t != REMIFA_R_32F && lu_t != REMIFA_R_16F) return -2;
  if (n < 0) return -4;
  if (nrhs < 0) return -5;
  if (!d_lu && n > 0) return -6;
  if (ldlu < std::max<int64_t>(n, 1)) return -7;
  if (!d_b && n > 0 && nrhs > 0) return -8;
  if (ldb < std::max<int64_t>(n, 1)) return -9;
  if (n > kNoAllocMaxN) {
    set_last_error("remifa_getrs_nopiv_noalloc: n = " + std::to_string(n) +
                   " exceeds the largest order of its static scratch, " +
                   std::to_string(kNoAllocMaxN));
    return REMIFA_ERR_WORKSPACE;
  }
  if (n == 0 || nrhs == 0) return 0;
  if (!h->scratch) {
    set_last_error("remifa_getrs_nopiv_noalloc: no scratch address for this handle's device");
    return REMIFA_ERR_HIP;
  }
  // the kernels of remifa_getrs_nopiv on the static scratch; the timer is
  // held off (its event pool would grow), the sync words and solution tags
  // are cleared by fill kernels inside getrs_launch, the inverted blocks
  // are rewritten, so nothing of an earlier call's contents is read
  const bool timed = h->timer.enabled;
  h->timer.enabled = false;
  const TrsmWs w = trsm_ws(h->scratch, n);
  int st = trsm_invert_launch(h, lu_t, n, d_lu, ldlu, w);
  if (!st) st = getrs_launch(h, lu_t, trans, n, nrhs, d_lu, ldlu, d_b, ldb, w);
  h->timer.enabled = timed;
  RF_CHECK(st);
  int err = 0;
  RF_HIP(hipMemcpyAsync(&err, w.sync + 1, sizeof(int), hipMemcpyDeviceToHost, h->stream));
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

static int ir_solve_impl(remifa_handle_t h, int trans, int lu_t, int64_t n, int64_t nrhs,
                         const float* d_a, int64_t lda, const void* d_lu, int64_t ldlu,
                         const double* d_b, int64_t ldb, double* d_x, int64_t ldx, double tol,
                         int maxit, int* iters, double* bwd_err, double* history) {
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
  const size_t b_sync = al(trsm_ws_bytes(n));
  const size_t b_w = al(sizeof(float) * n * nrhs);
  const size_t b_r = al(sizeof(double) * n * nrhs);
  const size_t b_part = al(resid_part_bytes(n, std::min<int64_t>(nrhs, RES_RHS)));
  const size_t b_cs = al(sizeof(double) * n);
  const size_t b_nrm = al(sizeof(double) * (2 * nrhs + 2));
  RF_CHECK(ensure_workspace(h, b_sync + b_w + b_r + b_part + b_cs + b_nrm));
  char* p = (char*)h->ws;
  const TrsmWs tw = trsm_ws(p, n);
  float* w = (float*)(p + b_sync);
  double* r = (double*)(p + b_sync + b_w);
  double* part = (double*)(p + b_sync + b_w + b_r);
  double* cs = (double*)(p + b_sync + b_w + b_r + b_part);
  double* nrm = (double*)(p + b_sync + b_w + b_r + b_part + b_cs);
  hipStream_t s = h->stream;
  const dim3 gv((unsigned)ceil_div(n, 256), (unsigned)nrhs);

  // ||op(A)||_1 and ||b||_inf per column (fixed for the whole solve): on
  // the handle's second stream (reading A, an HBM-bound pass) while the
  // diagonal blocks are inverted and the first triangular solves run
  // (latency-bound); joined before the first residual. One stream under
  // graph capture.
  hipStreamCaptureStatus cs_st = hipStreamCaptureStatusNone;
  const bool two = hipStreamIsCapturing(s, &cs_st) == hipSuccess &&
                   cs_st == hipStreamCaptureStatusNone && side_stream(h) == 0;
  hipStream_t sn = s;
  if (two) {
    RF_HIP(hipEventRecord(h->ev_main, s));  // inputs ready
    RF_HIP(hipStreamWaitEvent(h->side, h->ev_main, 0));
    sn = h->side;
  }
  if (trans)
    rowsum_kernel<<<(unsigned)ceil_div(n, 256), 256, 0, sn>>>(n, d_a, lda, cs);
  else
    colsum_kernel<<<(unsigned)n, 256, 0, sn>>>(n, d_a, lda, cs);
  maxreduce_kernel<<<1, 256, 0, sn>>>(n, cs, nrm + 2 * nrhs);
  colmax_kernel<<<(unsigned)nrhs, 256, 0, sn>>>(n, d_b, ldb, nrm + nrhs);
  RF_HIP(hipGetLastError());
  if (two) RF_HIP(hipEventRecord(h->ev_side, sn));
  // x0 = op(LU)^-1 b in fp32
  d2f_kernel<<<gv, 256, 0, s>>>(n, (int)nrhs, d_b, ldb, w, n);
  const int st0 = trsm_invert_launch(h, lu_t, n, d_lu, ldlu, tw);
  const int st1 = st0 ? st0 : getrs_launch(h, lu_t, trans, n, nrhs, d_lu, ldlu, w, n, tw);
  // (joined also when a launch failed: the workspace may be reused next)
  if (two) RF_HIP(hipStreamWaitEvent(s, h->ev_side, 0));
  RF_CHECK(st1);
  axpy_f2d_kernel<<<gv, 256, 0, s>>>(n, w, n, d_x, ldx, 1);
  RF_HIP(hipGetLastError());

  std::vector<double> hn(2 * nrhs + 2);
  int it = 0, done = -1;
  double worst = 0.0;
  for (;; ++it) {
    RF_CHECK(residual_launch(h, n, nrhs, d_a, lda, d_x, ldx, d_b, ldb, r, n, part, trans));
    colmax_kernel<<<(unsigned)nrhs, 256, 0, s>>>(n, r, n, nrm);  // ||r||_inf
    colmax_kernel<<<(unsigned)nrhs, 256, 0, s>>>(n, d_x, ldx, part);  // ||x||_inf (part is free now)
    RF_HIP(hipMemcpyAsync(hn.data(), nrm, sizeof(double) * (2 * nrhs + 1), hipMemcpyDeviceToHost, s));
    std::vector<double> xn(nrhs);
    RF_HIP(hipMemcpyAsync(xn.data(), part, sizeof(double) * nrhs, hipMemcpyDeviceToHost, s));
    int serr = 0;
    RF_HIP(hipMemcpyAsync(&serr, tw.sync + 1, sizeof(int), hipMemcpyDeviceToHost, s));
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
    RF_CHECK(getrs_launch(h, lu_t, trans, n, nrhs, d_lu, ldlu, w, n, tw));
    axpy_f2d_kernel<<<gv, 256, 0, s>>>(n, w, n, d_x, ldx, 0);
    RF_HIP(hipGetLastError());
  }
  if (iters) *iters = done;
  if (bwd_err) *bwd_err = worst;
  return 0;
}

int remifa_ir_solve(remifa_handle_t h, int lu_t, int64_t n, int64_t nrhs, const float* d_a,
                    int64_t lda, const void* d_lu, int64_t ldlu, const double* d_b, int64_t ldb,
                    double* d_x, int64_t ldx, double tol, int maxit, int* iters, double* bwd_err,
                    double* history) {
  return ir_solve_impl(h, 0, lu_t, n, nrhs, d_a, lda, d_lu, ldlu, d_b, ldb, d_x, ldx, tol, maxit,
                       iters, bwd_err, history);
}

int remifa_ir_solve_trans(remifa_handle_t h, int lu_t, int64_t n, int64_t nrhs, const float* d_a,
                          int64_t lda, const void* d_lu, int64_t ldlu, const double* d_b,
                          int64_t ldb, double* d_x, int64_t ldx, double tol, int maxit,
                          int* iters, double* bwd_err, double* history) {
  return ir_solve_impl(h, 1, lu_t, n, nrhs, d_a, lda, d_lu, ldlu, d_b, ldb, d_x, ldx, tol, maxit,
                       iters, bwd_err, history);
}

}  // extern "C"
