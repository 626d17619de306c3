// Update GEMMs C = alpha A B + beta C (src/gemm.cu gemmex; alpha = -1 and
// beta = 1 at every call site of the drivers). Column-major, no transposes.
//
// Matrix-core kernel (fp16 / bf16 / fp32 operands, fp32 accumulation):
//  * 256-thread workgroups (4 wave64s) own a BM x BN tile of C; each wave a
//    (BM/WM) x (BN/WN) sub-tile built from 32x32 MFMA blocks.
//  * The MFMA computes D^T = B^T A^T: its first operand is a B^T fragment
//    (lane = column n of C), its second an A^T fragment (lane = row m of C),
//    so a lane's accumulators hold 32 consecutive rows m across the wave and
//    the epilogue stores coalesce along the column-major C.
//  * Global -> register -> LDS double buffering with one barrier per K step:
//    the next K slab is fetched from HBM while the matrix cores work on the
//    current one. A slabs are stored m-contiguous ([k][m]) and B slabs
//    k-contiguous ([n][k]), both with 16-byte stores; B fragments are one
//    16-byte LDS read, A fragments two LDS transpose reads (conflict-free
//    thanks to the row padding).
//  * Workgroup ids are remapped so that each of the 8 XCDs (which own
//    private L2s) processes a contiguous range of tiles.
//  * Optional split-K: slice z of K writes alpha*partial to a workspace and a
//    reduction kernel adds the slices in order (deterministic). Unless a rule
//    or the caller fixes it, the split count comes from a cost model: whole
//    rounds of resident workgroups times the K slabs per workgroup, plus the
//    reduction's traffic. The update GEMMs of an LU have few output tiles
//    and long K late in the factorization, where one round of workgroups
//    would leave most CUs idle.
//
// fp16-accumulating compute types (FP16, TC16: cublasHgemm / GemmEx 16F) run
// on a VALU kernel with one correctly rounded fp16 fma per k, in k order.
#include <algorithm>
#include <type_traits>

#include "common.hpp"

namespace remifa {
namespace {

#include "gemm_mfma.hpp"
#include "gemm_big.hpp"
#include "gemm_fp8.hpp"

// C = beta*C + sum_z W[z] (slices added in order z = 0, 1, ...)
template <typename OutT>
__global__ void __launch_bounds__(256) splitk_reduce_kernel(int64_t M, int64_t N, int split,
                                                            const float* __restrict__ W,
                                                            int64_t ldw, float beta,
                                                            OutT* __restrict__ C, int64_t ldc,
                                                            const void* Cin, int64_t ldcin,
                                                            int cin_half) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  for (int64_t j = blockIdx.y; j < N; j += gridDim.y) {
    float s = 0.0f;
    for (int z = 0; z < split; ++z) s += W[(z * N + j) * ldw + i];
    OutT* cp = C + i + j * ldc;
    float cv = 0.0f;
    if (beta != 0.0f)
      cv = Cin == nullptr ? (float)*cp
                          : (cin_half ? (float)((const half_t*)Cin)[i + j * ldcin]
                                      : ((const float*)Cin)[i + j * ldcin]);
    *cp = (OutT)((beta != 0.0f) ? s + beta * cv : s);
  }
}

// The same sums, four rows per thread with 16-byte loads and JPB columns
// per workgroup (rows a multiple of 4, 16-byte aligned partials, C and Cin):
// few, large workgroups instead of one per 256 elements.
template <typename OutT, int JPB>
__global__ void __launch_bounds__(256) splitk_reduce4_kernel(int64_t M, int64_t N, int split,
                                                             const float* __restrict__ W,
                                                             int64_t ldw, float beta,
                                                             OutT* __restrict__ C, int64_t ldc,
                                                             const void* Cin, int64_t ldcin,
                                                             int cin_half) {
  const int64_t i = ((int64_t)blockIdx.x * 256 + threadIdx.x) * 4;
  if (i >= M) return;
  const int64_t j0 = (int64_t)blockIdx.y * JPB;
#pragma unroll
  for (int jj = 0; jj < JPB; ++jj) {
    const int64_t j = j0 + jj;
    if (j >= N) break;
    float s[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (int z = 0; z < split; ++z) {
      const float4 w = *(const float4*)&W[(z * N + j) * ldw + i];
      s[0] += w.x;
      s[1] += w.y;
      s[2] += w.z;
      s[3] += w.w;
    }
    OutT* cp = C + i + j * ldc;
    if (beta != 0.0f) {
      float cv[4];
      if (Cin == nullptr || cin_half) {
        const bool h16 = Cin ? true : sizeof(OutT) == 2;
        if (h16) {
          const half_t* hp = Cin ? (const half_t*)Cin + i + j * ldcin : (const half_t*)cp;
          typedef _Float16 hv4 __attribute__((ext_vector_type(4)));
          const hv4 v = *(const hv4*)hp;
          for (int e = 0; e < 4; ++e) cv[e] = (float)v[e];
        } else {
          const float4 v = *(const float4*)cp;
          cv[0] = v.x, cv[1] = v.y, cv[2] = v.z, cv[3] = v.w;
        }
      } else {
        const float4 v = *(const float4*)((const float*)Cin + i + j * ldcin);
        cv[0] = v.x, cv[1] = v.y, cv[2] = v.z, cv[3] = v.w;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) s[e] = s[e] + beta * cv[e];
    }
    if constexpr (sizeof(OutT) == 4) {
      *(float4*)cp = make_float4(s[0], s[1], s[2], s[3]);
    } else {
      typedef _Float16 hv4 __attribute__((ext_vector_type(4)));
      const hv4 v = {(_Float16)s[0], (_Float16)s[1], (_Float16)s[2], (_Float16)s[3]};
      *(hv4*)cp = v;
    }
  }
}

// C = beta C (or Cin) + the split partials, on the handle's stream
int splitk_reduce_launch(remifa_handle_t h, int out_t, int64_t m, int64_t n, int split,
                         const float* W, int64_t ldw, float beta, void* c, int64_t ldc,
                         const void* cin, int64_t ldcin, int cin_half) {
  const size_t es = out_t == REMIFA_R_32F ? 4 : 2;
  const size_t ces = cin ? (cin_half ? 2 : 4) : es;
  const bool vec = m % 4 == 0 && ldw % 4 == 0 && ldc % 4 == 0 && (uintptr_t)W % 16 == 0 &&
                   (uintptr_t)c % (4 * es) == 0 && (!cin || (ldcin % 4 == 0 &&
                                                             (uintptr_t)cin % (4 * ces) == 0));
  if (vec) {
    constexpr int JPB = 4;
    dim3 grid((unsigned)ceil_div(m, 1024), (unsigned)ceil_div(n, JPB));
    if (out_t == REMIFA_R_32F)
      splitk_reduce4_kernel<float, JPB><<<grid, 256, 0, h->stream>>>(
          m, n, split, W, ldw, beta, (float*)c, ldc, cin, ldcin, cin_half);
    else
      splitk_reduce4_kernel<half_t, JPB><<<grid, 256, 0, h->stream>>>(
          m, n, split, W, ldw, beta, (half_t*)c, ldc, cin, ldcin, cin_half);
  } else {
    dim3 grid((unsigned)ceil_div(m, 256), (unsigned)(n < 65535 ? n : 65535));
    if (out_t == REMIFA_R_32F)
      splitk_reduce_kernel<float><<<grid, 256, 0, h->stream>>>(m, n, split, W, ldw, beta,
                                                               (float*)c, ldc, cin, ldcin,
                                                               cin_half);
    else
      splitk_reduce_kernel<half_t><<<grid, 256, 0, h->stream>>>(m, n, split, W, ldw, beta,
                                                                (half_t*)c, ldc, cin, ldcin,
                                                                cin_half);
  }
  return hipGetLastError() == hipSuccess ? 0 : REMIFA_ERR_HIP;
}

// fp16-accumulating GEMM (FP16 / TC16): 64x64 tile per 256 threads, each
// thread a 4x4 block; acc_ij <- fma16(a_ik, b_kj, acc_ij) for k = 0, 1, ...
// then C <- fma16(alpha, acc, beta * C).
__global__ void __launch_bounds__(256) gemm_h16_kernel(int64_t M, int64_t N, int64_t K,
                                                       const half_t* __restrict__ A, int64_t lda,
                                                       const half_t* __restrict__ B, int64_t ldb,
                                                       half_t alpha, half_t beta,
                                                       half_t* __restrict__ C, int64_t ldc) {
  __shared__ half_t As[16][64 + 2];
  __shared__ half_t Bs[64][16 + 2];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;  // 16 x 16 threads, 4x4 each
  const int64_t m0 = (int64_t)blockIdx.x * 64, n0 = (int64_t)blockIdx.y * 64;
  half_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = (half_t)0;
  for (int64_t k0 = 0; k0 < K; k0 += 16) {
    for (int e = tid; e < 16 * 64; e += 256) {
      int mm = e % 64, kk = e / 64;
      int64_t gm = m0 + mm, gk = k0 + kk;
      As[kk][mm] = (gm < M && gk < K) ? A[gm + gk * lda] : (half_t)0;
      int kb = e % 16, nn = e / 16;
      int64_t gk2 = k0 + kb, gn = n0 + nn;
      Bs[nn][kb] = (gn < N && gk2 < K) ? B[gk2 + gn * ldb] : (half_t)0;
    }
    __syncthreads();
    const int kn = (int)((K - k0) < 16 ? (K - k0) : 16);
    for (int kk = 0; kk < kn; ++kk) {
      half_t a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[ty + 16 * j][kk];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_fmaf16(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      int64_t gm = m0 + tx + 16 * i, gn = n0 + ty + 16 * j;
      if (gm < M && gn < N) {
        half_t* cp = C + gm + gn * ldc;
        half_t bc = (beta != (half_t)0) ? beta * *cp : (half_t)0;
        *cp = __builtin_fmaf16(alpha, acc[i][j], bc);
      }
    }
}

// K slab depth: 16 for fp32 operands; for 16-bit operands 64, except the
// tiles of 256 x 128 and up (32: their accumulators leave no room for a
// deeper prefetch, and two 64-deep slabs would not fit the LDS)
constexpr int slab_k(int opk, int bm, int bn) {
  return opk == OPK_F32 ? 16 : (bm * bn >= 256 * 128 ? 32 : 64);
}

template <int OPK, typename OutT, int BM, int BN, int WM, int WN>
int launch_mfma(remifa_handle_t h, GemmArgs g, int split, bool vec) {
  constexpr int BK = slab_k(OPK, BM, BN);
  g.tiles_m = (int)ceil_div(g.M, BM);
  g.tiles_n = (int)ceil_div(g.N, BN);
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), 1, (unsigned)split);
  // No accumulator preload from C: the products are summed from zero and C is
  // added once in the epilogue, as cuBLAS computes alpha (A B) + beta C.
  // Accumulating onto C directly costs accuracy when |C| >> |A||B| (the
  // diagonally dominant trailing matrices of this library).
  constexpr bool PRE = false;
  if (vec)
    gemm_mfma_kernel<OPK, OutT, BM, BN, WM, WN, BK, true, PRE><<<grid, 256, 0, h->stream>>>(g);
  else
    gemm_mfma_kernel<OPK, OutT, BM, BN, WM, WN, BK, false, PRE><<<grid, 256, 0, h->stream>>>(g);
  RF_HIP(hipGetLastError());
  return 0;
}

// tile shapes (BM x BN); the timer's variant code of a launch is
// opk << 8 | fp16-output << 4 | tile
enum Tile {
  T256x32 = 0,
  T32x256 = 1,
  T256x64 = 2,
  T64x256 = 3,
  T256x256 = 4,
  T128x128 = 5,
  T256x128 = 6,
  T128x256 = 7
};

// 128x128 for large outputs (the measured best of the square tiles); narrow
// shapes (N or M = ib .. 64) use 256x32 / 256x64 and their transposes so that
// skinny updates keep the chip busy.
int choose_tile(int opk, int64_t M, int64_t N) {
  if (N <= 32) return T256x32;
  if (M <= 32) return T32x256;
  if (opk == OPK_F32) return T128x128;
  if (N <= 64) return T256x64;
  if (M <= 64) return T64x256;
  return T128x128;
}

// tile geometry for the split-K cost model: BM, BN and the workgroups a CU
// holds at once (LDS / register limited)
struct TileGeom {
  int bm, bn, slots;
};
TileGeom tile_geom(int tile) {
  switch (tile) {
    case T256x32: return {256, 32, 1};
    case T32x256: return {32, 256, 1};
    case T256x64: return {256, 64, 1};
    case T64x256: return {64, 256, 1};
    case T256x256: return {256, 256, 1};
    case T256x128: return {256, 128, 1};
    case T128x256: return {128, 256, 1};
    default: return {128, 128, 2};
  }
}

// split count minimising the modelled time: ceil(tiles * s / resident) rounds
// of ceil(slabs / s) K slabs each, plus the reduction (s partials read, C
// read and written) and its launch. rate: matrix-core flop rate of one CU
// running this kernel at full occupancy (measured: ~3.5 TFLOP/s for 16-bit
// operands on MI355X, ~0.45 for fp32). kmax_split: splits whose partials
// exceed `avail` bytes are not considered.
int split_model(int64_t M, int64_t N, int64_t K, TileGeom tg, int num_cu, double rate, int bk,
                size_t avail) {
  const int64_t tiles = ceil_div(M, tg.bm) * ceil_div(N, tg.bn);
  const int64_t res = (int64_t)tg.slots * num_cu;
  const int64_t nk = ceil_div(K, bk);
  const double t_slab = 2.0 * tg.bm * tg.bn * bk * tg.slots / rate;
  int best = 1;
  double tbest = (double)ceil_div(tiles, res) * nk * t_slab;
  for (int s = 2; s <= 16 && s <= nk; ++s) {
    const int64_t kc = ceil_div(nk, s);
    const int64_t se = ceil_div(nk, kc);  // slices actually used
    if (sizeof(float) * (size_t)se * (size_t)M * (size_t)N > avail) break;
    const double t = (double)ceil_div(tiles * se, res) * kc * t_slab +
                     (double)(se + 2) * M * N * 4.0 / 3.0e12 + 4.0e-6;
    if (t < tbest * 0.97) best = (int)se, tbest = t;
  }
  return best;
}

// 256 x 256 tiles on gemm_big_kernel: 16-bit operands in memory, M and N
// multiples of 256, K a multiple of the 32-deep slab (split-K chunks are
// multiples of 64), 16-byte aligned operands and column offsets of one slab
// within 32 bits (the DMA lane offsets)
bool big_fits(int opk, int64_t m, int64_t n, int64_t k, const void* a, int64_t lda, const void* b,
              int64_t ldb) {
  return (opk == OPK_F16 || opk == OPK_BF16S) && m % 256 == 0 && n % 256 == 0 && k % 32 == 0 &&
         (uintptr_t)a % 16 == 0 && (uintptr_t)b % 16 == 0 && lda % 8 == 0 && ldb % 8 == 0 &&
         (64 * lda + 256) * 2 < (int64_t)1 << 31 && (256 * ldb + 64) * 2 < (int64_t)1 << 31;
}

// 32-deep slabs, four LDS stages (128 KB), eight wavefronts: with four the
// LDS-DMA requests stalled the one wavefront per SIMD that also issues the
// matrix-core work (964 -> 1265 TFLOP/s at 8192^3 with eight, MI355X)
template <int OPK, typename OutT>
int launch_big(remifa_handle_t h, GemmArgs g, int split) {
  g.tiles_m = (int)(g.M / 256);
  g.tiles_n = (int)(g.N / 256);
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), 1, (unsigned)split);
  gemm_big_kernel<OPK, OutT, 32, 4, 8><<<grid, 512, 0, h->stream>>>(g);
  RF_HIP(hipGetLastError());
  return 0;
}

constexpr int kSplitGranule = 64;  // split-K granularity: a multiple of every slab depth

// The tile and split-K count of an update GEMM (operand kind opk, K >= 0),
// before the workspace limits the split: the autotuned rule for the shape
// (first match), else the heuristics; the caller's split, the rule's, else
// the cost model's on the CUs not held by concurrent work.
void plan_gemm(remifa_handle_t h, int opk, int64_t m, int64_t n, int64_t K, const void* a,
               int64_t lda, const void* b, int64_t ldb, int& tile, int& split) {
  int rule_tile = -1, rule_split = 0;
  for (const auto& r : h->gemm_rules) {
    if ((r[0] < 0 || r[0] == opk) && m >= r[1] && m <= r[2] && n >= r[3] && n <= r[4]) {
      rule_tile = (int)r[5];
      rule_split = (int)r[6];
      break;
    }
  }
  tile = choose_tile(opk, m, n);
  auto fits = [&](int t) {
    return t >= 0 && (opk != OPK_F32 || t == T256x32 || t == T32x256 || t == T128x128);
  };
  if (fits(rule_tile)) tile = rule_tile;  // autotuned table
  // below K = 2048 the 256 x 256 kernel's start-up and epilogue outweigh its
  // loop (measured on MI355X at N = 512): 128 x 128 unless forced
  if (tile == T256x256 && K < 2048) tile = T128x128;
  if (fits(h->gemm_tile)) tile = h->gemm_tile;  // explicit override (autotuner runs)
  if (tile == T256x256 && !big_fits(opk, m, n, K, a, lda, b, ldb)) tile = T128x128;
  split = h->split_k > 0 ? h->split_k : rule_split;
  if (split <= 0) {
    const bool standalone = h->ws_used == 0 && !h->ws_user;
    const size_t avail = standalone ? std::max<size_t>(h->ws_bytes, (size_t)256 << 20)
                                    : (h->ws_bytes > h->ws_used ? h->ws_bytes - h->ws_used : 0);
    split = split_model(m, n, K, tile_geom(tile), std::max(1, h->num_cu - h->cu_busy),
                        opk == OPK_F32 ? 0.45e12 : 3.5e12, kSplitGranule, avail);
  }
  if (K == 0) split = 1;
  const int64_t kc = ceil_div(ceil_div(K > 0 ? K : 1, split), kSplitGranule) * kSplitGranule;
  split = (int)ceil_div(K > 0 ? K : 1, kc);
}

template <int OPK, typename OutT>
int dispatch_basic(remifa_handle_t h, int tile, GemmArgs g, int split, bool vec) {
  switch (tile) {
    case T256x32: return launch_mfma<OPK, OutT, 256, 32, 4, 1>(h, g, split, vec);
    case T32x256: return launch_mfma<OPK, OutT, 32, 256, 1, 4>(h, g, split, vec);
    default: return launch_mfma<OPK, OutT, 128, 128, 2, 2>(h, g, split, vec);
  }
}

template <int OPK, typename OutT>
int dispatch_tiles(remifa_handle_t h, int tile, GemmArgs g, int split, bool vec) {
  switch (tile) {
    case T256x32: return launch_mfma<OPK, OutT, 256, 32, 4, 1>(h, g, split, vec);
    case T32x256: return launch_mfma<OPK, OutT, 32, 256, 1, 4>(h, g, split, vec);
    case T128x128: return launch_mfma<OPK, OutT, 128, 128, 2, 2>(h, g, split, vec);
  }
  if constexpr (OPK != OPK_F32) {
    switch (tile) {
      case T256x64: return launch_mfma<OPK, OutT, 256, 64, 4, 1>(h, g, split, vec);
      case T64x256: return launch_mfma<OPK, OutT, 64, 256, 1, 4>(h, g, split, vec);
      case T256x128: return launch_mfma<OPK, OutT, 256, 128, 2, 2>(h, g, split, vec);
      case T128x256: return launch_mfma<OPK, OutT, 128, 256, 2, 2>(h, g, split, vec);
    }
  }
  return REMIFA_ERR_UNSUPPORTED;
}

template <int OPK, typename OutT>
int launch_pair(remifa_handle_t h, GemmArgs g0, GemmArgs g1, bool vec) {
  // 32-deep slabs: 49 KB of LDS, three workgroups per CU, so both halves of
  // the pair (256 + 256 workgroups at the LU's shapes) run in one round
  constexpr int BK = OPK == OPK_F32 ? 16 : 32;
  g0.tiles_m = (int)ceil_div(g0.M, 256);
  g0.tiles_n = (int)ceil_div(g0.N, 32);
  g1.tiles_m = (int)ceil_div(g1.M, 32);
  g1.tiles_n = (int)ceil_div(g1.N, 256);
  const int nt0 = g0.tiles_m * g0.tiles_n;
  dim3 grid((unsigned)(nt0 + g1.tiles_m * g1.tiles_n));
  if (vec)
    gemm_pair_kernel<OPK, OutT, BK, true><<<grid, 256, 0, h->stream>>>(g0, g1, nt0);
  else
    gemm_pair_kernel<OPK, OutT, BK, false><<<grid, 256, 0, h->stream>>>(g0, g1, nt0);
  RF_HIP(hipGetLastError());
  return 0;
}

template <int OPK, typename OutT>
int launch_dual(remifa_handle_t h, GemmArgs g0, GemmArgs g1, bool vec) {
  constexpr int BK = slab_k(OPK, 128, 128);
  g0.tiles_m = (int)ceil_div(g0.M, 128);
  g0.tiles_n = (int)ceil_div(g0.N, 128);
  g1.tiles_m = (int)ceil_div(g1.M, 128);
  g1.tiles_n = (int)ceil_div(g1.N, 128);
  const int nt0 = g0.tiles_m * g0.tiles_n;
  dim3 grid((unsigned)(nt0 + g1.tiles_m * g1.tiles_n));
  if (vec)
    gemm_dual_kernel<OPK, OutT, BK, true><<<grid, 256, 0, h->stream>>>(g0, g1, nt0);
  else
    gemm_dual_kernel<OPK, OutT, BK, false><<<grid, 256, 0, h->stream>>>(g0, g1, nt0);
  RF_HIP(hipGetLastError());
  return 0;
}

}  // namespace

// C0 -= A0 B0 (m0 x n0) and C1 -= A1 B1 (m1 x n1), fp32 C, inner size k: one
// launch when both would take an unsplit 128 x 128 launch with the same
// operand access (fp16 / bf16 operands, TC32 / BF16_TC32), else two
// gemm_launch calls. Same tiles, hence the same bits, either way.
int gemm_dual_launch(remifa_handle_t h, int ctype, int in_t, int64_t m0, int64_t n0, int64_t k,
                     const void* a0, int64_t lda0, const void* b0, int64_t ldb0, void* c0,
                     int64_t ldc0, int64_t m1, int64_t n1, const void* a1, int64_t lda1,
                     const void* b1, int64_t ldb1, void* c1, int64_t ldc1) {
  int opk = -1;
  if (ctype == REMIFA_TC32 && in_t == REMIFA_R_16F) opk = OPK_F16;
  if (ctype == REMIFA_BF16_TC32 && in_t == REMIFA_R_16F) opk = OPK_BF16;
  auto vec_ok = [&](const void* a, int64_t lda, const void* b, int64_t ldb, int64_t m) {
    return ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) && (lda % 8 == 0) &&
           (ldb % 8 == 0) && (m % 8 == 0) && (k % 8 == 0);
  };
  bool dual = opk >= 0 && k > 0 && m0 > 0 && n0 > 0 && m1 > 0 && n1 > 0;
  bool v0 = false, v1 = false;
  if (dual) {
    int t0, s0, t1, s1;
    plan_gemm(h, opk, m0, n0, k, a0, lda0, b0, ldb0, t0, s0);
    plan_gemm(h, opk, m1, n1, k, a1, lda1, b1, ldb1, t1, s1);
    v0 = vec_ok(a0, lda0, b0, ldb0, m0);
    v1 = vec_ok(a1, lda1, b1, ldb1, m1);
    dual = t0 == T128x128 && t1 == T128x128 && s0 == 1 && s1 == 1 && v0 == v1;
  }
  if (!dual) {
    RF_CHECK(gemm_launch(h, ctype, in_t, REMIFA_R_32F, m0, n0, k, -1.0f, a0, lda0, b0, ldb0, 1.0f,
                         c0, ldc0));
    return gemm_launch(h, ctype, in_t, REMIFA_R_32F, m1, n1, k, -1.0f, a1, lda1, b1, ldb1, 1.0f,
                       c1, ldc1);
  }
  auto args = [&](int64_t m, int64_t n, const void* a, int64_t lda, const void* b, int64_t ldb,
                  void* c, int64_t ldc) {
    GemmArgs g{};
    g.M = m, g.N = n, g.K = k, g.lda = lda, g.ldb = ldb, g.ldc = ldc;
    g.alpha = -1.0f, g.beta = 1.0f, g.A = a, g.B = b, g.C = c;
    g.kchunk = ceil_div(k, kSplitGranule) * kSplitGranule;  // one K slice
    return g;
  };
  const GemmArgs g0 = args(m0, n0, a0, lda0, b0, ldb0, c0, ldc0);
  const GemmArgs g1 = args(m1, n1, a1, lda1, b1, ldb1, c1, ldc1);
  timer_begin(h, KC_GEMM, 2.0 * (double)k * (m0 * n0 + m1 * n1), (opk << 8) | T128x128);
  const int st = opk == OPK_F16 ? launch_dual<OPK_F16, float>(h, g0, g1, v0)
                                : launch_dual<OPK_BF16, float>(h, g0, g1, v0);
  timer_end(h);
  return st;
}

// The two slice updates of the inner LL factorization in one launch when
// they take the narrow tiles without split-K (C0 -= A0 B0 with N0 <= 32 per
// tile column, C1 -= A1 B1 with M1 <= 64); otherwise two gemm_launch calls.
int gemm_pair_launch(remifa_handle_t h, int ctype, int in_t, int out_t, int64_t m0, int64_t n0,
                     int64_t k, const void* a0, int64_t lda0, const void* b0, int64_t ldb0,
                     void* c0, int64_t ldc0, int64_t m1, int64_t n1, const void* a1, int64_t lda1,
                     const void* b1, int64_t ldb1, void* c1, int64_t ldc1) {
  int opk = -1;
  if (ctype == REMIFA_FP32 && in_t == REMIFA_R_32F) opk = OPK_F32;
  if (ctype == REMIFA_TC32 && in_t == REMIFA_R_16F) opk = OPK_F16;
  if (ctype == REMIFA_BF16_TC32 && in_t == REMIFA_R_16F) opk = OPK_BF16;
  const size_t es = (opk == OPK_F32) ? 4 : 2;
  const int chunk = 16 / (int)es;
  auto vec_ok = [&](const void* a, int64_t lda, const void* b, int64_t ldb, int64_t m) {
    return ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) && (lda % chunk == 0) &&
           (ldb % chunk == 0) && (m % chunk == 0) && (k % chunk == 0);
  };
  // the rule table / overrides must pick the narrow tiles without split-K
  auto rule = [&](int64_t m, int64_t n, int& tile, int& split) {
    tile = choose_tile(opk, m, n);
    split = 0;
    for (const auto& r : h->gemm_rules)
      if ((r[0] < 0 || r[0] == opk) && m >= r[1] && m <= r[2] && n >= r[3] && n <= r[4]) {
        tile = (int)r[5];
        split = (int)r[6];
        break;
      }
  };
  int t0, s0, t1, s1;
  if (opk >= 0) {
    rule(m0, n0, t0, s0);
    rule(m1, n1, t1, s1);
  }
  const bool v0 = opk >= 0 && vec_ok(a0, lda0, b0, ldb0, m0);
  const bool v1 = opk >= 0 && vec_ok(a1, lda1, b1, ldb1, m1);
  const bool pair = opk >= 0 && k > 0 && m0 > 0 && n0 > 0 && m1 > 0 && n1 > 0 && t0 == T256x32 &&
                    t1 == T32x256 && s0 <= 1 && s1 <= 1 && h->split_k <= 0 && h->gemm_tile < 0 &&
                    v0 == v1 && (out_t == REMIFA_R_32F || out_t == REMIFA_R_16F);
  if (!pair) {
    RF_CHECK(gemm_launch(h, ctype, in_t, out_t, m0, n0, k, -1.0f, a0, lda0, b0, ldb0, 1.0f, c0,
                         ldc0));
    return gemm_launch(h, ctype, in_t, out_t, m1, n1, k, -1.0f, a1, lda1, b1, ldb1, 1.0f, c1,
                       ldc1);
  }
  auto args = [&](int64_t m, int64_t n, const void* a, int64_t lda, const void* b, int64_t ldb,
                  void* c, int64_t ldc) {
    GemmArgs g{};
    g.M = m, g.N = n, g.K = k, g.lda = lda, g.ldb = ldb, g.ldc = ldc;
    g.alpha = -1.0f, g.beta = 1.0f, g.A = a, g.B = b, g.C = c;
    g.kchunk = k;  // one K slice
    return g;
  };
  GemmArgs g0 = args(m0, n0, a0, lda0, b0, ldb0, c0, ldc0);
  GemmArgs g1 = args(m1, n1, a1, lda1, b1, ldb1, c1, ldc1);
  timer_begin(h, KC_GEMM, 2.0 * (double)k * (m0 * n0 + m1 * n1),
              (opk << 8) | ((out_t == REMIFA_R_16F) << 4) | 8);
  int st;
  if (opk == OPK_F32)
    st = out_t == REMIFA_R_32F ? launch_pair<OPK_F32, float>(h, g0, g1, v0)
                               : launch_pair<OPK_F32, half_t>(h, g0, g1, v0);
  else if (opk == OPK_F16)
    st = out_t == REMIFA_R_32F ? launch_pair<OPK_F16, float>(h, g0, g1, v0)
                               : launch_pair<OPK_F16, half_t>(h, g0, g1, v0);
  else
    st = out_t == REMIFA_R_32F ? launch_pair<OPK_BF16, float>(h, g0, g1, v0)
                               : launch_pair<OPK_BF16, half_t>(h, g0, g1, v0);
  timer_end(h);
  return st;
}

int fp8_quantize_launch(remifa_handle_t h, int in_t, bool rows, int64_t R, int64_t K,
                        const void* src, int64_t ld, uint8_t* q, int64_t ldq, uint8_t* s,
                        int64_t lds) {
  if (R <= 0 || K <= 0) return 0;
  if (!src || !q || !s) return -6;
  if (ldq % 16 || (uintptr_t)q % 16 || ldq < ceil_div(K, 32) * 32 || lds < ceil_div(K, 32))
    return -7;
  if (ld < (rows ? R : K)) return -5;
  dim3 grid((unsigned)ceil_div(R, 256), (unsigned)ceil_div(K, 32));
  timer_begin(h, KC_CONVERT, (double)R * K * ((in_t == REMIFA_R_32F ? 4 : 2) + 1), 8);
  if (in_t == REMIFA_R_16F) {
    if (rows)
      quant_fp8_kernel<half_t, true><<<grid, 256, 0, h->stream>>>(R, K, (const half_t*)src, ld, q,
                                                                  ldq, s, lds);
    else
      quant_fp8_kernel<half_t, false><<<grid, 256, 0, h->stream>>>(R, K, (const half_t*)src, ld,
                                                                   q, ldq, s, lds);
  } else if (in_t == REMIFA_R_32F) {
    if (rows)
      quant_fp8_kernel<float, true><<<grid, 256, 0, h->stream>>>(R, K, (const float*)src, ld, q,
                                                                 ldq, s, lds);
    else
      quant_fp8_kernel<float, false><<<grid, 256, 0, h->stream>>>(R, K, (const float*)src, ld, q,
                                                                  ldq, s, lds);
  } else {
    timer_end(h);
    return -4;
  }
  timer_end(h);
  RF_HIP(hipGetLastError());
  return 0;
}

int fp8_gemm_launch(remifa_handle_t h, int out_t, int64_t m, int64_t n, int64_t k, float alpha,
                    const uint8_t* qa, int64_t ldqa, const uint8_t* sa, int64_t ldsa,
                    const uint8_t* qb, int64_t ldqb, const uint8_t* sb, int64_t ldsb, float beta,
                    void* c, int64_t ldc, const void* cin, int cin_t, int64_t ldcin) {
  if (m <= 0 || n <= 0) return 0;
  if (k < 0 || k % 32) return -3;
  if (out_t != REMIFA_R_32F && out_t != REMIFA_R_16F) return -5;
  if (!c || (k > 0 && (!qa || !qb || !sa || !sb))) return -6;
  if (cin && ((cin_t != REMIFA_R_16F && cin_t != REMIFA_R_32F) || ldcin < m)) return -7;
  const int64_t nkb4 = ceil_div(ceil_div(k, 32), 4) * 4;
  if (ldqa % 16 || ldqb % 16 || (uintptr_t)qa % 16 || (uintptr_t)qb % 16 || ldqa < k ||
      ldqb < k || ldsa % 4 || ldsb % 4 || (uintptr_t)sa % 4 || (uintptr_t)sb % 4 ||
      ldsa < nkb4 || ldsb < nkb4 || ldc < m)
    return -7;
  Fp8Args g{};
  g.M = m, g.N = n, g.K = k;
  g.A = qa, g.lda = ldqa, g.SA = sa, g.ldsa = ldsa;
  g.B = qb, g.ldb = ldqb, g.SB = sb, g.ldsb = ldsb;
  g.alpha = alpha, g.beta = beta, g.C = c, g.ldc = ldc;
  g.Cin = cin, g.ldcin = ldcin, g.cin_half = cin_t == REMIFA_R_16F;
  g.tiles_m = (int)ceil_div(m, 128);
  g.tiles_n = (int)ceil_div(n, 128);
  constexpr int BK = 128;
  int split = h->split_k;
  const size_t avail = h->ws_bytes > h->ws_used ? h->ws_bytes - h->ws_used : 0;
  if (split <= 0)
    split = k > 0 ? split_model(m, n, k, TileGeom{128, 128, 2}, std::max(1, h->num_cu - h->cu_busy),
                                7.0e12, BK, avail)
                  : 1;
  g.kchunk = ceil_div(ceil_div(k > 0 ? k : 1, split), BK) * BK;
  split = (int)ceil_div(k > 0 ? k : 1, g.kchunk);
  while (split > 1 && sizeof(float) * (size_t)split * m * n > avail) {
    split = (split + 1) / 2;
    g.kchunk = ceil_div(ceil_div(k, split), BK) * BK;
    split = (int)ceil_div(k, g.kchunk);
  }
  if (split > 1) {
    g.W = (float*)((char*)h->ws + h->ws_used);
    g.ldw = m;
  }
  timer_begin(h, KC_GEMM, 2.0 * (double)m * n * k, (3 << 8) | ((out_t == REMIFA_R_16F) << 4) | T128x128);
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), 1, (unsigned)split);
  if (out_t == REMIFA_R_32F)
    gemm_fp8_kernel<float><<<grid, 256, 0, h->stream>>>(g);
  else
    gemm_fp8_kernel<half_t><<<grid, 256, 0, h->stream>>>(g);
  int st = hipGetLastError() == hipSuccess ? 0 : REMIFA_ERR_HIP;
  timer_end(h);
  if (st == 0 && split > 1) {
    // the reduction is timed as its own variant (tile code 15)
    timer_begin(h, KC_GEMM, 0.0, (3 << 8) | ((out_t == REMIFA_R_16F) << 4) | 15);
    st = splitk_reduce_launch(h, out_t, m, n, split, g.W, g.ldw, beta, c, ldc, g.Cin, g.ldcin,
                              g.cin_half);
    timer_end(h);
  }
  return st;
}

int gemm_launch(remifa_handle_t h, int ctype, int in_t, int out_t, int64_t m, int64_t n, int64_t k,
                float alpha, const void* a, int64_t lda, const void* b, int64_t ldb, float beta,
                void* c, int64_t ldc) {
  return gemm_launch_cin(h, ctype, in_t, out_t, m, n, k, alpha, a, lda, b, ldb, beta, nullptr, 0,
                         0, c, ldc);
}

int gemm_launch_cin(remifa_handle_t h, int ctype, int in_t, int out_t, int64_t m, int64_t n,
                    int64_t k, float alpha, const void* a, int64_t lda, const void* b, int64_t ldb,
                    float beta, const void* cin, int cin_t, int64_t ldcin, void* c, int64_t ldc) {
  if (m <= 0 || n <= 0) return 0;
  if (cin && (ctype == REMIFA_FP16 || ctype == REMIFA_TC16 || ctype == REMIFA_FP8_TC32 ||
              (cin_t != REMIFA_R_16F && cin_t != REMIFA_R_32F)))
    return REMIFA_ERR_UNSUPPORTED;
  const double flops = 2.0 * (double)m * (double)n * (double)(k > 0 ? k : 0);
  if (k <= 0 || alpha == 0.0f) {
    if (beta == 1.0f && !cin) return 0;
    // C = beta * C: run the GEMM with K = 0 (the epilogue scales)
  }
  int opk;
  switch (ctype) {
    case REMIFA_FP32: opk = OPK_F32; if (in_t != REMIFA_R_32F) return -4; break;
    case REMIFA_TC32: opk = OPK_F16; if (in_t != REMIFA_R_16F) return -4; break;
    case REMIFA_BF16_TC32:
      // fp16 storage rounded to bf16 while staging, or bf16 storage as is
      if (in_t == REMIFA_R_16F) opk = OPK_BF16;
      else if (in_t == REMIFA_R_16BF) opk = OPK_BF16S;
      else return -4;
      break;
    case REMIFA_FP8_TC32: {
      // quantize A by rows and B by columns into the free workspace, then
      // the scaled matrix-core GEMM (K padded to whole 32-blocks with zeros)
      if (in_t != REMIFA_R_16F && in_t != REMIFA_R_32F) return -4;
      if (out_t != REMIFA_R_32F && out_t != REMIFA_R_16F) return -5;
      const int64_t kp = ceil_div(k > 0 ? k : 0, 32) * 32;
      const int64_t ldq = std::max<int64_t>(ceil_div(kp, 128) * 128, 128);
      const int64_t lds = ldq / 32;
      auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
      const size_t qbytes = al((size_t)m * ldq) + al((size_t)m * lds) + al((size_t)n * ldq) +
                            al((size_t)n * lds);
      if (h->ws_used == 0 && !h->ws_user) {
        const size_t part = std::min<size_t>((size_t)256 << 20, (size_t)16 * 4 * m * n);
        RF_CHECK(ensure_workspace(h, qbytes + part));
      }
      if (h->ws_used + qbytes > h->ws_bytes) return REMIFA_ERR_WORKSPACE;
      char* base = (char*)h->ws + h->ws_used;
      uint8_t* qa = (uint8_t*)base;
      uint8_t* sa = qa + al((size_t)m * ldq);
      uint8_t* qb = sa + al((size_t)m * lds);
      uint8_t* sb = qb + al((size_t)n * ldq);
      if (kp > 0) {
        RF_CHECK(fp8_quantize_launch(h, in_t, true, m, k, a, lda, qa, ldq, sa, lds));
        RF_CHECK(fp8_quantize_launch(h, in_t, false, n, k, b, ldb, qb, ldq, sb, lds));
      }
      const size_t saved = h->ws_used;
      h->ws_used += qbytes;
      const int st = fp8_gemm_launch(h, out_t, m, n, kp, alpha, qa, ldq, sa, lds, qb, ldq, sb, lds,
                                     beta, c, ldc);
      h->ws_used = saved;
      return st;
    }
    case REMIFA_FP16:
    case REMIFA_TC16: {
      if (in_t != REMIFA_R_16F) return -4;
      if (out_t != REMIFA_R_16F) return -5;
      timer_begin(h, KC_GEMM, flops);
      dim3 grid((unsigned)ceil_div(m, 64), (unsigned)ceil_div(n, 64));
      gemm_h16_kernel<<<grid, 256, 0, h->stream>>>(m, n, k, (const half_t*)a, lda,
                                                   (const half_t*)b, ldb, (half_t)alpha,
                                                   (half_t)beta, (half_t*)c, ldc);
      timer_end(h);
      RF_HIP(hipGetLastError());
      return 0;
    }
    default: return -2;
  }
  if (out_t != REMIFA_R_32F && out_t != REMIFA_R_16F) return -5;

  GemmArgs g{};
  g.M = m;
  g.N = n;
  g.K = k > 0 ? k : 0;
  g.lda = lda;
  g.ldb = ldb;
  g.ldc = ldc;
  g.alpha = alpha;
  g.beta = beta;
  g.A = a;
  g.B = b;
  g.C = c;
  g.Cin = cin;
  g.ldcin = ldcin;
  g.cin_half = cin_t == REMIFA_R_16F;
  const size_t es = (opk == OPK_F32) ? 4 : 2;
  const int chunk = 16 / (int)es;
  // whole 16-byte chunks: aligned operands, M and K multiples of the chunk
  const bool vec = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) && (lda % chunk == 0) &&
                   (ldb % chunk == 0) && (m % chunk == 0) && (g.K % chunk == 0);
  const int bk = kSplitGranule;
  int tile, split;
  plan_gemm(h, opk, m, n, g.K, a, lda, b, ldb, tile, split);
  g.kchunk = ceil_div(ceil_div(g.K > 0 ? g.K : 1, split), bk) * bk;
  split = (int)ceil_div(g.K > 0 ? g.K : 1, g.kchunk);
  if (split > 1) {
    // partials go to the part of the workspace a running driver has not
    // reserved; a standalone call may grow it, otherwise split less
    auto need = [&]() { return sizeof(float) * (size_t)split * (size_t)n * (size_t)m; };
    size_t avail = h->ws_bytes > h->ws_used ? h->ws_bytes - h->ws_used : 0;
    if (need() > avail && h->ws_used == 0 && !h->ws_user) {
      RF_CHECK(ensure_workspace(h, need()));
      avail = h->ws_bytes;
    }
    while (split > 1 && need() > avail) {
      split = (split + 1) / 2;
      g.kchunk = ceil_div(ceil_div(g.K, split), bk) * bk;
      split = (int)ceil_div(g.K, g.kchunk);
    }
    g.W = (float*)((char*)h->ws + h->ws_used);
    g.ldw = m;
  }

  timer_begin(h, KC_GEMM, flops, (opk << 8) | ((out_t == REMIFA_R_16F) << 4) | tile);
  int st;
  if (tile == T256x256) {
    if (opk == OPK_F16)
      st = out_t == REMIFA_R_32F ? launch_big<OPK_F16, float>(h, g, split)
                                 : launch_big<OPK_F16, half_t>(h, g, split);
    else
      st = out_t == REMIFA_R_32F ? launch_big<OPK_BF16S, float>(h, g, split)
                                 : launch_big<OPK_BF16S, half_t>(h, g, split);
  } else if (opk == OPK_F32) {
    st = out_t == REMIFA_R_32F ? dispatch_tiles<OPK_F32, float>(h, tile, g, split, vec)
                               : dispatch_tiles<OPK_F32, half_t>(h, tile, g, split, vec);
  } else if (opk == OPK_F16) {
    st = out_t == REMIFA_R_32F ? dispatch_tiles<OPK_F16, float>(h, tile, g, split, vec)
                               : dispatch_tiles<OPK_F16, half_t>(h, tile, g, split, vec);
  } else if (opk == OPK_BF16) {
    st = out_t == REMIFA_R_32F ? dispatch_tiles<OPK_BF16, float>(h, tile, g, split, vec)
                               : dispatch_tiles<OPK_BF16, half_t>(h, tile, g, split, vec);
  } else {
    // bf16 storage (the LU's bf16 mirror of its factors): the three basic tiles
    const int t = (tile == T256x32 || tile == T32x256) ? tile : T128x128;
    st = out_t == REMIFA_R_32F ? dispatch_basic<OPK_BF16S, float>(h, t, g, split, vec)
                               : dispatch_basic<OPK_BF16S, half_t>(h, t, g, split, vec);
  }
  timer_end(h);
  if (st == 0 && split > 1) {
    // the reduction is timed as its own variant (tile code 15)
    timer_begin(h, KC_GEMM, 0.0, (opk << 8) | ((out_t == REMIFA_R_16F) << 4) | 15);
    st = splitk_reduce_launch(h, out_t, m, n, split, g.W, g.ldw, beta, c, ldc, g.Cin, g.ldcin,
                              g.cin_half);
    timer_end(h);
  }
  return st;
}

}  // namespace remifa
