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
//    16-byte LDS read, A fragments are gathered element-wise (conflict-free
//    thanks to the row padding).
//  * Workgroup ids are remapped so that each of the 8 XCDs (which own
//    private L2s) processes a contiguous range of tiles.
//  * Optional split-K: slice z of K writes alpha*partial to a workspace and a
//    reduction kernel adds the slices in order (deterministic).
//
// fp16-accumulating compute types (FP16, TC16: cublasHgemm / GemmEx 16F) run
// on a VALU kernel with one correctly rounded fp16 fma per k, in k order.
#include "common.hpp"

namespace remifa {
namespace {

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

enum { OPK_F16 = 0, OPK_BF16 = 1, OPK_F32 = 2 };

struct GemmArgs {
  int64_t M, N, K;
  int64_t lda, ldb, ldc;
  float alpha, beta;
  const void* A;
  const void* B;
  void* C;
  float* W;        // split-K partials [split][N][ldw]
  int64_t ldw;
  int64_t kchunk;  // K range per split (multiple of BK)
  int tiles_m, tiles_n;
};

// --- per operand kind traits --------------------------------------------
template <int OPK>
struct OpTraits;
template <>
struct OpTraits<OPK_F16> {
  using Store = half_t;     // storage type in HBM
  using Lds = uint16_t;     // element in LDS
  static constexpr int BK = 32;
  static constexpr int CH = 8;  // elements per 16-byte chunk
  using Chunk = u16x8;
  static __device__ __forceinline__ Chunk cvt(Chunk c) { return c; }
};
template <>
struct OpTraits<OPK_BF16> {
  using Store = half_t;
  using Lds = uint16_t;
  static constexpr int BK = 32;
  static constexpr int CH = 8;
  using Chunk = u16x8;
  // fp16 storage -> bf16 operand, round to nearest even
  static __device__ __forceinline__ Chunk cvt(Chunk c) {
    Chunk r;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      float f = (float)__builtin_bit_cast(half_t, (uint16_t)c[v]);
      r[v] = f32_to_bf16_bits(f);
    }
    return r;
  }
};
template <>
struct OpTraits<OPK_F32> {
  using Store = float;
  using Lds = float;
  static constexpr int BK = 16;
  static constexpr int CH = 4;
  using Chunk = f32x4;
  static __device__ __forceinline__ Chunk cvt(Chunk c) { return c; }
};

// XCD-aware tile index: workgroup w runs on XCD w % 8; give every XCD a
// contiguous range of tiles (a bijection for any tile count).
__device__ __forceinline__ int xcd_tile(int w, int nt) {
  int q = nt >> 3, r = nt & 7;
  int x = w & 7, s = w >> 3;
  return x * q + (x < r ? x : r) + s;
}

template <typename OutT>
__device__ __forceinline__ float ld_out(const OutT* p) {
  return (float)*p;
}

template <int OPK, typename OutT, int BM, int BN, int WM, int WN, bool VEC>
__global__ void __launch_bounds__(256) gemm_mfma_kernel(GemmArgs g) {
  using TR = OpTraits<OPK>;
  using S = typename TR::Store;
  using L = typename TR::Lds;
  using Chunk = typename TR::Chunk;
  constexpr int BK = TR::BK;
  constexpr int CH = TR::CH;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / 32, FN = WTN / 32;
  constexpr int PADA = (OPK == OPK_F32) ? 32 : 8;
  constexpr int PADB = (OPK == OPK_F32) ? 4 : 8;
  constexpr int SA = BM + PADA;  // As[k][m]
  constexpr int SB = BK + PADB;  // Bs[n][k]
  constexpr int TA = BK * BM / CH, TB = BN * BK / CH;  // 16-byte chunks per slab
  constexpr int NA = (TA + 255) / 256;
  constexpr int NB = (TB + 255) / 256;
  static_assert(TA % 64 == 0 && TB % 64 == 0, "slab chunks must fill whole waves");
  static_assert(FM >= 1 && FN >= 1, "wave tile below 32x32");

  __shared__ __attribute__((aligned(16))) L As[2][BK * SA];
  __shared__ __attribute__((aligned(16))) L Bs[2][BN * SB];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;

  const int nt = g.tiles_m * g.tiles_n;
  const int tile = xcd_tile(blockIdx.x, nt);
  const int tm = tile % g.tiles_m, tn = tile / g.tiles_m;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * g.kchunk;
  int64_t kend = kbeg + g.kchunk;
  if (kend > g.K) kend = g.K;
  const int nk = (int)((kend - kbeg + BK - 1) / BK);

  const S* __restrict__ A = (const S*)g.A;
  const S* __restrict__ B = (const S*)g.B;

  Chunk ra[NA], rb[NB];

  auto load_tiles = [&](int kt) {
    const int64_t kk0 = kbeg + (int64_t)kt * BK;
#pragma unroll
    for (int p = 0; p < NA; ++p) {
      int c = tid + 256 * p;
      if (TA % 256 && c >= TA) break;
      int mq = c % (BM / CH), k = c / (BM / CH);
      int64_t gm = m0 + mq * CH, gk = kk0 + k;
      Chunk v;
      const S* src = A + gm + gk * g.lda;
      if (VEC && gk < kend && gm + CH <= g.M) {
        v = *(const Chunk*)src;
      } else {
#pragma unroll
        for (int e = 0; e < CH; ++e) {
          S x = (gk < kend && gm + e < g.M) ? src[e] : (S)0;
          if constexpr (OPK == OPK_F32)
            v[e] = x;
          else
            v[e] = __builtin_bit_cast(uint16_t, x);
        }
      }
      ra[p] = v;
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      int c = tid + 256 * p;
      if (TB % 256 && c >= TB) break;
      int kq = c % (BK / CH), n = c / (BK / CH);
      int64_t gk = kk0 + kq * CH, gn = n0 + n;
      Chunk v;
      const S* src = B + gk + gn * g.ldb;
      if (VEC && gn < g.N && gk + CH <= kend) {
        v = *(const Chunk*)src;
      } else {
#pragma unroll
        for (int e = 0; e < CH; ++e) {
          S x = (gn < g.N && gk + e < kend) ? src[e] : (S)0;
          if constexpr (OPK == OPK_F32)
            v[e] = x;
          else
            v[e] = __builtin_bit_cast(uint16_t, x);
        }
      }
      rb[p] = v;
    }
  };

  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int p = 0; p < NA; ++p) {
      int c = tid + 256 * p;
      if (TA % 256 && c >= TA) break;
      int mq = c % (BM / CH), k = c / (BM / CH);
      *(Chunk*)&As[buf][k * SA + mq * CH] = TR::cvt(ra[p]);
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      int c = tid + 256 * p;
      if (TB % 256 && c >= TB) break;
      int kq = c % (BK / CH), n = c / (BK / CH);
      *(Chunk*)&Bs[buf][n * SB + kq * CH] = TR::cvt(rb[p]);
    }
  };

  f16v acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;

  if (nk > 0) {
    load_tiles(0);
    store_tiles(0);
  }
  __syncthreads();

  const int l32 = lane & 31, lh = lane >> 5;
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nk) load_tiles(kt + 1);
    const L* as = As[buf];
    const L* bs = Bs[buf];
    if constexpr (OPK == OPK_F32) {
#pragma unroll
      for (int ks = 0; ks < BK; ks += 2) {
        float xb[FN], ya[FM];
#pragma unroll
        for (int j = 0; j < FN; ++j) xb[j] = bs[(wn * WTN + j * 32 + l32) * SB + ks + lh];
#pragma unroll
        for (int i = 0; i < FM; ++i) ya[i] = as[(ks + lh) * SA + wm * WTM + i * 32 + l32];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(xb[j], ya[i], acc[i][j], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < BK; ks += 16) {
        u16x8 xb[FN], ya[FM];
#pragma unroll
        for (int j = 0; j < FN; ++j)
          xb[j] = *(const u16x8*)&bs[(wn * WTN + j * 32 + l32) * SB + ks + 8 * lh];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
#pragma unroll
          for (int v = 0; v < 8; ++v)
            ya[i][v] = as[(ks + 8 * lh + v) * SA + wm * WTM + i * 32 + l32];
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            if constexpr (OPK == OPK_F16)
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                  __builtin_bit_cast(h8, xb[j]), __builtin_bit_cast(h8, ya[i]), acc[i][j], 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                  __builtin_bit_cast(bf8, xb[j]), __builtin_bit_cast(bf8, ya[i]), acc[i][j], 0, 0,
                  0);
          }
      }
    }
    if (kt + 1 < nk) store_tiles(buf ^ 1);
    __syncthreads();
  }

  // epilogue: lane holds C[m = .. + l32][n = .. + 8*(q/4) + 4*lh + q%4]
  const bool split = gridDim.z > 1;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int64_t m = m0 + wm * WTM + i * 32 + l32;
    if (m >= g.M) continue;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int64_t n = n0 + wn * WTN + j * 32 + 8 * (q >> 2) + 4 * lh + (q & 3);
        if (n >= g.N) continue;
        float v = acc[i][j][q];
        if (split) {
          g.W[((int64_t)blockIdx.z * g.N + n) * g.ldw + m] = g.alpha * v;
        } else {
          OutT* cp = (OutT*)g.C + m + n * g.ldc;
          float r = (g.beta != 0.0f) ? __builtin_fmaf(g.alpha, v, g.beta * ld_out(cp)) : g.alpha * v;
          *cp = (OutT)r;
        }
      }
    }
  }
}

// C = beta*C + sum_z W[z] (slices added in order z = 0, 1, ...)
template <typename OutT>
__global__ void __launch_bounds__(256) splitk_reduce_kernel(int64_t M, int64_t N, int split,
                                                            const float* __restrict__ W,
                                                            int64_t ldw, float beta,
                                                            OutT* __restrict__ C, int64_t ldc) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= M) return;
  for (int64_t j = blockIdx.y; j < N; j += gridDim.y) {
    float s = 0.0f;
    for (int z = 0; z < split; ++z) s += W[(z * N + j) * ldw + i];
    OutT* cp = C + i + j * ldc;
    float r = (beta != 0.0f) ? s + beta * (float)*cp : s;
    *cp = (OutT)r;
  }
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

template <int OPK, typename OutT, int BM, int BN, int WM, int WN>
int launch_mfma(remifa_handle_t h, GemmArgs g, int split, bool vec) {
  g.tiles_m = (int)ceil_div(g.M, BM);
  g.tiles_n = (int)ceil_div(g.N, BN);
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), 1, (unsigned)split);
  if (vec)
    gemm_mfma_kernel<OPK, OutT, BM, BN, WM, WN, true><<<grid, 256, 0, h->stream>>>(g);
  else
    gemm_mfma_kernel<OPK, OutT, BM, BN, WM, WN, false><<<grid, 256, 0, h->stream>>>(g);
  RF_HIP(hipGetLastError());
  return 0;
}

// tile shapes (BM x BN); the timer's variant code of a launch is
// opk << 8 | fp16-output << 4 | tile
enum Tile { T256x32 = 0, T32x256 = 1, T256x64 = 2, T64x256 = 3, T256x256 = 4, T128x128 = 5 };

// 256x256 for large outputs; narrow shapes (N or M = ib .. 64) use 256x32 /
// 256x64 and their transposes so that skinny updates keep the chip busy.
int choose_tile(int opk, int64_t M, int64_t N) {
  if (N <= 32) return T256x32;
  if (M <= 32) return T32x256;
  if (opk == OPK_F32) return T128x128;
  if (N <= 64) return T256x64;
  if (M <= 64) return T64x256;
  if (M * N >= (int64_t)256 * 256 * 512) return T256x256;
  return T128x128;
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
      case T256x256: return launch_mfma<OPK, OutT, 256, 256, 2, 2>(h, g, split, vec);
    }
  }
  return REMIFA_ERR_UNSUPPORTED;
}

}  // namespace

int gemm_launch(remifa_handle_t h, int ctype, int in_t, int out_t, int64_t m, int64_t n, int64_t k,
                float alpha, const void* a, int64_t lda, const void* b, int64_t ldb, float beta,
                void* c, int64_t ldc) {
  if (m <= 0 || n <= 0) return 0;
  const double flops = 2.0 * (double)m * (double)n * (double)(k > 0 ? k : 0);
  if (k <= 0 || alpha == 0.0f) {
    if (beta == 1.0f) return 0;
    // C = beta * C: run the GEMM with K = 0 (the epilogue scales)
  }
  int opk;
  switch (ctype) {
    case REMIFA_FP32: opk = OPK_F32; if (in_t != REMIFA_R_32F) return -4; break;
    case REMIFA_TC32: opk = OPK_F16; if (in_t != REMIFA_R_16F) return -4; break;
    case REMIFA_BF16_TC32: opk = OPK_BF16; if (in_t != REMIFA_R_16F) return -4; break;
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
  const size_t es = (opk == OPK_F32) ? 4 : 2;
  const int chunk = 16 / (int)es;
  const bool vec = ((uintptr_t)a % 16 == 0) && ((uintptr_t)b % 16 == 0) && (lda % chunk == 0) &&
                   (ldb % chunk == 0);
  const int bk = (opk == OPK_F32) ? 16 : 32;

  // split-K when the output tiles cannot fill the chip and K is long
  int split = h->split_k;
  if (split <= 0) {
    int64_t tiles = ceil_div(m, 128) * ceil_div(n, 128);
    split = 1;
    while (tiles * split < 2 * (int64_t)h->num_cu && g.K / (split * 2) >= 1024 && split < 16)
      split *= 2;
  }
  if (g.K == 0) split = 1;
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

  const int tile = choose_tile(opk, m, n);
  timer_begin(h, KC_GEMM, flops, (opk << 8) | ((out_t == REMIFA_R_16F) << 4) | tile);
  int st;
  if (opk == OPK_F32) {
    st = out_t == REMIFA_R_32F ? dispatch_tiles<OPK_F32, float>(h, tile, g, split, vec)
                               : dispatch_tiles<OPK_F32, half_t>(h, tile, g, split, vec);
  } else if (opk == OPK_F16) {
    st = out_t == REMIFA_R_32F ? dispatch_tiles<OPK_F16, float>(h, tile, g, split, vec)
                               : dispatch_tiles<OPK_F16, half_t>(h, tile, g, split, vec);
  } else {
    st = out_t == REMIFA_R_32F ? dispatch_tiles<OPK_BF16, float>(h, tile, g, split, vec)
                               : dispatch_tiles<OPK_BF16, half_t>(h, tile, g, split, vec);
  }
  if (st == 0 && split > 1) {
    dim3 grid((unsigned)ceil_div(m, 256), (unsigned)(n < 65535 ? n : 65535));
    if (out_t == REMIFA_R_32F)
      splitk_reduce_kernel<float><<<grid, 256, 0, h->stream>>>(m, n, split, g.W, g.ldw, beta,
                                                               (float*)c, ldc);
    else
      splitk_reduce_kernel<half_t><<<grid, 256, 0, h->stream>>>(m, n, split, g.W, g.ldw, beta,
                                                                (half_t*)c, ldc);
    if (hipGetLastError() != hipSuccess) st = REMIFA_ERR_HIP;
  }
  timer_end(h);
  return st;
}

}  // namespace remifa
