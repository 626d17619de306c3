// Shared by the slice kernels (subpanel.hip) and the block kernel
// (blockll.hip): the arithmetic of the inner factorization and the
// factorization of a 64 x 64 diagonal tile held in registers.
#pragma once
#include <type_traits>

#include "common.hpp"

namespace remifa {
namespace {

constexpr int SW = 64;  // tile edge (slices are at most 64 wide)
constexpr int PD = 4;   // LDS row padding of the tile
constexpr int G = 8;    // vector entries held in registers at a time

// published tile: D, D^T, op(D), op(D)^T (64 x 64 row-major each), 1/u_kk, ones
constexpr int T_D = 0, T_DT = 4096, T_DOP = 8192, T_DOPT = 12288, T_R = 16384, T_ONE = 16448;
constexpr int T_FLOATS = 16512;

enum SubMode {
  SM_F32 = 0,     // fp32 data, fp32 updates (I = FP32)
  SM_F32_TC = 1,  // fp32 data, fp16 operands with fp32 accumulation (I = TC32)
  SM_F16_TC = 2,  // fp16 data, fp16 arithmetic in the panel, TC32 updates
  SM_F16_H = 3,   // fp16 data, fp16-accumulated updates (I = FP16 / TC16)
  SM_F16_BF = 4   // fp16 data, bf16 operands with fp32 accumulation
};

__device__ __forceinline__ float r16(float x) { return (float)(half_t)x; }
__device__ __forceinline__ float rbf(float x) { return bf16_bits_to_f32(f32_to_bf16_bits(x)); }

template <int MODE>
struct Sem {
  static constexpr bool HALF = MODE >= SM_F16_TC;  // fp16 storage and panel arithmetic
  // operand rounding of the cross-block updates
  static __device__ __forceinline__ float op(float x) {
    if constexpr (MODE == SM_F32_TC) return r16(x);
    if constexpr (MODE == SM_F16_BF) return rbf(x);
    return x;  // fp32, or already fp16-valued
  }
  // panel arithmetic
  static __device__ __forceinline__ float rcp(float x) {
    float r = 1.0f / x;
    return HALF ? r16(r) : r;
  }
  static __device__ __forceinline__ float mul(float a, float b) {
    float r = a * b;  // exact for fp16 inputs
    return HALF ? r16(r) : r;
  }
  static __device__ __forceinline__ float fnma(float a, float b, float c) {
    if constexpr (HALF)
      return (float)__builtin_fmaf16(-(half_t)a, (half_t)b, (half_t)c);
    else
      return __builtin_fmaf(-a, b, c);
  }
  // cross-block product sums: operands are already op-rounded
  static __device__ __forceinline__ float acc_add(float acc, float a, float b) {
    if constexpr (MODE == SM_F16_H)
      return (float)__builtin_fmaf16((half_t)a, (half_t)b, (half_t)acc);
    else
      return __builtin_fmaf(a, b, acc);
  }
  // c - acc, rounded to the panel precision
  static __device__ __forceinline__ float apply(float c, float acc) {
    if constexpr (MODE == SM_F16_H) return (float)((half_t)c - (half_t)acc);
    float r = c - acc;
    return HALF ? r16(r) : r;
  }
};

// ---------------------------------------------------------------------------
// 1. the diagonal tile, blocked by 8-column panels: eight wavefronts, lane i
// holds row i, wavefront w owns columns 8w..8w+7 in registers. Stage g:
// wavefront g factors its panel (8 pivots, the pivot row broadcast with
// readlane inside the wavefront) and shares the multipliers through LDS; after one barrier every
// wavefront right of it applies the panel's 8 pivots to its own columns in
// pivot order. Every entry receives the same sequence of fused updates as in
// a pivot-by-pivot elimination. ib (a multiple of 8) decides where the block
// boundaries fall: at a block boundary the block sums (LL: before the block,
// RL: after it) run as runtime loops over LDS copies of the operands.
constexpr int TP = 8;  // panel width = columns per wavefront

// Row pitches are 16-byte multiples so that a lane publishes its 8 values
// with two 16-byte stores; 48-, 144- and 272-byte pitches keep 16
// consecutive lanes on distinct banks. The op-rounded copies feed only the
// block sums, which read columns / rows below the last block's start: KM =
// SW - ib of them are kept.
template <int KM>
struct __attribute__((aligned(16))) TileLdsT {
  __attribute__((aligned(16))) float Lp[2][SW][TP + 4];  // multipliers of the current panel
  __attribute__((aligned(16))) float Uo[KM][SW + PD];    // op-rounded rows of U (rows < KM)
  __attribute__((aligned(16))) float Xo[SW][KM + PD];    // op-rounded multipliers (columns < KM)
  float R[SW];                                           // reciprocals of the pivots
};

__device__ __forceinline__ void st8(float* p, const float* v) {
  *(float4*)p = make_float4(v[0], v[1], v[2], v[3]);
  *(float4*)(p + 4) = make_float4(v[4], v[5], v[6], v[7]);
}

// value of v in lane `lane` (wave-uniform) for every lane of the wavefront
__device__ __forceinline__ float bcast_lane(float v, int lane) {
  return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}

// The tile factorization proper, on eight wavefronts: lane i of wavefront w
// holds x[jj] = T[i][8 w + jj]; on return the same registers hold the
// factors (multipliers below the diagonal, U on and above it) and S.R the
// reciprocals of the pivots. ib is the block size (8, 16, 32 or 64, with
// SW - ib <= KM).
template <int MODE, bool RL, int KM>
__device__ __forceinline__ void tile_factor(TileLdsT<KM>& S, float (&x)[TP], int ib, int w, int i) {
  using Sm = Sem<MODE>;
  const int cb = w * TP;
  for (int g = 0; g < SW / TP; ++g) {
    const int p0 = g * TP;         // first pivot of the panel
    const int kb = p0 - p0 % ib;   // ib block holding it: [kb, kend)
    const int kend = kb + ib;
    // (every wavefront takes every stage: the barrier below needs all eight;
    // columns left of the panel are final and take no work)
    if (!RL && p0 == kb && kb > 0 && w >= g) {
      // left-looking update at the start of a block: its columns for the
      // rows >= kb, every column right of it for the rows of the block;
      // products summed from zero in k order, subtracted once
      float acc[TP];
#pragma unroll
      for (int jj = 0; jj < TP; ++jj) acc[jj] = 0.0f;
#pragma unroll 2
      for (int j = 0; j < kb; ++j) {
        const float xj = S.Xo[i][j];
        const float* ur = &S.Uo[j][cb];
#pragma unroll
        for (int jj = 0; jj < TP; ++jj) acc[jj] = Sm::acc_add(acc[jj], xj, ur[jj]);
      }
      if (i >= kb && (cb < kend || i < kend)) {
#pragma unroll
        for (int jj = 0; jj < TP; ++jj) x[jj] = Sm::apply(x[jj], acc[jj]);
      }
    }
    if (w == g) {
      // factor the panel: pivot by pivot inside this wavefront; lane k's
      // row is read by every lane with readlane (no LDS round trip). Only
      // the arithmetic is on this chain: the panel is published to LDS
      // afterwards, by all lanes at once.
#pragma unroll
      for (int kk = 0; kk < TP; ++kk) {
        const int k = p0 + kk;
        float u[TP];
#pragma unroll
        for (int jj = kk; jj < TP; ++jj) u[jj] = bcast_lane(x[jj], k);
        const float r = Sm::rcp(u[kk]);
        if (i > k) {
          const float lk = Sm::mul(x[kk], r);
          x[kk] = lk;
#pragma unroll
          for (int jj = kk + 1; jj < TP; ++jj) x[jj] = Sm::fnma(lk, u[jj], x[jj]);
        }
      }
      // multipliers (rows below each pivot; the other entries are not read)
      float o[TP];
#pragma unroll
      for (int kk = 0; kk < TP; ++kk) o[kk] = Sm::op(x[kk]);
      st8(&S.Lp[g & 1][i][0], x);
      if (p0 < KM) st8(&S.Xo[i][p0], o);
      if (i >= p0 && i < p0 + TP) {
        // this lane's row of U in the panel and its pivot's reciprocal
        if (p0 < KM) st8(&S.Uo[i][cb], o);
        float d = x[0];
#pragma unroll
        for (int kk = 1; kk < TP; ++kk) d = (i - p0 == kk) ? x[kk] : d;
        S.R[i] = Sm::rcp(d);
      }
    }
    __syncthreads();
    if (w > g) {
      // the panel's pivots on this wavefront's columns, in pivot order: for
      // every row inside the block, for the rows below it only within it
      const bool upd = i < kend || cb < kend;
      float lp[TP];
#pragma unroll
      for (int kk = 0; kk < TP; ++kk) lp[kk] = S.Lp[g & 1][i][kk];
#pragma unroll
      for (int kk = 0; kk < TP; ++kk) {
        const int k = p0 + kk;
        float u[TP];
#pragma unroll
        for (int jj = 0; jj < TP; ++jj) u[jj] = bcast_lane(x[jj], k);
        if (i > k && upd) {
#pragma unroll
          for (int jj = 0; jj < TP; ++jj) x[jj] = Sm::fnma(lp[kk], u[jj], x[jj]);
        }
      }
      // rows p0 .. p0 + 7 of these columns are final: publish them
      if (p0 < KM && i >= p0 && i < p0 + TP) {
        float o[TP];
#pragma unroll
        for (int jj = 0; jj < TP; ++jj) o[jj] = Sm::op(x[jj]);
        st8(&S.Uo[i][cb], o);
      }
      if (RL && p0 + TP == kend && cb >= kend) {
        // block complete: right-looking rank-ib update of the trailing part
        float acc[TP];
#pragma unroll
        for (int jj = 0; jj < TP; ++jj) acc[jj] = 0.0f;
#pragma unroll 2
        for (int q = kb; q < kend; ++q) {
          const float xq = S.Xo[i][q];
          const float* ur = &S.Uo[q][cb];
#pragma unroll
          for (int jj = 0; jj < TP; ++jj) acc[jj] = Sm::acc_add(acc[jj], xq, ur[jj]);
        }
        if (i >= kend) {
#pragma unroll
          for (int jj = 0; jj < TP; ++jj) x[jj] = Sm::apply(x[jj], acc[jj]);
        }
      }
    }
  }

}

}  // namespace
}  // namespace remifa
