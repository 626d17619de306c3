// Fused sub-panel factorization: the inner ib loop of the reference drivers
// (src/lu_nopiv_inner_ll.hxx:197-370, src/lu_nopiv_inner_rl.hxx:104-196)
// for an sw-wide slice (sw <= 64) of the outer block, in one launch.
//
// Input: the sw-wide block column l[0:cm, 0:sw] (its top sw x sw tile is the
// diagonal tile) and the sw-tall block row u[0:sw, sw:cm], already updated by
// everything left of / above the slice. Output: the factors, in place, and
// optionally their fp16 mirror cp (the reference's d_a_cpy).
//
// Every workgroup (256 lanes) factors the diagonal tile redundantly in LDS,
// then handles 256 rows of L (one row per lane, sw values in registers) or
// 256 columns of U (one column per lane). Inside the slice the work follows
// the reference's ib-granular semantics:
//  * within an ib block (the reference's panel kernel, src/lu_nopiv.cu:123-
//    170): reciprocal of the pivot, multiply, fused multiply-subtract, in the
//    factor precision IF (fp32, or fp16 with one rounding per operation);
//  * across ib blocks (the reference's inner GEMMs): products of operands
//    rounded to the update compute type (fp16 / bf16 for TC types), summed
//    from zero (fp32, or fp16 for FP16/TC16), subtracted once and rounded to
//    the panel precision IP -- left-looking (LL: one update per column block)
//    or right-looking (RL: a rank-ib update after each block).
// The diagonal tile is read by every workgroup; the last workgroup to have
// read it (an agent-scope counter) writes it back and re-arms the counter,
// so no snapshot copy is needed.
#include "common.hpp"

namespace remifa {
namespace {

constexpr int SW = 64;  // widest sub-panel

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
  // c - sum_j a_j b_j over the cross-block operands, rounded to IP
  struct Acc {
    float f = 0.0f;
    half_t h = (half_t)0.0f;
    __device__ __forceinline__ void add(float a, float b) {
      if constexpr (MODE == SM_F16_H)
        h = __builtin_fmaf16((half_t)a, (half_t)b, h);
      else
        f = __builtin_fmaf(op(a), op(b), f);
    }
    __device__ __forceinline__ float apply(float c) const {
      if constexpr (MODE == SM_F16_H) return (float)((half_t)c - h);
      float r = c - f;
      return HALF ? r16(r) : r;
    }
  };
};

template <typename T, int IB, int MODE, bool RL>
__global__ void __launch_bounds__(256) subpanel_kernel(int64_t cm, int sw, T* __restrict__ l,
                                                       int64_t ldl, T* __restrict__ u, int64_t ldu,
                                                       half_t* __restrict__ cp, int64_t ldcp,
                                                       int* counter, int nl_wg) {
  using Sm = Sem<MODE>;
  using Acc = typename Sm::Acc;
  constexpr int NBLK = SW / IB;
  __shared__ float D[SW][SW + 1];  // D[i][j]
  __shared__ float R[SW];          // reciprocals of the pivots
  __shared__ int s_last;
  const int t = threadIdx.x;

  // ---- load the diagonal tile (lane t < sw owns row t)
  if (t < sw) {
#pragma unroll
    for (int j = 0; j < SW; ++j)
      if (j < sw) D[t][j] = (float)l[t + (int64_t)j * ldl];
  }
  __syncthreads();
  if (t == 0) {
    // every workgroup has read the tile before the last one rewrites it
    int old = __hip_atomic_fetch_add(counter, 1, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (old == (int)gridDim.x - 1);
  }

  // ---- factor the diagonal tile, block by block: row t % 64 is shared by
  // four lanes, lane quarter q = t / 64 takes the columns c = q (mod 4)
  const bool own = t < sw;
  const int drow = t & 63, dq = t >> 6;
  const bool down = drow < sw;
  for (int k0 = 0; k0 < sw; k0 += IB) {
    const int kend = k0 + IB < sw ? k0 + IB : sw;
    if (!RL && k0 > 0) {
      // left-looking cross-block update: columns of block kb for the rows
      // >= k0, and the columns right of the block for the rows of the block
      if (down && drow >= k0) {
        const int cend = (drow < kend) ? sw : kend;
        for (int c = k0 + dq; c < cend; c += 4) {
          Acc acc;
          for (int j = 0; j < k0; ++j) acc.add(D[drow][j], D[j][c]);
          D[drow][c] = acc.apply(D[drow][c]);
        }
      }
      __syncthreads();
    }
    for (int k = k0; k < kend; ++k) {
      const float r = Sm::rcp(D[k][k]);
      if (t == 0) R[k] = r;
      if (down && drow > k && dq == 0) D[drow][k] = Sm::mul(D[drow][k], r);
      __syncthreads();
      if (down && drow > k) {
        // within the block for every row; the rows of the block also carry
        // on to the right of it (their U part)
        const float ltk = D[drow][k];
        const int jend = (drow < kend) ? sw : kend;
        for (int j = k + 1 + dq; j < jend; j += 4) D[drow][j] = Sm::fnma(ltk, D[k][j], D[drow][j]);
      }
      __syncthreads();
    }
    if (RL && kend < sw) {
      // right-looking rank-ib update of the trailing part of the tile
      if (down && drow >= kend) {
        for (int c = kend + dq; c < sw; c += 4) {
          Acc acc;
          for (int k = k0; k < kend; ++k) acc.add(D[drow][k], D[k][c]);
          D[drow][c] = acc.apply(D[drow][c]);
        }
      }
      __syncthreads();
    }
  }

  // ---- this workgroup's rows of L or columns of U
  const int wg = blockIdx.x;
  if (wg < nl_wg) {
    const int64_t row = sw + (int64_t)wg * 256 + t;
    if (row < cm) {
      float x[SW];
#pragma unroll
      for (int j = 0; j < SW; ++j) x[j] = (j < sw) ? (float)l[row + (int64_t)j * ldl] : 0.0f;
#pragma unroll
      for (int kb = 0; kb < NBLK; ++kb) {
        const int k0 = kb * IB;
        if (k0 < sw) {
          if (!RL && k0 > 0) {
#pragma unroll
            for (int c = 0; c < IB; ++c) {
              if (k0 + c < sw) {
                Acc acc;
#pragma unroll
                for (int j = 0; j < SW; ++j)
                  if (j < k0) acc.add(x[j], D[j][k0 + c]);
                x[k0 + c] = acc.apply(x[k0 + c]);
              }
            }
          }
#pragma unroll
          for (int kk = 0; kk < IB; ++kk) {
            const int k = k0 + kk;
            if (k < sw) {
              const float lk = Sm::mul(x[k], R[k]);
              x[k] = lk;
#pragma unroll
              for (int j = 0; j < SW; ++j)
                if (j > k && j < k0 + IB && j < sw) x[j] = Sm::fnma(lk, D[k][j], x[j]);
            }
          }
          if (RL && k0 + IB < sw) {
#pragma unroll
            for (int c = 0; c < SW; ++c) {
              if (c >= k0 + IB && c < sw) {
                Acc acc;
#pragma unroll
                for (int kk = 0; kk < IB; ++kk) acc.add(x[k0 + kk], D[k0 + kk][c]);
                x[c] = acc.apply(x[c]);
              }
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < SW; ++j) {
        if (j < sw) {
          l[row + (int64_t)j * ldl] = (T)x[j];
          if (cp) cp[row + (int64_t)j * ldcp] = (half_t)x[j];
        }
      }
    }
  } else {
    const int64_t col = sw + (int64_t)(wg - nl_wg) * 256 + t;
    if (col < cm) {
      float v[SW];
#pragma unroll
      for (int i = 0; i < SW; ++i) v[i] = (i < sw) ? (float)u[i + col * ldu] : 0.0f;
#pragma unroll
      for (int kb = 0; kb < NBLK; ++kb) {
        const int k0 = kb * IB;
        if (k0 < sw) {
          if (!RL && k0 > 0) {
#pragma unroll
            for (int r = 0; r < IB; ++r) {
              if (k0 + r < sw) {
                Acc acc;
#pragma unroll
                for (int j = 0; j < SW; ++j)
                  if (j < k0) acc.add(D[k0 + r][j], v[j]);
                v[k0 + r] = acc.apply(v[k0 + r]);
              }
            }
          }
#pragma unroll
          for (int kk = 0; kk < IB; ++kk) {
            const int k = k0 + kk;
            if (k < sw) {
              const float vk = v[k];
#pragma unroll
              for (int i = 0; i < SW; ++i)
                if (i > k && i < k0 + IB && i < sw) v[i] = Sm::fnma(D[i][k], vk, v[i]);
            }
          }
          if (RL && k0 + IB < sw) {
#pragma unroll
            for (int i = 0; i < SW; ++i) {
              if (i >= k0 + IB && i < sw) {
                Acc acc;
#pragma unroll
                for (int kk = 0; kk < IB; ++kk) acc.add(D[i][k0 + kk], v[k0 + kk]);
                v[i] = acc.apply(v[i]);
              }
            }
          }
        }
      }
#pragma unroll
      for (int i = 0; i < SW; ++i) {
        if (i < sw) {
          u[i + col * ldu] = (T)v[i];
          if (cp) cp[i + col * ldcp] = (half_t)v[i];
        }
      }
    }
  }

  // ---- the last reader of the diagonal tile stores it and re-arms the counter
  __syncthreads();
  if (s_last) {
    if (own) {
#pragma unroll
      for (int j = 0; j < SW; ++j) {
        if (j < sw) {
          l[t + (int64_t)j * ldl] = (T)D[t][j];
          if (cp) cp[t + (int64_t)j * ldcp] = (half_t)D[t][j];
        }
      }
    }
    if (t == 0) __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

template <typename T, int MODE, bool RL>
int launch_ib(remifa_handle_t h, int ib, int64_t cm, int sw, void* l, int64_t ldl, void* u,
              int64_t ldu, half_t* cp, int64_t ldcp, int* counter) {
  const int64_t rest = cm - sw;
  const int nl = (int)ceil_div(rest > 0 ? rest : 0, 256);
  const int nu = nl;
  const int grid = (nl + nu) > 0 ? nl + nu : 1;
  const int nlw = nl > 0 ? nl : 1;  // with no rows the single workgroup takes the L role
  switch (ib) {
    case 8:
      subpanel_kernel<T, 8, MODE, RL><<<grid, 256, 0, h->stream>>>(cm, sw, (T*)l, ldl, (T*)u, ldu,
                                                                    cp, ldcp, counter, nlw);
      break;
    case 16:
      subpanel_kernel<T, 16, MODE, RL><<<grid, 256, 0, h->stream>>>(cm, sw, (T*)l, ldl, (T*)u, ldu,
                                                                     cp, ldcp, counter, nlw);
      break;
    case 32:
      subpanel_kernel<T, 32, MODE, RL><<<grid, 256, 0, h->stream>>>(cm, sw, (T*)l, ldl, (T*)u, ldu,
                                                                     cp, ldcp, counter, nlw);
      break;
    case 64:
      subpanel_kernel<T, 64, MODE, RL><<<grid, 256, 0, h->stream>>>(cm, sw, (T*)l, ldl, (T*)u, ldu,
                                                                     cp, ldcp, counter, nlw);
      break;
    default: return REMIFA_ERR_UNSUPPORTED;
  }
  RF_HIP(hipGetLastError());
  return 0;
}

template <int MODE, bool RL>
int launch_mode(remifa_handle_t h, int ib, int64_t cm, int sw, void* l, int64_t ldl, void* u,
                int64_t ldu, half_t* cp, int64_t ldcp, int* counter) {
  if constexpr (MODE <= SM_F32_TC)
    return launch_ib<float, MODE, RL>(h, ib, cm, sw, l, ldl, u, ldu, cp, ldcp, counter);
  else
    return launch_ib<half_t, MODE, RL>(h, ib, cm, sw, l, ldl, u, ldu, cp, ldcp, counter);
}

}  // namespace

// mode from (storage, inner compute type): see SubMode
int subpanel_mode(int T, int I) {
  if (T == REMIFA_R_32F) return I == REMIFA_FP32 ? SM_F32 : SM_F32_TC;
  if (I == REMIFA_FP16 || I == REMIFA_TC16) return SM_F16_H;
  if (I == REMIFA_BF16_TC32) return SM_F16_BF;
  return SM_F16_TC;
}

int subpanel_launch(remifa_handle_t h, int T, int I, bool rl, int ib, int64_t cm, int sw, void* l,
                    int64_t ldl, void* u, int64_t ldu, half_t* cp, int64_t ldcp, int* counter) {
  if (sw <= 0) return 0;
  if (sw > SW || (ib != 8 && ib != 16 && ib != 32 && ib != 64)) return REMIFA_ERR_UNSUPPORTED;
  const size_t es = dtype_size(T);
  double bytes = 2.0 * ((double)cm * sw + (double)sw * (double)(cm > sw ? cm - sw : 0)) * es;
  timer_begin(h, KC_PANEL, bytes, sw);
  const int mode = subpanel_mode(T, I);
  int st;
#define RF_SUB(M)                                                                              \
  st = rl ? launch_mode<M, true>(h, ib, cm, sw, l, ldl, u, ldu, cp, ldcp, counter)             \
          : launch_mode<M, false>(h, ib, cm, sw, l, ldl, u, ldu, cp, ldcp, counter)
  switch (mode) {
    case SM_F32: RF_SUB(SM_F32); break;
    case SM_F32_TC: RF_SUB(SM_F32_TC); break;
    case SM_F16_TC: RF_SUB(SM_F16_TC); break;
    case SM_F16_H: RF_SUB(SM_F16_H); break;
    default: RF_SUB(SM_F16_BF); break;
  }
#undef RF_SUB
  timer_end(h);
  return st;
}

}  // namespace remifa
