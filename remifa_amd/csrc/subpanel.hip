// Fused sub-panel factorization: the inner ib loop of the reference drivers
// (src/lu_nopiv_inner_ll.hxx:197-370, src/lu_nopiv_inner_rl.hxx:104-196)
// for an sw-wide slice (sw <= 64) of the outer block, in two launches.
//
// Input: the sw-wide block column l[0:cm, 0:sw] (its top sw x sw tile is the
// diagonal tile) and the sw-tall block row u[0:sw, sw:cm], already updated by
// everything left of / above the slice. Output: the factors, in place, and
// optionally their fp16 mirror cp (the reference's d_a_cpy).
//
//  1. tile_kernel (eight wavefronts, the tile in registers) factors the
//     diagonal tile, padded with the identity to 64 x 64, stores it back, and
//     publishes it to a small global table: the factors D and D^T, their operand-rounded
//     copies (the cross-block GEMM operands) and the pivot reciprocals.
//  2. vec_kernel (two wavefronts per workgroup): every lane owns one row of L
//     below the tile or one column of U right of it and solves it against
//     the tile. A row x of L needs x <- x U^-1 and a column v of U needs
//     v <- L^-1 v; with M = D for rows and M = D^T for columns (unit
//     reciprocals for columns) both are the same recurrence over the 64
//     positions of the lane's vector. The workgroup copies its half of the
//     table to LDS (read back as broadcasts); the lane's vector goes through
//     registers in groups of 8 and is kept in a lane-private LDS row.
//
// Inside the slice the work follows the reference's ib-granular semantics:
//  * within an ib block (the reference's panel kernel, src/lu_nopiv.cu:123-
//    170): reciprocal of the pivot, multiply, fused multiply-subtract, in the
//    factor precision IF (fp32, or fp16 with one rounding per operation),
//    each entry receiving its updates in pivot order;
//  * across ib blocks (the reference's inner GEMMs): products of operands
//    rounded to the update compute type (fp16 / bf16 for TC types), summed
//    from zero in k order (fp32, or fp16 for FP16/TC16), subtracted once and
//    rounded to the panel precision IP -- left-looking (LL: one update per
//    column block) or right-looking (RL: a rank-ib update after each block).
#include "slice_common.hpp"

namespace remifa {
namespace {

template <int MODE, bool RL>
__global__ void __launch_bounds__(512) tile_kernel(int sw, int ib, void* __restrict__ lv,
                                                   int64_t ldl, half_t* __restrict__ cp,
                                                   int64_t ldcp, float* __restrict__ tab) {
  using Sm = Sem<MODE>;
  using T = typename std::conditional<(MODE <= SM_F32_TC), float, half_t>::type;
  T* const l = (T*)lv;
  __shared__ TileLdsT<SW> S;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);  // wavefront
  const int i = threadIdx.x % 64;                                  // row
  const int cb = w * TP;                                           // first column
  float x[TP];
  {
    const int ic = i < sw ? i : sw - 1;
#pragma unroll
    for (int jj = 0; jj < TP; ++jj) {
      const int j = cb + jj;
      const int jc = j < sw ? j : sw - 1;
      const float v = (float)l[ic + (int64_t)jc * ldl];
      x[jj] = (i < sw && j < sw) ? v : (i == j ? 1.0f : 0.0f);
    }
  }

  tile_factor<MODE, RL, SW>(S, x, ib, w, i);

  // store the factors and publish the table: D, op(D) (multipliers and row
  // entries op-rounded), the transposes, the reciprocals
  if (i < sw) {
    if (cb + TP <= sw) {  // the whole panel is inside the slice
#pragma unroll
      for (int jj = 0; jj < TP; ++jj) l[i + (int64_t)(cb + jj) * ldl] = (T)x[jj];
      if (cp) {
#pragma unroll
        for (int jj = 0; jj < TP; ++jj) cp[i + (int64_t)(cb + jj) * ldcp] = (half_t)x[jj];
      }
    } else {
#pragma unroll
      for (int jj = 0; jj < TP; ++jj) {
        const int j = cb + jj;
        if (j < sw) {
          l[i + (int64_t)j * ldl] = (T)x[jj];
          if (cp) cp[i + (int64_t)j * ldcp] = (half_t)x[jj];
        }
      }
    }
  }
  {
    float o[TP];
#pragma unroll
    for (int jj = 0; jj < TP; ++jj) o[jj] = Sm::op(x[jj]);
    st8(tab + T_D + i * SW + cb, x);  // row i: 32 contiguous bytes per lane
    st8(tab + T_DOP + i * SW + cb, o);
#pragma unroll
    for (int jj = 0; jj < TP; ++jj) {  // transposes: coalesced across lanes
      tab[T_DT + (cb + jj) * SW + i] = x[jj];
      tab[T_DOPT + (cb + jj) * SW + i] = o[jj];
    }
  }
  __syncthreads();
  if (w == 0) {
    tab[T_R + i] = S.R[i];
    tab[T_ONE + i] = 1.0f;
  }
}

// ---------------------------------------------------------------------------
// 2. rows of L and columns of U
constexpr int VL = 128;  // lanes (vectors) per workgroup

template <typename T>
__device__ __forceinline__ void ld16(const T* p, float* v);  // 16 bytes -> floats
template <>
__device__ __forceinline__ void ld16<float>(const float* p, float* v) {
  const float4 q = *(const float4*)p;
  v[0] = q.x, v[1] = q.y, v[2] = q.z, v[3] = q.w;
}
template <>
__device__ __forceinline__ void ld16<half_t>(const half_t* p, float* v) {
  typedef _Float16 h8 __attribute__((ext_vector_type(8)));
  const h8 q = *(const h8*)p;
#pragma unroll
  for (int e = 0; e < 8; ++e) v[e] = (float)q[e];
}
template <typename T>
__device__ __forceinline__ void st16(T* p, const float* v);  // floats -> 16 bytes
template <>
__device__ __forceinline__ void st16<float>(float* p, const float* v) {
  *(float4*)p = make_float4(v[0], v[1], v[2], v[3]);
}
template <>
__device__ __forceinline__ void st16<half_t>(half_t* p, const float* v) {
  typedef _Float16 h8 __attribute__((ext_vector_type(8)));
  h8 q;
#pragma unroll
  for (int e = 0; e < 8; ++e) q[e] = (_Float16)v[e];
  *(h8*)p = q;
}

// vflags: bit 0 = the U columns may use 16-byte accesses (u and ldu
// aligned), bit 1 = likewise for the fp16 mirror cp
template <typename T, int MODE, bool RL>
__global__ void __launch_bounds__(VL) vec_kernel(int64_t cm, int sw, int ib, T* __restrict__ l,
                                                 int64_t ldl, T* __restrict__ u, int64_t ldu,
                                                 half_t* __restrict__ cp, int64_t ldcp,
                                                 const float* __restrict__ tab, int nl_wg,
                                                 int vflags) {
  using Sm = Sem<MODE>;
  constexpr int EV = 16 / sizeof(T);  // entries per 16-byte access
  // the tile of this workgroup's role (M = D for rows of L, D^T for columns
  // of U; Mo = op(M); Rs = 1/u_kk for rows, ones for columns), and the
  // lane-private vectors (odd stride: lane t, entry j sits in bank (t+j)%64)
  __shared__ float Mt[SW][SW + PD];
  __shared__ float Mot[SW][SW + PD];
  __shared__ float Rs[SW];
  __shared__ float Xa[VL][SW + 1];  // finished values, then current ones
  const int t = threadIdx.x;
  const int wg = blockIdx.x;
  const bool is_l = wg < nl_wg;  // workgroup-uniform
  {
    const float* M = tab + (is_l ? T_D : T_DT);
    const float* Mo = tab + (is_l ? T_DOP : T_DOPT);
    constexpr int NQ = SW * SW / 4 / VL;  // 16-byte pieces per lane and matrix
    float4 qm[NQ], qo[NQ];
#pragma unroll
    for (int e = 0; e < NQ; ++e) {
      const int f = t + e * VL;
      qm[e] = *(const float4*)&M[f * 4];
      qo[e] = *(const float4*)&Mo[f * 4];
    }
#pragma unroll
    for (int e = 0; e < NQ; ++e) {
      const int f = t + e * VL;
      const int r = f / (SW / 4), c = (f % (SW / 4)) * 4;
      *(float4*)&Mt[r][c] = qm[e];
      *(float4*)&Mot[r][c] = qo[e];
    }
    if (t < SW) Rs[t] = tab[(is_l ? T_R : T_ONE) + t];
  }
  __syncthreads();
  const int64_t idx = sw + (int64_t)(is_l ? wg : wg - nl_wg) * VL + t;
  if (idx >= cm) return;  // no barriers below: lanes are independent
  // entry j of the lane's vector is p[j * st]; its mirror q[j * sq]
  T* const p = is_l ? l + idx : u + idx * ldu;
  const int64_t st = is_l ? ldl : 1;
  half_t* const q = cp ? (is_l ? cp + idx : cp + idx * ldcp) : nullptr;
  const int64_t sq = is_l ? ldcp : 1;
  const bool vp = !is_l && (vflags & 1);   // 16-byte accesses along the vector
  const bool vq = !is_l && (vflags & 2);
  float* const xa = &Xa[t][0];  // (operands of the block sums: op(xa[j]))

  // the whole vector in one go (all loads in flight together, addresses
  // clamped in bounds); xa[j] is overwritten by the finished value later
  {
    // all loads first (one memory latency), then the LDS row
    float v[SW];
    if (vp && sw == SW) {
#pragma unroll
      for (int j = 0; j < SW; j += EV) ld16<T>(p + j, v + j);
    } else {
      T raw[SW];
#pragma unroll
      for (int j = 0; j < SW; ++j) raw[j] = p[(j < sw ? j : sw - 1) * st];
#pragma unroll
      for (int j = 0; j < SW; ++j) v[j] = j < sw ? (float)raw[j] : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < SW; ++j) xa[j] = v[j];
  }
  for (int c0 = 0; c0 < sw; c0 += G) {
    const int kb = c0 - c0 % ib;  // start of the ib block holding c0
    float xb[G];
#pragma unroll
    for (int c = 0; c < G; ++c) xb[c] = xa[c0 + c];
    if (!RL && kb > 0) {
      // left-looking cross-block update, products summed from zero
      float acc[G];
#pragma unroll
      for (int c = 0; c < G; ++c) acc[c] = 0.0f;
#pragma unroll 4
      for (int j = 0; j < kb; ++j) {
        const float xj = Sm::op(xa[j]);
        const float* mr = &Mot[j][c0];
#pragma unroll
        for (int c = 0; c < G; ++c) acc[c] = Sm::acc_add(acc[c], xj, mr[c]);
      }
#pragma unroll
      for (int c = 0; c < G; ++c) xb[c] = Sm::apply(xb[c], acc[c]);
    }
    // pivots of this ib block already finished (earlier groups), in order
#pragma unroll 4
    for (int k = kb; k < c0; ++k) {
      const float lk = xa[k];
      const float* mr = &Mt[k][c0];
#pragma unroll
      for (int c = 0; c < G; ++c) xb[c] = Sm::fnma(lk, mr[c], xb[c]);
    }
    // pivots of this group
#pragma unroll
    for (int kk = 0; kk < G; ++kk) {
      const float lk = Sm::mul(xb[kk], Rs[c0 + kk]);
      xb[kk] = lk;
      const float* mr = &Mt[c0 + kk][c0];
#pragma unroll
      for (int c = kk + 1; c < G; ++c) xb[c] = Sm::fnma(lk, mr[c], xb[c]);
    }
    if (vp && c0 + G <= sw) {
#pragma unroll
      for (int c = 0; c < G; c += EV) st16<T>(p + c0 + c, xb + c);
    } else {
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c0 + c < sw) p[(c0 + c) * st] = (T)xb[c];
    }
    if (q) {
      if (vq && c0 + G <= sw) {
        st16<half_t>(q + c0, xb);
      } else {
#pragma unroll
        for (int c = 0; c < G; ++c)
          if (c0 + c < sw) q[(c0 + c) * sq] = (half_t)xb[c];
      }
    }
#pragma unroll
    for (int c = 0; c < G; ++c) xa[c0 + c] = xb[c];
    const int kend = kb + ib < sw ? kb + ib : sw;
    if (RL && c0 + G >= kend && kend < sw) {
      // the ib block is complete: rank-ib update of the entries right of it
      for (int c1 = kend; c1 < sw; c1 += G) {
        float acc[G];
#pragma unroll
        for (int c = 0; c < G; ++c) acc[c] = 0.0f;
#pragma unroll 4
        for (int k = kb; k < kend; ++k) {
          const float xk = Sm::op(xa[k]);
          const float* mr = &Mot[k][c1];
#pragma unroll
          for (int c = 0; c < G; ++c) acc[c] = Sm::acc_add(acc[c], xk, mr[c]);
        }
#pragma unroll
        for (int c = 0; c < G; ++c) xa[c1 + c] = Sm::apply(xa[c1 + c], acc[c]);
      }
    }
  }
}

// 2b. full 64-wide slices: the same recurrence, one ib block of the vector
// at a time in registers (the finished blocks kept op-rounded for the
// cross-block sums), the tile in LDS read as broadcasts. Each entry receives
// the same operations in the same order as in vec_kernel: the cross-block
// sums (LL: one sum over every finished block, RL: one per finished block,
// in block order), then the pivots of its own block in pivot order. A
// workgroup of 256 lanes takes one role (rows of L or columns of U).
constexpr int VS = 256;  // lanes per workgroup

template <typename T, int MODE, bool RL, int IB>
__global__ void __launch_bounds__(VS) vecs_kernel(int64_t cm, T* __restrict__ l, int64_t ldl,
                                                  T* __restrict__ u, int64_t ldu,
                                                  half_t* __restrict__ cp, int64_t ldcp,
                                                  const float* __restrict__ tab, int nl_wg,
                                                  int vflags) {
  using Sm = Sem<MODE>;
  constexpr bool OPR = MODE == SM_F32_TC || MODE == SM_F16_BF;  // op(M) differs from M
  constexpr bool XSUM = RL || IB < SW;                          // cross-block sums occur
  __shared__ __attribute__((aligned(16))) float Ms[SW * SW];
  __shared__ __attribute__((aligned(16))) float Mos[(OPR && XSUM) ? SW * SW : 4];
  __shared__ float Rs[SW];
  const int t = threadIdx.x;
  const int wg = blockIdx.x;
  const bool is_l = wg < nl_wg;  // workgroup-uniform
  {
    const float4* M4 = (const float4*)(tab + (is_l ? T_D : T_DT));
    const float4* Mo4 = (const float4*)(tab + (is_l ? T_DOP : T_DOPT));
#pragma unroll
    for (int e = t; e < SW * SW / 4; e += VS) {
      ((float4*)Ms)[e] = M4[e];
      if constexpr (OPR && XSUM) ((float4*)Mos)[e] = Mo4[e];
    }
    if (t < SW) Rs[t] = tab[(is_l ? T_R : T_ONE) + t];
  }
  const float* const Mo = (OPR && XSUM) ? Mos : Ms;
  __syncthreads();
  const int64_t idx = SW + (int64_t)(is_l ? wg : wg - nl_wg) * VS + t;
  if (idx >= cm) return;  // no barriers below
  T* const p = is_l ? l + idx : u + idx * ldu;
  const int64_t st = is_l ? ldl : 1;
  const bool vp = !is_l && (vflags & 1);
  half_t* const q = cp ? (is_l ? cp + idx : cp + idx * ldcp) : nullptr;
  const int64_t sq = is_l ? ldcp : 1;
  const bool vq = !is_l && (vflags & 2);
  constexpr int EV = 16 / sizeof(T);
  float xo[XSUM ? SW : 1];  // op-rounded finished entries
#pragma unroll
  for (int b0 = 0; b0 < SW; b0 += IB) {
    asm volatile("" ::: "memory");  // one block at a time
    float x[IB];
    if (vp) {
#pragma unroll
      for (int j = 0; j < IB; j += EV) ld16<T>(p + b0 + j, x + j);
    } else {
#pragma unroll
      for (int j = 0; j < IB; ++j) x[j] = (float)p[(b0 + j) * st];
    }
    if constexpr (XSUM) {
      // cross-block sums: LL one over every finished block, RL one per block
#pragma unroll
      for (int pb = 0; pb < b0; pb += (RL ? IB : SW)) {
        const int pe = RL ? pb + IB : b0;
        float acc[IB];
#pragma unroll
        for (int c = 0; c < IB; ++c) acc[c] = 0.0f;
#pragma unroll
        for (int j = pb; j < pe; ++j) {
#pragma unroll
          for (int c = 0; c < IB; c += 4) {
            const float4 m = *(const float4*)&Mo[j * SW + b0 + c];
            acc[c] = Sm::acc_add(acc[c], xo[j], m.x);
            acc[c + 1] = Sm::acc_add(acc[c + 1], xo[j], m.y);
            acc[c + 2] = Sm::acc_add(acc[c + 2], xo[j], m.z);
            acc[c + 3] = Sm::acc_add(acc[c + 3], xo[j], m.w);
          }
        }
#pragma unroll
        for (int c = 0; c < IB; ++c) x[c] = Sm::apply(x[c], acc[c]);
      }
    }
    // this block's pivots, in order, groups of 8 entries
#pragma unroll
    for (int g0 = 0; g0 < IB; g0 += G) {
#pragma unroll
      for (int k = 0; k < g0; ++k) {
        const float4 a = *(const float4*)&Ms[(b0 + k) * SW + b0 + g0];
        const float4 b = *(const float4*)&Ms[(b0 + k) * SW + b0 + g0 + 4];
        const float mr[G] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int c = 0; c < G; ++c) x[g0 + c] = Sm::fnma(x[k], mr[c], x[g0 + c]);
      }
#pragma unroll
      for (int kk = 0; kk < G; ++kk) {
        const float lk = Sm::mul(x[g0 + kk], Rs[b0 + g0 + kk]);
        x[g0 + kk] = lk;
        const float4 a = *(const float4*)&Ms[(b0 + g0 + kk) * SW + b0 + g0];
        const float4 b = *(const float4*)&Ms[(b0 + g0 + kk) * SW + b0 + g0 + 4];
        const float mr[G] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
        for (int c = kk + 1; c < G; ++c) x[g0 + c] = Sm::fnma(lk, mr[c], x[g0 + c]);
      }
    }
    // the finished block: factors, mirror, op-rounded copy
    if (vp) {
#pragma unroll
      for (int j = 0; j < IB; j += EV) st16<T>(p + b0 + j, x + j);
    } else {
#pragma unroll
      for (int j = 0; j < IB; ++j) p[(b0 + j) * st] = (T)x[j];
    }
    if (q) {
      if (vq) {
#pragma unroll
        for (int j = 0; j < IB; j += 8) st16<half_t>(q + b0 + j, x + j);
      } else {
#pragma unroll
        for (int j = 0; j < IB; ++j) q[(b0 + j) * sq] = (half_t)x[j];
      }
    }
    if constexpr (XSUM) {
#pragma unroll
      for (int j = 0; j < IB; ++j) xo[b0 + j] = Sm::op(x[j]);
    }
  }
}

template <typename T, int MODE, bool RL>
int launch_typed(remifa_handle_t h, int ib, int64_t cm, int sw, void* l, int64_t ldl, void* u,
                 int64_t ldu, half_t* cp, int64_t ldcp, float* tab, double bytes) {
  timer_begin(h, KC_PANEL, (double)sw * sw * 2 * sizeof(T), 0);
  tile_kernel<MODE, RL><<<1, 512, 0, h->stream>>>(sw, ib, l, ldl, cp, ldcp, tab);
  RF_HIP(hipGetLastError());
  timer_end(h);
  const int64_t rest = cm - sw;
  if (rest <= 0) return 0;
  const int nl = (int)ceil_div(rest, VL);
  // 16-byte accesses along the U columns need aligned column starts
  const int vflags = ((uintptr_t)u % 16 == 0 && (ldu * sizeof(T)) % 16 == 0 ? 1 : 0) |
                     (cp && (uintptr_t)cp % 16 == 0 && (ldcp * 2) % 16 == 0 ? 2 : 0);
  timer_begin(h, KC_PANEL, bytes, 1);
  if (sw == SW) {
    // full slices: one ib block of each vector at a time in registers
    const int nw = (int)ceil_div(rest, VS);
#define RF_VECS(IBV)                                                                          \
  vecs_kernel<T, MODE, RL, IBV><<<2 * nw, VS, 0, h->stream>>>(cm, (T*)l, ldl, (T*)u, ldu, cp, \
                                                               ldcp, tab, nw, vflags)
    switch (ib) {
      case 8: RF_VECS(8); break;
      case 16: RF_VECS(16); break;
      case 32: RF_VECS(32); break;
      default: RF_VECS(64); break;
    }
#undef RF_VECS
  } else {
    vec_kernel<T, MODE, RL><<<2 * nl, VL, 0, h->stream>>>(cm, sw, ib, (T*)l, ldl, (T*)u, ldu, cp,
                                                           ldcp, tab, nl, vflags);
  }
  RF_HIP(hipGetLastError());
  timer_end(h);
  return 0;
}

template <int MODE, bool RL>
int launch_mode(remifa_handle_t h, int ib, int64_t cm, int sw, void* l, int64_t ldl, void* u,
                int64_t ldu, half_t* cp, int64_t ldcp, float* tab, double bytes) {
  if constexpr (MODE <= SM_F32_TC)
    return launch_typed<float, MODE, RL>(h, ib, cm, sw, l, ldl, u, ldu, cp, ldcp, tab, bytes);
  else
    return launch_typed<half_t, MODE, RL>(h, ib, cm, sw, l, ldl, u, ldu, cp, ldcp, tab, bytes);
}

}  // namespace

// bytes of the published-tile table a slice needs in device memory
size_t subpanel_table_bytes() { return (size_t)T_FLOATS * sizeof(float); }

// mode from (storage, inner compute type): see SubMode
int subpanel_mode(int T, int I) {
  if (T == REMIFA_R_32F) return I == REMIFA_FP32 ? SM_F32 : SM_F32_TC;
  if (I == REMIFA_FP16 || I == REMIFA_TC16) return SM_F16_H;
  if (I == REMIFA_BF16_TC32) return SM_F16_BF;
  return SM_F16_TC;
}

int subpanel_launch(remifa_handle_t h, int T, int I, bool rl, int ib, int64_t cm, int sw, void* l,
                    int64_t ldl, void* u, int64_t ldu, half_t* cp, int64_t ldcp, float* tab) {
  if (sw <= 0) return 0;
  if (sw > SW || (ib != 8 && ib != 16 && ib != 32 && ib != 64) || cm < sw)
    return REMIFA_ERR_UNSUPPORTED;
  const size_t es = dtype_size(T);
  const double rest = (double)(cm - sw);
  const double bytes = 2.0 * (2.0 * rest * sw) * es + (cp ? 2.0 * rest * sw * 2 : 0.0);
  const int mode = subpanel_mode(T, I);
#define RF_SUB(M)                                                                              \
  return rl ? launch_mode<M, true>(h, ib, cm, sw, l, ldl, u, ldu, cp, ldcp, tab, bytes)        \
            : launch_mode<M, false>(h, ib, cm, sw, l, ldl, u, ldu, cp, ldcp, tab, bytes)
  switch (mode) {
    case SM_F32: RF_SUB(SM_F32);
    case SM_F32_TC: RF_SUB(SM_F32_TC);
    case SM_F16_TC: RF_SUB(SM_F16_TC);
    case SM_F16_H: RF_SUB(SM_F16_H);
    default: RF_SUB(SM_F16_BF);
  }
#undef RF_SUB
}

}  // namespace remifa
