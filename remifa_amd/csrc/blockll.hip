// Inner factorization of one outer block of the left-looking driver in two
// launches: the reference's ib loop (src/lu_nopiv_inner_ll.hxx:197-370)
// with ib = 32, for the buffered layout (fp16 storage, the block
// column and block row already updated into fp32 buffers, TC32 inner
// updates, fp32 panel arithmetic).
//
// The nb x nb diagonal block is cut into slices of 64 columns (two ib
// blocks). A "vector" is a row of L below a slice's 64 x 64 diagonal tile
// or a column of U right of it; a "unit" is 32 consecutive vectors of one
// kind, the N side of one MFMA tile. For a unit and a slice at column so,
// with M the tile's U (rows of L) or L^T (columns of U):
//   block 0 (entries so .. so+31):  x = C - sum_{k < so}      op(.) op(.)
//   block 1 (entries so+32 .. +63): x = C - sum_{k < so + 32} op(.) op(.)
// each one fp32 accumulation on the matrix cores (v_mfma_f32_32x32x16_f16:
// op = fp16 rounding) that starts from -C, as the drivers' update GEMMs
// start from C, then the block's pivots in order: x_k *= 1/u_kk (rows of
// L), x_e -= x_k M[k][e] (fp32 fused multiply-add, entry e in pivot
// order). That is the
// reference's per-ib-block GEMM (one update over every column of the outer
// block left of the ib block) followed by its panel kernel
// (src/lu_nopiv.cu:123-170, 251-326); the slice kernels (subpanel.hip)
// split the same update sum at the slice boundary instead.
//
// Two launches per outer block, on the handle's stream:
//  * the diagonal kernel, one workgroup: for each slice, the tile's update
//    (MFMA, K = so), its factorization (tile_factor, eight wavefronts),
//    its slice table (the factored tile in the layouts the units read)
//    written to global memory, then the units inside the diagonal block
//    (rows / columns so+64 .. nb-1) for that slice;
//  * the far kernel: the units of the rows of L and columns of U beyond the
//    diagonal block, each through every slice (a workgroup stages a slice's
//    table and the other factor's tile vectors in LDS, then its wavefronts
//    take their units). Its workgroups are small enough (registers, LDS) to
//    share a CU with the second stream's update GEMM.
// A far unit reads the slice tables, the other factor's vectors inside the
// diagonal block and its own earlier results: everything it reads was
// written by the diagonal kernel before the launch or by itself.
//
// One-launch form (the default): both in one grid. The first workgroup to
// start (an atomic ticket, not blockIdx) takes the diagonal role: per slice
// the tile's update, factorization and table, published through a counter,
// then only the next tile's four units. The other workgroups (helpers) start
// slice s when the counter says so and take every other unit: the far units
// and the diagonal block's rows / columns beyond the next tile, whose
// progress they publish per 32-vector block; the diagonal role waits for
// that before it takes such a block as the next tile's. A helper's slice s
// reads only the table and the tile's vectors, all final by the time slice
// s is published. Helpers wait only for the diagonal role, which waits only
// for slices that helpers can finish without further waits; the grid stays
// below half the CUs, so a helper that has not started finds a CU once the
// other stream's workgroups there end.
//
// MFMA operand layout used below (v_mfma_f32_32x32x16_f16, D = A B + C):
// lane l supplies A[i = l % 32][k = 8 (l / 32) .. +7] and
// B[k = 8 (l / 32) .. +7][j = l % 32]; register r of lane l holds
// D[i = 8 (r / 4) + 4 (l / 32) + r % 4][j = l % 32]. With A = the other
// factor (rows i = tile entries) and B = the unit's own vectors (j = vector),
// lane (n, h) holds 16 of the 32 entries of vector n in a block:
// e = 8 q + 4 h + j for register 4 q + j. A pivot's multiplier goes to the
// other half with v_permlane32_swap (lanes 0-31 <-> 32-63).
#include <algorithm>

#include "slice_common.hpp"

namespace remifa {
namespace {

constexpr int BLT = 512;  // threads per workgroup (eight wavefronts)
constexpr int NV = 32;    // vectors per unit
constexpr int NBM = 512;  // largest outer block
constexpr int KSM = (NBM - SW) / 16;  // most k-steps of a unit's update
constexpr int PP = NBM - SW + 8;      // row pitch (halves) of the staged panel
constexpr int kBufWord3 = 0x00020000;  // buffer resource word 3 (raw 32-bit data)

typedef _Float16 bh8 __attribute__((ext_vector_type(8)));
typedef _Float16 bh4 __attribute__((ext_vector_type(4)));
typedef float bf16v __attribute__((ext_vector_type(16)));

// one role's part of a slice table (rows of L: M = D; columns of U: M = D^T)
struct RoleTab {
  // the diagonal 32 x 32 blocks, m[b][k][e] = M[32 b + k][32 b + e] for
  // e != k; the diagonal holds the pivot's scale instead (1 / u_kk for the
  // rows of L, 1 for the columns of U), so that a pivot's row read brings it
  float m[2][32][32];
  _Float16 mot[32][32];   // op(M[k][32 + e]) stored [e][k] (cross-block operand)
};
struct SliceTab {
  RoleTab role[2];
};
static_assert(sizeof(RoleTab) % 16 == 0, "role table size");

struct BlkArgs {
  int64_t updm;         // rows of the block column = columns of the block row
  int nb;               // width of the outer block (multiple of 64, <= NBM)
  int nfu;              // far units of each role
  int nfw;              // far workgroups of each role
  const float* pl;      // block column buffer (updm x nb, fp32)
  int64_t ldpl;
  const float* pu;      // block row buffer (nb x updm, fp32)
  int64_t ldpu;
  half_t* a;            // the factors (fp16), the outer block at (0, 0)
  int64_t lda;
  half_t* lop;          // op(L) rows, k-contiguous: lop[v * ldo + k]
  half_t* uop;          // op(U) columns, k-contiguous: uop[v * ldo + k]
  int64_t ldo;
  SliceTab* tabs;       // one per slice
  // one-launch form: sync[0] ticket counter (tickets of this launch start at
  // tbase), sync[1] published slices (epoch * 16 + s + 1 once slice s's
  // table is written; epoch = launch number within the factorization),
  // sync[2] wait-limit flag, sync[4 + 16 role + b] the slices done of
  // diagonal-block unit b (same encoding)
  int* sync;
  int tbase;
  int epoch;
  long long wait_limit;  // polls of a published counter per wait
};

// LDS of the diagonal kernel (one workgroup; 162,048 bytes of the CU's 160 KiB)
struct DiagLds {
  union {
    TileLdsT<32> fac;          // the tile factorization
    float tt[SW][SW + 4];      // the updated tile before factorization
  } f;
  _Float16 pan[2][SW][PP];     // [role]: the other factor's tile vectors, [e][k]
  SliceTab tab;                // the current slice's table
};

__device__ __forceinline__ unsigned f2u(float v) { return __builtin_bit_cast(unsigned, v); }
__device__ __forceinline__ float u2f(unsigned v) { return __builtin_bit_cast(float, v); }

// value of v in half HK of the wavefront (lanes 0-31 or 32-63) in every
// lane: lane n of either half receives lane n of half HK
template <int HK>
__device__ __forceinline__ float bcast_half(float v) {
  const auto r = __builtin_amdgcn_permlane32_swap(f2u(v), f2u(v), false, false);
  return u2f(HK ? r[1] : r[0]);
}

__device__ __forceinline__ unsigned pack_h2(float a, float b) {
  const unsigned lo = __builtin_bit_cast(unsigned short, (_Float16)a);
  const unsigned hi = __builtin_bit_cast(unsigned short, (_Float16)b);
  return lo | (hi << 16);
}

// The pivots of one 32-entry block of a unit, in order (see the header).
// The tile's rows are read from LDS two pivots ahead of their use; the
// owner of pivot k finds its scale in its own part of row k (entry k).
template <int MODE>
__device__ __forceinline__ void solve_block(float (&x)[16], const float (*m)[32], int lh) {
  using Sm = Sem<MODE>;
  constexpr int PF = 3;  // row buffers in flight
  float4 mv[PF][4];
  auto fetch = [&](int k) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (8 * q + 7 >= k) mv[k % PF][q] = *(const float4*)&m[k][8 * q + 4 * lh];
  };
  fetch(0);
  fetch(1);
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    // (keeps each row's reads two pivots ahead, not all at the top)
    asm volatile("" ::: "memory");
    if (k + 2 < 32) fetch(k + 2);
    const int hk = (k >> 2) & 1, ik = 4 * (k >> 3) + (k & 3);
    const float4 d4 = mv[k % PF][k >> 3];
    const float dk[4] = {d4.x, d4.y, d4.z, d4.w};
    // (in the owner's half, entry k of row k is the scale)
    const float own = Sm::mul(x[ik], dk[k & 3]);
    const float lk = hk ? bcast_half<1>(own) : bcast_half<0>(own);
    x[ik] = lh == hk ? lk : x[ik];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (8 * q + 7 <= k) continue;  // the group is finished in both halves
      const float4 v4 = mv[k % PF][q];
      const float mm[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e0 = 8 * q + j, e1 = e0 + 4;  // entry in half 0 / half 1
        if (e1 <= k) continue;
        const float y = Sm::fnma(lk, mm[j], x[4 * q + j]);
        x[4 * q + j] = (e0 > k || lh) ? y : x[4 * q + j];
      }
    }
  }
}

// One unit (32 vectors from v0; role 0 = rows of L, 1 = columns of U) for
// the slice at column so, on one wavefront. T = the role's table and pan
// the other factor's op-rounded vectors of the tile ([e][k], k < so), both
// in LDS.
template <int MODE, int KC>
__device__ __forceinline__ void do_unit(const BlkArgs& g, const RoleTab& T,
                                        const _Float16 (*pan)[PP], int role, int64_t v0,
                                        int so) {
  using Sm = Sem<MODE>;
  const int lane = threadIdx.x & 63, l32 = lane & 31, lh = lane >> 5;
  const int64_t v = v0 + l32;
  const bool ok = v < g.updm;
  const int64_t vc = ok ? v : g.updm - 1;
  const int64_t ldo = g.ldo;
  // this unit's own earlier entries (the B operand) and its entries of the
  // buffers, every load in flight at once. The accumulators start at -C, so
  // that x = -(sum - C) = C - sum (C enters the sum first, as in the update
  // GEMMs of the drivers).
  const half_t* own = (role ? g.uop : g.lop) + vc * ldo + 8 * lh;
  bf16v acc0, acc1;
  if (role == 0) {
    // rows of L: entry (v, so + e) of the column-major buffer, a per-lane
    // offset plus a uniform one per entry (buffer loads, no per-entry
    // address registers)
    const auto rc = __builtin_amdgcn_make_buffer_rsrc((void*)(g.pl + (int64_t)so * g.ldpl), 0,
                                                      (int)(SW * g.ldpl * 4), kBufWord3);
    const int vo = (int)(vc + 4 * lh * g.ldpl) * 4;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int eu = 8 * (r >> 2) + (r & 3);
      // (the builtin returns the 32 bits of the entry)
      acc0[r] = -u2f(__builtin_amdgcn_raw_buffer_load_b32(rc, vo, (int)(eu * g.ldpl * 4), 0));
      acc1[r] = -u2f(__builtin_amdgcn_raw_buffer_load_b32(rc, vo, (int)((32 + eu) * g.ldpl * 4), 0));
    }
  } else {
    const float* c = g.pu + so + vc * g.ldpu + 4 * lh;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float4 p = *(const float4*)(c + 8 * q);
      const float4 s = *(const float4*)(c + 32 + 8 * q);
      acc0[4 * q] = -p.x, acc0[4 * q + 1] = -p.y, acc0[4 * q + 2] = -p.z, acc0[4 * q + 3] = -p.w;
      acc1[4 * q] = -s.x, acc1[4 * q + 1] = -s.y, acc1[4 * q + 2] = -s.z, acc1[4 * q + 3] = -s.w;
    }
  }
  // the update, KC k-steps of 16 per round trip
  for (int c0 = 0; c0 < so; c0 += 16 * KC) {
    bh8 y[KC];
#pragma unroll
    for (int u = 0; u < KC; ++u)
      if (c0 + 16 * u < so) y[u] = *(const bh8*)(own + c0 + 16 * u);
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      if (u % 4 == 0) asm volatile("" ::: "memory");  // (a few LDS reads ahead at a time)
      if (c0 + 16 * u < so) {
        const bh8 a0 = *(const bh8*)&pan[l32][c0 + 16 * u + 8 * lh];
        const bh8 a1 = *(const bh8*)&pan[32 + l32][c0 + 16 * u + 8 * lh];
        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, y[u], acc0, 0, 0, 0);
        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, y[u], acc1, 0, 0, 0);
      }
    }
  }
  float x0[16], x1[16];
  // block 0
#pragma unroll
  for (int r = 0; r < 16; ++r) x0[r] = -acc0[r];
  solve_block<MODE>(x0, T.m[0], lh);
  // block 1: its part of the sum over block 0 joins the same accumulator
  // (B operand: op(x0) rearranged across the halves, see the header)
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    const int g0 = 8 * ks, g1 = 8 * ks + 4;  // registers of groups q = 2 ks, 2 ks + 1
    const unsigned p0 = pack_h2(x0[g0], x0[g0 + 1]), p1 = pack_h2(x0[g0 + 2], x0[g0 + 3]);
    const unsigned q0 = pack_h2(x0[g1], x0[g1 + 1]), q1 = pack_h2(x0[g1 + 2], x0[g1 + 3]);
    const auto sa = __builtin_amdgcn_permlane32_swap(p0, q0, false, false);
    const auto sb = __builtin_amdgcn_permlane32_swap(p1, q1, false, false);
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    const u4 yb = {sa[0], sb[0], sa[1], sb[1]};
    const bh8 av = *(const bh8*)&T.mot[l32][16 * ks + 8 * lh];
    acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, __builtin_bit_cast(bh8, yb), acc1, 0, 0, 0);
  }
#pragma unroll
  for (int r = 0; r < 16; ++r) x1[r] = -acc1[r];
  solve_block<MODE>(x1, T.m[1], lh);
  if (!ok) return;
  // the factors (fp16) and the op-rounded copy for later slices
  half_t* oc = (role ? g.uop : g.lop) + vc * ldo + so + 4 * lh;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const bh4 h0 = {(_Float16)x0[4 * q], (_Float16)x0[4 * q + 1], (_Float16)x0[4 * q + 2],
                    (_Float16)x0[4 * q + 3]};
    const bh4 h1 = {(_Float16)x1[4 * q], (_Float16)x1[4 * q + 1], (_Float16)x1[4 * q + 2],
                    (_Float16)x1[4 * q + 3]};
    *(bh4*)(oc + 8 * q) = h0;
    *(bh4*)(oc + 32 + 8 * q) = h1;
    if (role) {
      half_t* ac = g.a + so + vc * g.lda + 4 * lh;
      *(bh4*)(ac + 8 * q) = h0;
      *(bh4*)(ac + 32 + 8 * q) = h1;
    }
  }
  if (!role) {
    const auto ra = __builtin_amdgcn_make_buffer_rsrc((void*)(g.a + (int64_t)so * g.lda), 0,
                                                      (int)(SW * g.lda * 2), kBufWord3);
    const int vo = (int)(vc + 4 * lh * g.lda) * 2;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int eu = 8 * (r >> 2) + (r & 3);
      __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)x0[r]),
                                            ra, vo, (int)(eu * g.lda * 2), 0);
      __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)x1[r]),
                                            ra, vo, (int)((32 + eu) * g.lda * 2), 0);
    }
  }
}

// solve_block for N units of one role at once (the same table rows feed
// all of them; their pivot chains interleave)
template <int MODE, int N>
__device__ __forceinline__ void solve_blocks(float (&x)[N][16], const float (*m)[32], int lh) {
  using Sm = Sem<MODE>;
  constexpr int PF = 3;  // row buffers in flight
  float4 mv[PF][4];
  auto fetch = [&](int k) {
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (8 * q + 7 >= k) mv[k % PF][q] = *(const float4*)&m[k][8 * q + 4 * lh];
  };
  fetch(0);
  fetch(1);
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    asm volatile("" ::: "memory");
    if (k + 2 < 32) fetch(k + 2);
    const int hk = (k >> 2) & 1, ik = 4 * (k >> 3) + (k & 3);
    const float4 d4 = mv[k % PF][k >> 3];
    const float dk[4] = {d4.x, d4.y, d4.z, d4.w};
    float lk[N];
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const float own = Sm::mul(x[n][ik], dk[k & 3]);
      lk[n] = hk ? bcast_half<1>(own) : bcast_half<0>(own);
      x[n][ik] = lh == hk ? lk[n] : x[n][ik];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (8 * q + 7 <= k) continue;
      const float4 v4 = mv[k % PF][q];
      const float mm[4] = {v4.x, v4.y, v4.z, v4.w};
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e0 = 8 * q + j, e1 = e0 + 4;
        if (e1 <= k) continue;
#pragma unroll
        for (int n = 0; n < N; ++n) {
          const float y = Sm::fnma(lk[n], mm[j], x[n][4 * q + j]);
          x[n][4 * q + j] = (e0 > k || lh) ? y : x[n][4 * q + j];
        }
      }
    }
  }
}

// do_unit for N units of one role at once (vectors from v0[n]): the same
// arithmetic per unit; the other factor's tile vectors and the table rows
// are read once for all of them and their chains interleave.
template <int MODE, int KC, int N>
__device__ __forceinline__ void do_units(const BlkArgs& g, const RoleTab& T,
                                         const _Float16 (*pan)[PP], int role,
                                         const int64_t (&v0)[N], int so) {
  const int lane = threadIdx.x & 63, l32 = lane & 31, lh = lane >> 5;
  const int64_t ldo = g.ldo;
  bool ok[N];
  int64_t vc[N];
  const half_t* own[N];
#pragma unroll
  for (int n = 0; n < N; ++n) {
    const int64_t v = v0[n] + l32;
    ok[n] = v < g.updm;
    vc[n] = ok[n] ? v : g.updm - 1;
    own[n] = (role ? g.uop : g.lop) + vc[n] * ldo + 8 * lh;
  }
  bf16v acc0[N], acc1[N];
  if (role == 0) {
    const auto rc = __builtin_amdgcn_make_buffer_rsrc((void*)(g.pl + (int64_t)so * g.ldpl), 0,
                                                      (int)(SW * g.ldpl * 4), kBufWord3);
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const int vo = (int)(vc[n] + 4 * lh * g.ldpl) * 4;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int eu = 8 * (r >> 2) + (r & 3);
        acc0[n][r] = -u2f(__builtin_amdgcn_raw_buffer_load_b32(rc, vo, (int)(eu * g.ldpl * 4), 0));
        acc1[n][r] =
            -u2f(__builtin_amdgcn_raw_buffer_load_b32(rc, vo, (int)((32 + eu) * g.ldpl * 4), 0));
      }
    }
  } else {
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const float* c = g.pu + so + vc[n] * g.ldpu + 4 * lh;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float4 p = *(const float4*)(c + 8 * q);
        const float4 sv = *(const float4*)(c + 32 + 8 * q);
        acc0[n][4 * q] = -p.x, acc0[n][4 * q + 1] = -p.y, acc0[n][4 * q + 2] = -p.z,
                  acc0[n][4 * q + 3] = -p.w;
        acc1[n][4 * q] = -sv.x, acc1[n][4 * q + 1] = -sv.y, acc1[n][4 * q + 2] = -sv.z,
                  acc1[n][4 * q + 3] = -sv.w;
      }
    }
  }
  for (int c0 = 0; c0 < so; c0 += 16 * KC) {
    bh8 y[N][KC];
#pragma unroll
    for (int n = 0; n < N; ++n)
#pragma unroll
      for (int u = 0; u < KC; ++u)
        if (c0 + 16 * u < so) y[n][u] = *(const bh8*)(own[n] + c0 + 16 * u);
#pragma unroll
    for (int u = 0; u < KC; ++u) {
      if (u % 2 == 0) asm volatile("" ::: "memory");
      if (c0 + 16 * u < so) {
        const bh8 a0 = *(const bh8*)&pan[l32][c0 + 16 * u + 8 * lh];
        const bh8 a1 = *(const bh8*)&pan[32 + l32][c0 + 16 * u + 8 * lh];
#pragma unroll
        for (int n = 0; n < N; ++n) {
          acc0[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a0, y[n][u], acc0[n], 0, 0, 0);
          acc1[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a1, y[n][u], acc1[n], 0, 0, 0);
        }
      }
    }
  }
  float x0[N][16], x1[N][16];
#pragma unroll
  for (int n = 0; n < N; ++n)
#pragma unroll
    for (int r = 0; r < 16; ++r) x0[n][r] = -acc0[n][r];
  solve_blocks<MODE, N>(x0, T.m[0], lh);
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    const int g0 = 8 * ks, g1 = 8 * ks + 4;
    const bh8 av = *(const bh8*)&T.mot[l32][16 * ks + 8 * lh];
#pragma unroll
    for (int n = 0; n < N; ++n) {
      const unsigned p0 = pack_h2(x0[n][g0], x0[n][g0 + 1]), p1 = pack_h2(x0[n][g0 + 2], x0[n][g0 + 3]);
      const unsigned q0 = pack_h2(x0[n][g1], x0[n][g1 + 1]), q1 = pack_h2(x0[n][g1 + 2], x0[n][g1 + 3]);
      const auto sa = __builtin_amdgcn_permlane32_swap(p0, q0, false, false);
      const auto sb = __builtin_amdgcn_permlane32_swap(p1, q1, false, false);
      typedef unsigned u4 __attribute__((ext_vector_type(4)));
      const u4 yb = {sa[0], sb[0], sa[1], sb[1]};
      acc1[n] = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, __builtin_bit_cast(bh8, yb), acc1[n], 0,
                                                       0, 0);
    }
  }
#pragma unroll
  for (int n = 0; n < N; ++n)
#pragma unroll
    for (int r = 0; r < 16; ++r) x1[n][r] = -acc1[n][r];
  solve_blocks<MODE, N>(x1, T.m[1], lh);
#pragma unroll
  for (int n = 0; n < N; ++n) {
    if (!ok[n]) continue;
    half_t* oc = (role ? g.uop : g.lop) + vc[n] * ldo + so + 4 * lh;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const bh4 h0 = {(_Float16)x0[n][4 * q], (_Float16)x0[n][4 * q + 1], (_Float16)x0[n][4 * q + 2],
                      (_Float16)x0[n][4 * q + 3]};
      const bh4 h1 = {(_Float16)x1[n][4 * q], (_Float16)x1[n][4 * q + 1], (_Float16)x1[n][4 * q + 2],
                      (_Float16)x1[n][4 * q + 3]};
      *(bh4*)(oc + 8 * q) = h0;
      *(bh4*)(oc + 32 + 8 * q) = h1;
      if (role) {
        half_t* ac = g.a + so + vc[n] * g.lda + 4 * lh;
        *(bh4*)(ac + 8 * q) = h0;
        *(bh4*)(ac + 32 + 8 * q) = h1;
      }
    }
    if (!role) {
      const auto ra = __builtin_amdgcn_make_buffer_rsrc((void*)(g.a + (int64_t)so * g.lda), 0,
                                                        (int)(SW * g.lda * 2), kBufWord3);
      const int vo = (int)(vc[n] + 4 * lh * g.lda) * 2;
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int eu = 8 * (r >> 2) + (r & 3);
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)x0[n][r]),
                                              ra, vo, (int)(eu * g.lda * 2), 0);
        __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, (_Float16)x1[n][r]),
                                              ra, vo, (int)((32 + eu) * g.lda * 2), 0);
      }
    }
  }
}

// the other factor's op-rounded vectors of the slice's tile into LDS:
// pan[e][k] = src[(so + e) * ldo + k], e < 64, k < so
__device__ __forceinline__ void stage_panel(_Float16 (*pan)[PP], const half_t* src, int64_t ldo,
                                            int so) {
  const int kc = so / 8;  // 16-byte chunks per row
  for (int c = threadIdx.x; c < SW * kc; c += BLT) {
    const int e = c / kc, k = 8 * (c % kc);
    *(bh8*)&pan[e][k] = *(const bh8*)(src + (int64_t)(so + e) * ldo + k);
  }
}

constexpr long long kWaitLimit = 1LL << 21;  // default polls of a published counter per wait

// stage_panel for both of the diagonal role's panels with every load of a
// thread issued before its first LDS store (so <= NBM - SW: at most
// PANEL_CHUNKS 16-byte chunks per thread and panel), one memory round trip
// on the tile chain instead of one per chunk
constexpr int PANEL_CHUNKS = (SW * (NBM - SW) / 8 + BLT - 1) / BLT;
__device__ __forceinline__ void stage_panels(_Float16 (*pan0)[PP], const half_t* src0,
                                             _Float16 (*pan1)[PP], const half_t* src1,
                                             int64_t ldo, int so) {
  const int kc = so / 8;  // 16-byte chunks per row
  const int tot = SW * kc;
  bh8 v0[PANEL_CHUNKS], v1[PANEL_CHUNKS];
#pragma unroll
  for (int u = 0; u < PANEL_CHUNKS; ++u) {
    const int c = threadIdx.x + u * BLT;
    if (c < tot) {
      const int e = c / kc, k = 8 * (c % kc);
      v0[u] = *(const bh8*)(src0 + (int64_t)(so + e) * ldo + k);
      v1[u] = *(const bh8*)(src1 + (int64_t)(so + e) * ldo + k);
    }
  }
#pragma unroll
  for (int u = 0; u < PANEL_CHUNKS; ++u) {
    const int c = threadIdx.x + u * BLT;
    if (c < tot) {
      const int e = c / kc, k = 8 * (c % kc);
      *(bh8*)&pan0[e][k] = v0[u];
      *(bh8*)&pan1[e][k] = v1[u];
    }
  }
}

template <int MODE, bool PUB>
__device__ void diagonal_role(const BlkArgs& g, DiagLds& S) {
  using Sm = Sem<MODE>;
  const int t = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(t / 64);
  const int i = t & 63, l32 = i & 31, lh = i >> 5;
  const int cb = TP * w;
  const int ns = g.nb / SW;
  for (int s = 0; s < ns; ++s) {
    const int so = SW * s;
    // 0. the slice's operands: op(U) columns of the tile (the rows of L
    // take them) and op(L) rows of the tile (the columns of U take them),
    // k < so
    stage_panels(S.pan[0], g.uop, S.pan[1], g.lop, g.ldo, so);
    __syncthreads();
    // 1. the tile's update T = C - op(L[so.., 0:so]) op(U[0:so, so..]):
    // wavefront w takes quarter (ti, tj) = w % 4 over half kh = w / 4 of
    // the k range; the halves' sums are added, then subtracted from C once
    {
      const int qd = w & 3, kh = w >> 2, ti = qd >> 1, tj = qd & 1;
      const int kl = so / 2;  // k per half (a multiple of 32)
      bf16v acc;
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[r] = 0.0f;
      for (int k0 = kh * kl; k0 < kh * kl + kl; k0 += 16) {
        const bh8 av = *(const bh8*)&S.pan[1][32 * ti + l32][k0 + 8 * lh];
        const bh8 bv = *(const bh8*)&S.pan[0][32 * tj + l32][k0 + 8 * lh];
        acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(av, bv, acc, 0, 0, 0);
      }
      const int j = 32 * tj + l32;
      if (kh == 1) {
#pragma unroll
        for (int r = 0; r < 16; ++r) S.f.tt[32 * ti + 8 * (r >> 2) + 4 * lh + (r & 3)][j] = acc[r];
      }
      __syncthreads();
      if (kh == 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int ii = 32 * ti + 8 * (r >> 2) + 4 * lh + (r & 3);
          const float c = g.pl[(so + ii) + (int64_t)(so + j) * g.ldpl];
          S.f.tt[ii][j] = Sm::apply(c, acc[r] + S.f.tt[ii][j]);
        }
      }
    }
    __syncthreads();
    float x[TP];
#pragma unroll
    for (int jj = 0; jj < TP; ++jj) x[jj] = S.f.tt[i][cb + jj];
    __syncthreads();
    // 2. the factorization (ib = 32)
    tile_factor<MODE, false, 32>(S.f.fac, x, 32, w, i);
    __syncthreads();
    const float rk = S.f.fac.R[i];
    // 3. the factored tile: fp16 factors, the slice table (LDS, then copied
    // to global memory with 16-byte stores: every entry is written below)
    SliceTab* gt = g.tabs + s;
#pragma unroll
    for (int jj = 0; jj < TP; ++jj) g.a[(so + i) + (int64_t)(so + cb + jj) * g.lda] = (half_t)x[jj];
    auto put = [&](int role, int b, int k, int e, float v) { S.tab.role[role].m[b][k][e] = v; };
    auto put_o = [&](int role, int e, int k, float v) {
      S.tab.role[role].mot[e][k] = (_Float16)Sm::op(v);
    };
#pragma unroll
    for (int jj = 0; jj < TP; ++jj) {
      const int c = cb + jj;
      // (on the diagonal: the scales, 1 / u_ii for the rows of L, 1 for the
      // columns of U)
      const float v0 = c == i ? rk : x[jj], v1 = c == i ? 1.0f : x[jj];
      if (i < 32 && c < 32) {
        put(0, 0, i, c, v0);
        put(1, 0, c, i, v1);
      } else if (i >= 32 && c >= 32) {
        put(0, 1, i - 32, c - 32, v0);
        put(1, 1, c - 32, i - 32, v1);
      } else if (i < 32) {  // c >= 32: U right of block 0
        put_o(0, c - 32, i, x[jj]);
      } else {              // i >= 32, c < 32: L below block 0
        put_o(1, i - 32, c, x[jj]);
      }
    }
    __syncthreads();
    {
      constexpr int NT = (int)(sizeof(SliceTab) / 16);
      const float4* src = (const float4*)&S.tab;
      float4* dst = (float4*)gt;
      for (int e = t; e < NT; e += BLT) dst[e] = src[e];
    }
    // (publish: every wavefront's stores have reached the L2 before the
    // barrier; the release store after it writes the L2 back once)
    if constexpr (PUB) asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    if constexpr (PUB)
      if (t == 0)
        __hip_atomic_store(g.sync + 1, g.epoch * 16 + s + 1, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
    // 4. the units inside the diagonal block, rows of L and columns of U
    // interleaved, the next tile's rows and columns (the first unit pair of
    // each kind) first. In the one-launch form only those four: the
    // helpers take the rest (blocks of 32 vectors from 2 s + 4 on) and
    // publish each block's slices, which the next-tile units wait for.
    const int nun = (g.nb - so - SW) / NV;  // units of each kind (even)
    const int nq = PUB ? std::min(4, 2 * nun) : 2 * nun;
    for (int q = w; q < nq; q += BLT / 64) {
      int role, u;
      if (q < 4) {
        role = q >> 1, u = q & 1;
      } else {
        role = q & 1, u = 2 + ((q - 4) >> 1);
      }
      if constexpr (PUB) {
        if (s > 0) {
          // block 2 s + 2 + u: its slices 0 .. s-1 came from a helper
          const int want = g.epoch * 16 + s;
          int* prog = g.sync + 4 + 16 * role + 2 * s + 2 + u;
          if (i == 0) {
            for (long long it = 0;; ++it) {
              if (__hip_atomic_load(prog, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= want)
                break;
              if (it > g.wait_limit ||
                  ((it & 255) == 0 &&
                   __hip_atomic_load(g.sync + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
                __hip_atomic_store(g.sync + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
              }
              __builtin_amdgcn_s_sleep(2);
            }
          }
          __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        }
      }
      do_unit<MODE, 14>(g, S.tab.role[role], S.pan[role], role, so + SW + NV * u, so);
    }
    __syncthreads();
  }
}

template <int MODE>
__global__ void __launch_bounds__(BLT) blockll_diag_kernel(BlkArgs g) {
  __shared__ DiagLds S;
  diagonal_role<MODE, false>(g, S);
}

// far workgroups 0 .. nfw-1 take rows of L, the others columns of U
template <int MODE, int KC>
__device__ __forceinline__ void far_body(const BlkArgs& g) {
  __shared__ _Float16 pan[SW][PP];
  __shared__ RoleTab tab;
  const int t = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(t / 64);
  const int ns = g.nb / SW;
  const int role = (int)blockIdx.x >= g.nfw;
  const int fr = (int)blockIdx.x - role * g.nfw;
  for (int s = 0; s < ns; ++s) {
    const int so = SW * s;
    {
      const float4* src = (const float4*)&g.tabs[s].role[role];
      float4* dst = (float4*)&tab;
      for (int e = t; e < (int)(sizeof(RoleTab) / 16); e += BLT) dst[e] = src[e];
    }
    stage_panel(pan, role ? g.lop : g.uop, g.ldo, so);
    __syncthreads();
    for (int u = fr * (BLT / 64) + w; u < g.nfu; u += g.nfw * (BLT / 64))
      do_unit<MODE, KC>(g, tab, pan, role, g.nb + (int64_t)NV * u, so);
    __syncthreads();
  }
}
// (at most 128 registers: two far workgroups, or one beside an update GEMM
// workgroup, fit on a CU)
template <int MODE, int KC>
__global__ void __launch_bounds__(BLT) __attribute__((amdgpu_waves_per_eu(4)))
blockll_far_kernel(BlkArgs g) {
  far_body<MODE, KC>(g);
}

// One-launch form, far role: helper hx of nhelp; its wavefronts take units
// q = (global wavefront) + j * (helper wavefronts) through every slice (both
// roles' tables and panels staged per slice): first the diagonal block's
// blocks 4 .. nb/32 - 1 of each role (role q % 2, block 4 + q / 2; block b
// while it is below the next tile, slices s <= b/2 - 2, its progress
// published after each), then the far units (role f % 2, unit f / 2 for
// f = q - 2 ndu).
template <int MODE>
__device__ void helper_role(const BlkArgs& g, DiagLds& S, int hx, int nhelp) {
  const int t = threadIdx.x;
  const int w = __builtin_amdgcn_readfirstlane(t / 64);
  const int ns = g.nb / SW;
  const int hw = nhelp * (BLT / 64), gw = hx * (BLT / 64) + w;
  const int ndu = std::max(0, g.nb / NV - 4);  // diagonal-block units per role
  const int tot = 2 * ndu + 2 * g.nfu;
  bool gave_up = false;
  for (int s = 0; s < ns; ++s) {
    const int so = SW * s;
    if (t == 0 && !gave_up) {
      const int want = g.epoch * 16 + s + 1;
      for (long long it = 0;; ++it) {
        if (__hip_atomic_load(g.sync + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= want)
          break;
        if (it > g.wait_limit ||
            ((it & 255) == 255 &&
             __hip_atomic_load(g.sync + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
          // give up (the host reports the error and the results are void)
          __hip_atomic_store(g.sync + 2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          gave_up = true;
          break;
        }
        __builtin_amdgcn_s_sleep(4);
      }
      // (one invalidation of this CU's caches for the whole workgroup)
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    {
      const float4* src = (const float4*)&g.tabs[s];
      float4* dst = (float4*)&S.tab;
      for (int e = t; e < (int)(sizeof(SliceTab) / 16); e += BLT) dst[e] = src[e];
    }
    stage_panel(S.pan[0], g.uop, g.ldo, so);
    stage_panel(S.pan[1], g.lop, g.ldo, so);
    __syncthreads();
    int q = gw;
    for (; q < tot && q < 2 * ndu; q += hw) {
      const int role = q & 1, b = 4 + (q >> 1);
      if (2 * s + 4 > b) continue;  // the diagonal role's from here on
      do_unit<MODE, 8>(g, S.tab.role[role], S.pan[role], role, (int64_t)NV * b, so);
      // (this wavefront's stores have reached the L2; the release store
      // writes it back)
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      if ((t & 63) == 0)
        __hip_atomic_store(g.sync + 4 + 16 * role + b, g.epoch * 16 + s + 1, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    // far units three or two at a time (q, q + hw, ...: hw is even, so the
    // same role)
    for (; q < tot; q += 2 * hw) {
      const int f = q - 2 * ndu, role = f & 1;
      if (q + 2 * hw < tot) {
        const int64_t v0[3] = {g.nb + (int64_t)NV * (f >> 1), g.nb + (int64_t)NV * ((f + hw) >> 1),
                               g.nb + (int64_t)NV * ((f + 2 * hw) >> 1)};
        do_units<MODE, 4, 3>(g, S.tab.role[role], S.pan[role], role, v0, so);
        q += hw;
      } else if (q + hw < tot) {
        const int64_t v0[2] = {g.nb + (int64_t)NV * (f >> 1), g.nb + (int64_t)NV * ((f + hw) >> 1)};
        do_units<MODE, 4, 2>(g, S.tab.role[role], S.pan[role], role, v0, so);
      } else {
        do_unit<MODE, 8>(g, S.tab.role[role], S.pan[role], role, g.nb + (int64_t)NV * (f >> 1),
                         so);
      }
    }
    __syncthreads();
  }
}

template <int MODE>
__global__ void __launch_bounds__(BLT) blockll_fused_kernel(BlkArgs g, int nhelp) {
  __shared__ DiagLds S;
  __shared__ int s_tk;
  if (threadIdx.x == 0)
    s_tk = __hip_atomic_fetch_add(g.sync, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - g.tbase;
  __syncthreads();
  const int tk = s_tk;
  if (tk == 0)
    diagonal_role<MODE, true>(g, S);
  else if (tk > 0 && tk <= nhelp)  // (any other ticket: counter not reset; the waits report it)
    helper_role<MODE>(g, S, tk - 1, nhelp);
}

__global__ void blockll_sync_reset_kernel(int* sync) {
  sync[threadIdx.x] = 0;  // (64 words)
}

}  // namespace

// Workspace of the block kernels for an m x m matrix and outer blocks of
// nb: op(L) rows and op(U) columns (m x nb fp16 each), the slice tables and
// the one-launch form's sync words.
size_t blockll_workspace_bytes(int64_t m, int64_t nb) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return 2 * al((size_t)m * nb * 2) + al((size_t)(nb / SW) * sizeof(SliceTab)) + 256;
}

// the sync words inside that workspace
int* blockll_sync_words(void* base, int64_t m, int64_t nb) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  return (int*)((char*)base + 2 * al((size_t)m * nb * 2) + al((size_t)(nb / SW) * sizeof(SliceTab)));
}

// whether an outer block of width in (outer blocks of nb) takes the block
// kernels: ib = 32, TC32 inner updates on fp32 buffers, 64-column slices,
// 8-byte aligned fp16 columns, buffer offsets within 31 bits
bool blockll_eligible(int ib, int I, int64_t in, int64_t nb, int64_t m, int64_t lda) {
  return ib == 32 && I == REMIFA_TC32 && in % SW == 0 && in <= NBM && nb <= NBM &&
         lda % 4 == 0 && (int64_t)SW * std::max(m, lda) * 4 < (1LL << 31);
}

// Helper workgroups of the one-launch form for an outer block of width in
// with updm rows: about blk_upw units per wavefront (automatic: 3, the
// fastest of 1 .. 8 for the n = 32768 factorization with the second
// stream). The grid stays below half the CUs: the diagonal role's waits on
// helpers that have not started end once a CU frees up.
int blockll_helpers(remifa_handle_t h, int64_t updm, int64_t in) {
  const int upw = h->blk_upw > 0 ? h->blk_upw : 3;
  const int64_t far = updm - in;
  const int64_t nfu = far > 0 ? (far + NV - 1) / NV : 0;
  const int64_t tot = 2 * std::max<int64_t>(0, in / NV - 4) + 2 * nfu;
  const int64_t nh = ceil_div(tot, (int64_t)upw * (BLT / 64));
  return (int)std::min<int64_t>(nh, std::max(1, h->num_cu / 2 - 1));
}

// CUs the block kernels hold for an outer block while they run: the
// one-launch form's workgroups (161 KB of LDS each, one per CU). The
// two-launch form is given the same figure (its far kernel holds about as
// many for part of the time), so that both plan the same update GEMMs and
// give the same bits.
int blockll_cus(remifa_handle_t h, int64_t updm, int64_t in) {
  return 1 + blockll_helpers(h, updm, in);
}

// Outer block kk: pl / pu the fp32 buffers, a the fp16 factors at the
// block's origin, in = block width (<= nb), updm = rows of the block column.
// seq: {launches so far, tickets so far} of this factorization's one-launch
// calls (both 0 before the first: its sync words are cleared then).
int blockll_launch(remifa_handle_t h, void* base, int64_t m, int64_t nb, int64_t updm,
                   int64_t in, const float* pl, int64_t ldpl, const float* pu, int64_t ldpu,
                   half_t* a, int64_t lda, int* seq) {
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  char* p = (char*)base;
  BlkArgs g;
  g.lop = (half_t*)p;
  g.uop = (half_t*)(p + al((size_t)m * nb * 2));
  g.tabs = (SliceTab*)(p + 2 * al((size_t)m * nb * 2));
  g.ldo = nb;
  g.updm = updm;
  g.nb = (int)in;
  g.pl = pl;
  g.ldpl = ldpl;
  g.pu = pu;
  g.ldpu = ldpu;
  g.a = a;
  g.lda = lda;
  const int64_t far = updm - in;
  g.nfu = far > 0 ? (int)((far + NV - 1) / NV) : 0;
  // far workgroups of each role: one unit per wavefront and slice
  g.nfw = (g.nfu + BLT / 64 - 1) / (BLT / 64);
  const double bytes = 2.0 * (double)updm * in * (4 + 2 + 2);
  g.sync = blockll_sync_words(base, m, nb);
  g.wait_limit = h->blk_wait_polls > 0 ? h->blk_wait_polls : kWaitLimit;
  if (h->blk_fused) {
    const int nhelp = blockll_helpers(h, updm, in);
    if (seq[0] == 0) blockll_sync_reset_kernel<<<1, 64, 0, h->stream>>>(g.sync);
    g.epoch = ++seq[0];
    g.tbase = seq[1];
    seq[1] += 1 + nhelp;
    timer_begin(h, KC_PANEL, bytes, 4);
    blockll_fused_kernel<SM_F32_TC><<<1 + nhelp, BLT, 0, h->stream>>>(g, nhelp);
    RF_HIP(hipGetLastError());
    timer_end(h);
    return 0;
  }
  g.tbase = 0;
  g.epoch = 0;
  timer_begin(h, KC_PANEL, bytes, 2);
  blockll_diag_kernel<SM_F32_TC><<<1, BLT, 0, h->stream>>>(g);
  RF_HIP(hipGetLastError());
  timer_end(h);
  if (g.nfu > 0) {
    timer_begin(h, KC_PANEL, bytes, 3);
    blockll_far_kernel<SM_F32_TC, 8><<<2 * g.nfw, BLT, 0, h->stream>>>(g);
    RF_HIP(hipGetLastError());
    timer_end(h);
  }
  return 0;
}

}  // namespace remifa
