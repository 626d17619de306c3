// Batched no-pivot LU and solve of many independent small matrices (orders
// up to 512), fp32 storage with matrix-core updates (TC32: fp16 operands,
// fp32 accumulation; BF16_TC32: bf16 operands).
//
// One persistent launch: a fixed grid of 8-wavefront workgroups takes
// matrices from an atomic queue (callers order the batch largest first, so
// the queue is a longest-processing-time schedule). A workgroup factors its
// matrix right-looking with 32-wide panels, which is the reference's RL
// driver with nb = ib = 32 (src/lu_nopiv.hxx:53-208): per panel
//  1. the 32 pivots of the panel: lane i keeps row i of the panel in
//     registers; wavefront 0 factors the top 32 x 32 block (pivot rows by
//     readlane), then every lane takes its row below against U11 from LDS,
//     pivots in order: reciprocal, multiply, fused update in fp32
//     (src/lu_nopiv.cu:123-170), two barriers per panel;
//  2. the block row U12 = L11^-1 A12, one lane per column;
//  3. the Schur update A22 -= L21 U12 on the matrix cores from fp16 (bf16)
//     copies of L21 and U12 staged in LDS, products summed from zero and
//     subtracted once (the update GEMMs of gemm.hip do the same); four
//     32 x 32 tiles of C per wavefront in flight together.
// The matrix stays in HBM (the L2 / last-level cache holds it between
// panels); every element of the trailing matrix is read and written once per
// panel.
//
// The batched solve runs one workgroup per matrix: x in LDS, 32-row blocks,
// the 32x32 diagonal block substitution inside one wavefront, the rest of
// each block column as a matrix-vector update by all lanes.
#include "common.hpp"

namespace remifa {
namespace {

constexpr int PW = 32;     // panel width
constexpr int BMAX = 512;  // largest order
constexpr int LP = PW + 8; // fp16 row stride of the staged operands (16-byte rows)

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

// LDS of a workgroup for orders up to BM
template <int BM>
struct BatchLds {
  uint16_t Lh[BM][LP];    // L21 (rows below the panel's top block), fp16/bf16 bits
  uint16_t Ut[BM][LP];    // U12 transposed: [column][k]
  __attribute__((aligned(16))) float L11[PW][PW];   // the panel's top block (unit L and U)
  __attribute__((aligned(16))) float L11t[PW][PW];  // its transpose
  float Rk[PW];           // reciprocals of the panel's pivots
  int item;
  int info;
};

template <bool BF>
__device__ __forceinline__ uint16_t op_bits(float v) {
  if constexpr (BF)
    return f32_to_bf16_bits(v);
  else
    return __builtin_bit_cast(uint16_t, (half_t)v);
}

// BM: largest order taken (a size bucket), BT: threads (BT >= BM: one lane
// per row of a panel and per column of a block row)
template <bool BF, int BM, int BT>
__global__ void __launch_bounds__(BT) batched_lu_kernel(int count, const int* __restrict__ d_n,
                                                        const int* __restrict__ d_ld,
                                                        float* const* __restrict__ d_ptrs,
                                                        int* __restrict__ d_info,
                                                        int* __restrict__ queue) {
  static_assert(BT >= BM && BT >= 64 && BT % 64 == 0, "one lane per row");
  __shared__ BatchLds<BM> S;
  const int t = threadIdx.x;
  const int lane = t & 63, wave = t >> 6;
  const int l32 = lane & 31, lh = lane >> 5;
  for (;;) {
    if (t == 0) {
      S.item = __hip_atomic_fetch_add(queue, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      S.info = 0;
    }
    __syncthreads();
    const int item = S.item;
    if (item >= count) break;  // every workgroup reaches this: the grid drains
    const int n = d_n[item];
    const int64_t ld = d_ld[item];
    float* __restrict__ a = d_ptrs[item];
    if (n <= 0 || n > BM) {
      if (t == 0 && d_info) d_info[item] = n > BM ? -1 : 0;
      __syncthreads();
      continue;
    }
    for (int k0 = 0; k0 < n; k0 += PW) {
      const int pw = n - k0 < PW ? n - k0 : PW;
      const int mp = n - k0;  // rows of the panel
      // ---- 1. the panel's pivots, row t in registers
      float x[PW];
      {
        const int r = t < mp ? t : mp - 1;
#pragma unroll
        for (int j = 0; j < PW; ++j) {
          const int jc = j < pw ? j : pw - 1;
          const float v = a[(k0 + r) + (int64_t)(k0 + jc) * ld];
          x[j] = (t < mp && j < pw) ? v : (t == j ? 1.0f : 0.0f);
        }
      }
      // (a) the top pw x pw block inside wavefront 0: lane i < pw holds row
      // i; pivot kk's row comes to every lane by readlane
      if (wave == 0) {
#pragma unroll
        for (int kk = 0; kk < PW; ++kk) {
          if (kk < pw) {
            float u[PW];
#pragma unroll
            for (int j = kk; j < PW; ++j)
              u[j] = __builtin_bit_cast(float, __builtin_amdgcn_readlane(
                                                   __builtin_bit_cast(int, x[j]), kk));
            const float p = u[kk];
            const float rk = 1.0f / p;
            if (t == 0) {
              S.Rk[kk] = rk;
              if ((p == 0.0f || !__builtin_isfinite(p)) && S.info == 0) S.info = k0 + kk + 1;
            }
            if (t > kk && t < PW && t < mp) {
              const float lk = x[kk] * rk;
              x[kk] = lk;
#pragma unroll
              for (int j = kk + 1; j < PW; ++j) x[j] = __builtin_fmaf(-lk, u[j], x[j]);
            }
          }
        }
        if (t < PW) {
#pragma unroll
          for (int j = 0; j < PW; ++j) S.L11[t][j] = x[j];
#pragma unroll
          for (int j = 0; j < PW; ++j) S.L11t[j][t] = x[j];
        }
      }
      __syncthreads();
      // (b) every row below it against U11, pivots in order (the same
      // operations as the pivot-by-pivot elimination)
      if (t >= PW && t < mp) {
#pragma unroll
        for (int kk = 0; kk < PW; ++kk) {
          if (kk < pw) {
            // row kk of U11 as 16-byte broadcast reads
            float ur[PW];
#pragma unroll
            for (int j = (kk + 1) / 4 * 4; j < PW; j += 4) {
              const float4 v = *(const float4*)&S.L11[kk][j];
              ur[j] = v.x, ur[j + 1] = v.y, ur[j + 2] = v.z, ur[j + 3] = v.w;
            }
            const float lk = x[kk] * S.Rk[kk];
            x[kk] = lk;
#pragma unroll
            for (int j = kk + 1; j < PW; ++j) x[j] = __builtin_fmaf(-lk, ur[j], x[j]);
          }
        }
      }
      // store the panel; keep L11/U11 (fp32) and L21 (operand precision)
      if (t < mp) {
#pragma unroll
        for (int j = 0; j < PW; ++j)
          if (j < pw) a[(k0 + t) + (int64_t)(k0 + j) * ld] = x[j];
      }
      const int mr = n - k0 - pw;  // rows (and columns) of the trailing matrix
      const int mr32 = (mr + 31) & ~31;
      if (t >= PW && t - PW < mr32) {
        const int r = t - PW;
        u16x8 v[PW / 8];
#pragma unroll
        for (int j = 0; j < PW; ++j) v[j / 8][j % 8] = r < mr ? op_bits<BF>(x[j]) : (uint16_t)0;
#pragma unroll
        for (int q = 0; q < PW / 8; ++q) *(u16x8*)&S.Lh[r][8 * q] = v[q];
      }
      __syncthreads();
      if (mr <= 0) continue;
      // ---- 2. U12 = L11^-1 A12, one lane per column
      if (t < mr32) {
        const int c = t < mr ? t : mr - 1;
        float* col = a + k0 + (int64_t)(k0 + pw + c) * ld;
        float u[PW];
#pragma unroll
        for (int i = 0; i < PW; ++i) u[i] = col[i];
#pragma unroll
        for (int kk = 0; kk < PW; ++kk) {
          {
            // column kk of L11 (row kk of its transpose), 16-byte reads
            float lc[PW];
#pragma unroll
            for (int i = (kk + 1) / 4 * 4; i < PW; i += 4) {
              const float4 v = *(const float4*)&S.L11t[kk][i];
              lc[i] = v.x, lc[i + 1] = v.y, lc[i + 2] = v.z, lc[i + 3] = v.w;
            }
#pragma unroll
            for (int i = kk + 1; i < PW; ++i) u[i] = __builtin_fmaf(-lc[i], u[kk], u[i]);
          }
        }
        if (t < mr) {
#pragma unroll
          for (int i = 0; i < PW; ++i) col[i] = u[i];
        }
        u16x8 v[PW / 8];
#pragma unroll
        for (int i = 0; i < PW; ++i) v[i / 8][i % 8] = t < mr ? op_bits<BF>(u[i]) : (uint16_t)0;
#pragma unroll
        for (int q = 0; q < PW / 8; ++q) *(u16x8*)&S.Ut[t][8 * q] = v[q];
      }
      __syncthreads();
      // ---- 3. A22 -= L21 U12 on the matrix cores, 32 x 32 tiles, up to four
      // per wavefront at a time (their C entries read in one round trip)
      const int nt = mr32 / 32;
      constexpr int TQ = 4;
      for (int tile0 = wave; tile0 < nt * nt; tile0 += TQ * (BT / 64)) {
        f16v acc[TQ];
        float cv[TQ][16];
#pragma unroll
        for (int z = 0; z < TQ; ++z) {
          const int tile = tile0 + z * (BT / 64);
          const int tc = tile < nt * nt ? tile : 0;
          const int tm = tc % nt, tn = tc / nt;
          const int m = tm * 32 + l32;
          const float* cm = a + (k0 + pw + (m < mr ? m : 0)) + (int64_t)(k0 + pw) * ld;
#pragma unroll
          for (int q = 0; q < 16; ++q) {
            const int nn = tn * 32 + 4 * lh + 8 * (q >> 2) + (q & 3);
            cv[z][q] = cm[(int64_t)(nn < mr ? nn : 0) * ld];
          }
#pragma unroll
          for (int q = 0; q < 16; ++q) acc[z][q] = 0.0f;
#pragma unroll
          for (int ks = 0; ks < PW; ks += 16) {
            const u16x8 xb = *(const u16x8*)&S.Ut[tn * 32 + l32][ks + 8 * lh];
            const u16x8 ya = *(const u16x8*)&S.Lh[tm * 32 + l32][ks + 8 * lh];
            if constexpr (BF)
              acc[z] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                  __builtin_bit_cast(bf8, xb), __builtin_bit_cast(bf8, ya), acc[z], 0, 0, 0);
            else
              acc[z] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                  __builtin_bit_cast(h8, xb), __builtin_bit_cast(h8, ya), acc[z], 0, 0, 0);
          }
        }
        // lane holds C[m = tm*32 + l32][n = tn*32 + 4 lh + 8 (q / 4) + q % 4]
#pragma unroll
        for (int z = 0; z < TQ; ++z) {
          const int tile = tile0 + z * (BT / 64);
          const int tm = tile % nt, tn = tile / nt;
          const int m = tm * 32 + l32;
          if (tile < nt * nt && m < mr) {
            float* cm = a + (k0 + pw + m) + (int64_t)(k0 + pw) * ld;
#pragma unroll
            for (int q = 0; q < 16; ++q) {
              const int nn = tn * 32 + 4 * lh + 8 * (q >> 2) + (q & 3);
              if (nn < mr) cm[(int64_t)nn * ld] = cv[z][q] - acc[z][q];
            }
          }
        }
      }
      __syncthreads();  // the trailing matrix is complete before the next panel
    }
    if (t == 0 && d_info) d_info[item] = S.info;
    __syncthreads();
  }
}

// One matrix per workgroup: b <- (L U)^-1 b with the factors in place.
__global__ void __launch_bounds__(256) batched_getrs_kernel(int count, const int* __restrict__ d_n,
                                                            const int* __restrict__ d_ld,
                                                            float* const* __restrict__ d_ptrs,
                                                            float* const* __restrict__ d_b) {
  __shared__ float xs[BMAX];
  const int item = blockIdx.x;
  const int n = d_n[item];
  if (n <= 0 || n > BMAX) return;  // uniform: the whole workgroup leaves
  const int64_t ld = d_ld[item];
  const float* __restrict__ a = d_ptrs[item];
  float* __restrict__ b = d_b[item];
  const int t = threadIdx.x;
  for (int i = t; i < n; i += 256) xs[i] = b[i];
  __syncthreads();
  // forward: unit lower L
  for (int k0 = 0; k0 < n; k0 += 32) {
    const int kw = n - k0 < 32 ? n - k0 : 32;
    if (t < 64) {
      // the diagonal block inside one wavefront (lanes 0..kw-1 own its rows;
      // the lane's row of the block is read before the substitution)
      const int tr = t < kw ? t : 0;
      float rowv[32];
#pragma unroll
      for (int k = 0; k < 32; ++k) rowv[k] = a[(k0 + tr) + (int64_t)(k0 + (k < kw ? k : 0)) * ld];
      float xi = t < kw ? xs[k0 + t] : 0.0f;
#pragma unroll
      for (int k = 0; k < 32; ++k) {
        if (k < kw) {
          const float xk = __shfl(xi, k, 64);
          if (t > k && t < kw) xi = __builtin_fmaf(-rowv[k], xk, xi);
        }
      }
      if (t < kw) xs[k0 + t] = xi;
    }
    __syncthreads();
    for (int i = k0 + kw + t; i < n; i += 256) {
      float s = xs[i];
      for (int k = 0; k < kw; ++k) s = __builtin_fmaf(-a[i + (int64_t)(k0 + k) * ld], xs[k0 + k], s);
      xs[i] = s;
    }
    __syncthreads();
  }
  // backward: upper U
  for (int k1 = n; k1 > 0; k1 -= 32) {
    const int k0 = k1 - 32 > 0 ? k1 - 32 : 0;
    const int kw = k1 - k0;
    if (t < 64) {
      const int tr = t < kw ? t : 0;
      float rowv[32];
#pragma unroll
      for (int k = 0; k < 32; ++k) rowv[k] = a[(k0 + tr) + (int64_t)(k0 + (k < kw ? k : 0)) * ld];
      float xi = t < kw ? xs[k0 + t] : 0.0f;
#pragma unroll
      for (int k = 31; k >= 0; --k) {
        if (k < kw) {
          if (t == k) xi = xi / rowv[k];
          const float xk = __shfl(xi, k, 64);
          if (t < k) xi = __builtin_fmaf(-rowv[k], xk, xi);
        }
      }
      if (t < kw) xs[k0 + t] = xi;
    }
    __syncthreads();
    for (int i = t; i < k0; i += 256) {
      float s = xs[i];
      for (int k = 0; k < kw; ++k) s = __builtin_fmaf(-a[i + (int64_t)(k0 + k) * ld], xs[k0 + k], s);
      xs[i] = s;
    }
    __syncthreads();
  }
  for (int i = t; i < n; i += 256) b[i] = xs[i];
}

__global__ void zero_queue_kernel(int* q) {
  if (threadIdx.x == 0) q[0] = 0;
}

}  // namespace
}  // namespace remifa

using namespace remifa;

namespace remifa {
void batched_shape(remifa_handle_t h, int bm, int* bt, int* per_cu);
}

extern "C" int remifa_lu_nopiv_batched(remifa_handle_t h, int ctype, int64_t count, const int* d_n,
                                       const int* d_ld, float* const* d_ptrs, int* d_info,
                                       int max_n) {
  if (!h) return -1;
  if (ctype != REMIFA_TC32 && ctype != REMIFA_BF16_TC32) {
    set_last_error("remifa_lu_nopiv_batched: update types TC32 and BF16_TC32 are built");
    return REMIFA_ERR_UNSUPPORTED;
  }
  if (count < 0 || count > INT32_MAX) return -3;
  if (count > 0 && (!d_n || !d_ld || !d_ptrs)) return -4;
  if (max_n < 0 || max_n > BMAX) return -8;
  if (count == 0) return 0;
  // the queue counter lives in the handle's workspace; it is reset by a
  // kernel (not a memset), so the call can be captured in a graph
  RF_CHECK(ensure_workspace(h, 256));
  int* queue = (int*)h->ws;
  zero_queue_kernel<<<1, 64, 0, h->stream>>>(queue);
  // the size bucket of max_n picks the kernel (LDS for orders up to the
  // bucket's bound) and, from the handle's tuning table, its threads and
  // persistent workgroups per CU
  int bm = 512;
  for (int b : {64, 128, 256})
    if (max_n <= b) {
      bm = b;
      break;
    }
  int bt = 0, per_cu = 0;
  batched_shape(h, bm, &bt, &per_cu);
  const int grid = per_cu * h->num_cu;
  timer_begin(h, KC_OTHER, 0.0, 1);
  int st = 0;
#define RF_BATCHED(BM, BT)                                                                       \
  if (bm == BM && bt == BT) {                                                                    \
    if (ctype == REMIFA_TC32)                                                                    \
      batched_lu_kernel<false, BM, BT>                                                           \
          <<<grid, BT, 0, h->stream>>>((int)count, d_n, d_ld, d_ptrs, d_info, queue);            \
    else                                                                                         \
      batched_lu_kernel<true, BM, BT>                                                            \
          <<<grid, BT, 0, h->stream>>>((int)count, d_n, d_ld, d_ptrs, d_info, queue);            \
  } else
  RF_BATCHED(64, 64)
  RF_BATCHED(64, 128)
  RF_BATCHED(128, 128)
  RF_BATCHED(128, 256)
  RF_BATCHED(256, 256)
  RF_BATCHED(256, 512)
  RF_BATCHED(512, 512)
  st = REMIFA_ERR_UNSUPPORTED;
#undef RF_BATCHED
  timer_end(h);
  if (st) {
    set_last_error("remifa_lu_nopiv_batched: no kernel for this bucket / thread count");
    return st;
  }
  RF_HIP(hipGetLastError());
  return 0;
}

// Threads and workgroups per CU of the batched factorization for orders up
// to bm: the handle's table (remifa_set_batched_tuning) or the defaults.
void remifa::batched_shape(remifa_handle_t h, int bm, int* bt, int* per_cu) {
  *bt = bm < 128 ? 128 : (bm < 512 ? 256 : 512);
  *per_cu = bm <= 64 ? 4 : (bm <= 128 ? 4 : (bm <= 256 ? 2 : 1));
  for (const auto& r : h->batched_rules)
    if (r[0] == bm) {
      *bt = r[1];
      *per_cu = r[2];
    }
}

extern "C" int remifa_getrs_nopiv_batched(remifa_handle_t h, int64_t count, const int* d_n,
                                          const int* d_ld, float* const* d_ptrs,
                                          float* const* d_b, int max_n) {
  if (!h) return -1;
  if (count < 0 || count > INT32_MAX) return -2;
  if (count > 0 && (!d_n || !d_ld || !d_ptrs || !d_b)) return -3;
  if (max_n < 0 || max_n > BMAX) return -7;
  if (count == 0) return 0;
  timer_begin(h, KC_OTHER, 0.0, 2);
  batched_getrs_kernel<<<(unsigned)count, 256, 0, h->stream>>>((int)count, d_n, d_ld, d_ptrs, d_b);
  timer_end(h);
  RF_HIP(hipGetLastError());
  return 0;
}

// Batched tuning table: rules[3 i .. 3 i + 2] = (bucket bound, threads,
// workgroups per CU), bucket bound in {64, 128, 256, 512}, threads a
// multiple of 64 between the bound and twice it (at most 512), 1..16
// workgroups per CU. count = 0 restores the defaults.
extern "C" int remifa_set_batched_tuning(remifa_handle_t h, int count, const int* rules) {
  if (!h) return -1;
  if (count < 0 || count > 64) return -2;
  if (count > 0 && !rules) return -3;
  std::vector<std::array<int, 3>> r((size_t)count);
  for (int i = 0; i < count; ++i) {
    const int bm = rules[3 * i], bt = rules[3 * i + 1], pc = rules[3 * i + 2];
    const bool bm_ok = bm == 64 || bm == 128 || bm == 256 || bm == 512;
    const bool bt_ok = (bm == 64 && (bt == 64 || bt == 128)) ||
                       (bm == 128 && (bt == 128 || bt == 256)) ||
                       (bm == 256 && (bt == 256 || bt == 512)) || (bm == 512 && bt == 512);
    if (!bm_ok || !bt_ok || pc < 1 || pc > 16) return -3;
    r[i] = {bm, bt, pc};
  }
  h->batched_rules.swap(r);
  return 0;
}
