// Left-looking and right-looking no-pivot LU drivers.
//
// LL (src/lu_nopiv_ll.hxx:55-582 with src/lu_nopiv_inner_ll.hxx:37-372) and
// RL (src/lu_nopiv.hxx:53-208 with src/lu_nopiv_inner_rl.hxx:34-198) keep
// the reference's outer loops and rounding points: for every outer block of
// nb columns an update GEMM (LL: block column and block row with K = ofs;
// RL: the Schur complement with K = nb after the block) and the inner
// factorization of the block with its ib-granular semantics.
//
// The inner factorization runs over slices of pw <= 64 columns (pw a multiple
// of ib, default 64): each slice is one fused kernel (subpanel.hip) that does
// the reference's ib loop inside, and the updates between slices are matrix-
// core GEMMs (LL: the slice's columns/rows minus everything left of it, K =
// slice offset; RL: a rank-pw update after it). pw = ib reproduces the
// reference's structure of one panel per ib columns.
//
// Differences from the reference, all deliberate:
//  * no allocation inside the loop: one workspace arena per call (slice
//    table, fp32 panel buffers, fp16 copy when none is passed, split-K partials);
//  * the precision tuple is a runtime argument (no template explosion);
//  * d_info is 0 or the 1-based column of the first zero / non-finite pivot
//    (the reference writes the last panel width, src/lu_nopiv.cu:173);
//  * the fp16 copy of the factors (TC updates on fp32 storage) is written by
//    the slice kernel as it produces the factors, not by separate converts.
#include <algorithm>

#include "common.hpp"

namespace remifa {

int subpanel_launch(remifa_handle_t h, int T, int I, bool rl, int ib, int64_t cm, int sw, void* l,
                    int64_t ldl, void* u, int64_t ldu, half_t* cp, int64_t ldcp, float* tab);
size_t subpanel_table_bytes();
size_t blockll_workspace_bytes(int64_t m, int64_t nb);
int* blockll_sync_words(void* base, int64_t m, int64_t nb);
int blockll_cus(remifa_handle_t h, int64_t updm, int64_t in);
bool blockll_eligible(int ib, int I, int64_t in, int64_t nb, int64_t m, int64_t lda);
int blockll_launch(remifa_handle_t h, void* base, int64_t m, int64_t nb, int64_t updm,
                   int64_t in, const float* pl, int64_t ldpl, const float* pu, int64_t ldpu,
                   half_t* a, int64_t lda, int* seq);

namespace {

__global__ void info_kernel(int64_t n, const void* a, int64_t lda, int t, int* info) {
  // one workgroup: first column k with u_kk == 0 or not finite (1-based)
  __shared__ int first;
  if (threadIdx.x == 0) first = 0x7fffffff;
  __syncthreads();
  for (int64_t k = threadIdx.x; k < n; k += blockDim.x) {
    float v = (t == REMIFA_R_16F) ? (float)((const half_t*)a)[k + k * lda]
                                  : ((const float*)a)[k + k * lda];
    if (v == 0.0f || !__builtin_isfinite(v)) atomicMin(&first, (int)(k + 1));
  }
  __syncthreads();
  if (threadIdx.x == 0) *info = first == 0x7fffffff ? 0 : first;
}

__global__ void zero_int_kernel(int* p, int n) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = 0;
}

bool is_h16acc(int c) { return c == REMIFA_FP16 || c == REMIFA_TC16; }

// storage type of the operands a compute type reads
int operand_type(int c) { return c == REMIFA_FP32 ? REMIFA_R_32F : REMIFA_R_16F; }

struct Arena {
  char* base;
  size_t off = 0;
  void* take(size_t bytes) {
    void* p = base + off;
    off += (bytes + 255) & ~(size_t)255;
    return p;
  }
};

// split-K partials of the update GEMMs: up to 16 slices of the largest
// block-column update, at most 256 MiB
size_t gemm_reserve(int64_t m, int64_t nb) {
  const size_t full = (size_t)16 * 4 * (size_t)m * (size_t)nb;
  return std::min<size_t>((size_t)256 << 20, std::max<size_t>((size_t)16 << 20, full));
}

struct Ctx {
  remifa_handle_t h;
  remifa_lu_config_t c;
  int64_t ib;
  int pw;      // slice width of the inner factorization
  float* tab;  // the slice's published diagonal tile
};

int gemm(Ctx& x, int ctype, int out_t, int64_t m, int64_t n, int64_t k, const void* a,
         int64_t lda, const void* b, int64_t ldb, void* c, int64_t ldc) {
  if (m <= 0 || n <= 0 || k <= 0) return 0;
  return gemm_launch(x.h, ctype, operand_type(ctype), out_t, m, n, k, -1.0f, a, lda, b, ldb, 1.0f,
                     c, ldc);
}

// element pointer arithmetic on untyped buffers
inline void* at(void* p, int t, int64_t i, int64_t j, int64_t ld) {
  return (char*)p + (i + j * ld) * (int64_t)dtype_size(t);
}

// Inner left-looking factorization of one outer block: m x n block column
// `l` and n x m block row `u` (columns n.. beyond the diagonal block), both
// in precision T (l == u when factoring in place). cpy/ldcp: fp16 mirror of
// this block's factors; when T = fp32 and the inner update is a TC type its
// operands are read from the mirror (the reference's d_a_cpy,
// src/lu_nopiv_inner_ll.hxx:225-265). `mirror`: write the factors to cpy.
int inner_ll(Ctx& x, int T, int64_t m, int64_t n, void* l, int64_t ldl, void* u, int64_t ldu,
             void* cpy, int64_t ldcp, bool mirror) {
  const int I = x.c.I;
  const bool tc_from_copy = (T == REMIFA_R_32F) && I != REMIFA_FP32;
  if (tc_from_copy) mirror = true;
  for (int64_t so = 0; so < n; so += x.pw) {
    const int sw = (int)std::min<int64_t>(x.pw, n - so);
    const int64_t cm = m - so;
    if (so > 0) {
      const void* la = tc_from_copy ? at(cpy, REMIFA_R_16F, so, 0, ldcp) : at(l, T, so, 0, ldl);
      const int64_t lla = tc_from_copy ? ldcp : ldl;
      const void* ub = tc_from_copy ? at(cpy, REMIFA_R_16F, 0, so, ldcp) : at(u, T, 0, so, ldu);
      const int64_t lub = tc_from_copy ? ldcp : ldu;
      // columns of the slice (:223-246) and its rows right of the diagonal
      // tile (:249-301): one launch when both take the narrow tiles
      if (cm > sw) {
        const void* ub2 = tc_from_copy ? at(cpy, REMIFA_R_16F, 0, so + sw, ldcp)
                                       : at(u, T, 0, so + sw, ldu);
        RF_CHECK(gemm_pair_launch(x.h, I, operand_type(I), T, cm, sw, so, la, lla, ub, lub,
                                  at(l, T, so, so, ldl), ldl, sw, cm - sw, la, lla, ub2, lub,
                                  at(u, T, so, so + sw, ldu), ldu));
      } else {
        RF_CHECK(gemm(x, I, T, cm, sw, so, la, lla, ub, lub, at(l, T, so, so, ldl), ldl));
      }
    }
    RF_CHECK(subpanel_launch(x.h, T, I, false, (int)x.ib, cm, sw, at(l, T, so, so, ldl), ldl,
                             at(u, T, so, so, ldu), ldu,
                             mirror ? (half_t*)at(cpy, REMIFA_R_16F, so, so, ldcp) : nullptr, ldcp,
                             x.tab));
  }
  return 0;
}

int validate(const remifa_lu_config_t& c, int algo) {
  auto prec_ok = [](int p) { return p == REMIFA_R_32F || p == REMIFA_R_16F; };
  auto ct_ok = [](int t) { return t >= REMIFA_FP32 && t <= REMIFA_FP8_TC32; };
  if (!prec_ok(c.W) || !ct_ok(c.I) || !ct_ok(c.O)) return 1;
  if (!prec_ok(c.OP) || !prec_ok(c.IP) || !prec_ok(c.OF) || !prec_ok(c.IF)) return 1;
  if (c.ib != 8 && c.ib != 16 && c.ib != 32 && c.ib != 64) return 1;
  if (c.W == REMIFA_R_32F) {
    // src/lu_nopiv_inner_ll.hxx:74-100: fp32 storage needs fp32 panels/factors
    if (c.OP != REMIFA_R_32F || c.IP != REMIFA_R_32F || c.OF != REMIFA_R_32F ||
        c.IF != REMIFA_R_32F)
      return 1;
    if (is_h16acc(c.I) || is_h16acc(c.O)) return 1;  // FP16/TC16 need fp16 storage
  } else {
    if (c.I == REMIFA_FP32 && c.IP != REMIFA_R_32F) return 1;  // :102-104
    if (c.O == REMIFA_FP32) return 1;  // no fp32 GEMM on fp16 operands
    if (algo == REMIFA_LL) {
      // supported layouts: everything fp16 (in place) or fp32 inner buffers
      bool inplace = c.IP == REMIFA_R_16F && c.IF == REMIFA_R_16F && c.OF == REMIFA_R_16F;
      bool buffered = c.IP == REMIFA_R_32F && c.IF == REMIFA_R_32F && c.OF == c.OP;
      if (!inplace && !buffered) return 1;
      if (is_h16acc(c.O) && c.OP != REMIFA_R_16F) return 1;
      if (buffered && is_h16acc(c.I)) return 1;
    }
  }
  // bf16 / fp8 updates read operands rounded from the fp16 storage
  auto lowp = [](int t) { return t == REMIFA_BF16_TC32 || t == REMIFA_FP8_TC32; };
  if (c.W == REMIFA_R_32F && (lowp(c.I) || lowp(c.O))) return 1;
  // fp8 operands only for the outer updates (one scale per 32-block of
  // the factors' rows / columns; the inner updates keep I)
  if (c.I == REMIFA_FP8_TC32) return 1;
  return 0;
}

}  // namespace

// Workspace needed by lu_nopiv (bytes).
// FP8 outer updates, LL: e4m3 mirrors of the finished factors, written once
// per outer block: L8[r * ld8 + k] (row r of L, K-contiguous) and
// U8[j * ld8 + k] (column j of U), scale bytes SL / SU [. * lds8 + k / 32].
// RL quantizes each Schur update's operands into the GEMM workspace.
struct Fp8Mirror {
  int64_t ld8 = 0, lds8 = 0;
  uint8_t *l8 = nullptr, *sl = nullptr, *u8 = nullptr, *su = nullptr;
  static int64_t ld_codes(int64_t n) { return ceil_div(n, 128) * 128; }
  static size_t bytes(int64_t m, int64_t n) {
    const int64_t ld8 = ld_codes(n), lds8 = ld8 / 32;
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    return al((size_t)m * ld8) + al((size_t)m * lds8) + al((size_t)n * ld8) + al((size_t)n * lds8);
  }
};

// LL lookahead layout: fp16 storage, fp32 inner buffers and panels, 16-bit
// outer operands (TC32 from A itself or BF16_TC32 from its bf16 mirror),
// at least two outer blocks
// (FP8_TC32 when nb is a multiple of 128: block k's part of the update
// starts its K range at a multiple of 128, where the scale words stay
// 4-byte aligned)
bool lookahead_layout(int algo, const remifa_lu_config_t& c, int64_t n, int64_t nb) {
  return algo == REMIFA_LL && c.W == REMIFA_R_16F && c.IP == REMIFA_R_32F &&
         c.OP == REMIFA_R_32F && (c.O != REMIFA_FP8_TC32 || nb % 128 == 0) && n > nb;
}
// LL lookahead on the factors in place, fp32 storage with fp32-accumulating
// outer updates: the update of block k+1 by blocks 0..k-1 writes only block
// k+1's column and row of A (and reads factors left of / above block k),
// block k's factorization only block k's column and row. The update then
// reaches A in two parts, one fp32 rounding more; with fp16 storage or fp16
// accumulation that extra rounding would be an fp16 one, so those keep the
// single update.
bool lookahead_inplace(int algo, const remifa_lu_config_t& c, int64_t n, int64_t nb) {
  return algo == REMIFA_LL && c.W == REMIFA_R_32F &&
         (c.O == REMIFA_FP32 || c.O == REMIFA_TC32 || c.O == REMIFA_BF16_TC32) && n > nb;
}

// The lookahead's second stream (default priority), masked to leave
// h->side_reserve CUs to the handle's stream. The mask bits i = x + 8 y
// reserved are y = 4 ((x + t) % 8), t < reserve / 8: reserve / 8 CUs of
// every class i % 8 and of every run of 32 bits, so every XCD keeps the
// same share whether the driver numbers CUs XCD by XCD or round-robin.
int side_stream(remifa_handle_t h) {
  if (h->side && h->side_reserve_made == h->side_reserve) return 0;
  if (h->side) {
    RF_HIP(hipStreamSynchronize(h->side));
    RF_HIP(hipStreamDestroy(h->side));
    h->side = nullptr;
  }
  if (!h->ev_main) RF_HIP(hipEventCreateWithFlags(&h->ev_main, hipEventDisableTiming));
  if (!h->ev_side) RF_HIP(hipEventCreateWithFlags(&h->ev_side, hipEventDisableTiming));
  const int ncu = h->num_cu;
  const int r = h->side_reserve;
  if (r == 0 || ncu != 256) {
    RF_HIP(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
  } else {
    uint32_t mask[8];
    for (auto& w : mask) w = 0xFFFFFFFFu;
    for (int x = 0; x < 8; ++x)
      for (int t = 0; t < r / 8; ++t) {
        const int i = x + 8 * 4 * ((x + t) % 8);
        mask[i / 32] &= ~(1u << (i % 32));
      }
    RF_HIP(hipExtStreamCreateWithCUMask(&h->side, 8, mask));
  }
  h->side_reserve_made = r;
  return 0;
}

size_t lu_workspace_bytes(int algo, const remifa_lu_config_t& c, int64_t m, int64_t n, int64_t nb,
                          bool have_cpy) {
  size_t b = (subpanel_table_bytes() + 255) & ~(size_t)255;  // slice table
  if (algo == REMIFA_LL && c.O == REMIFA_BF16_TC32 && c.W == REMIFA_R_16F &&
      c.IP == REMIFA_R_32F)
    b += ((size_t)m * n * 2 + 255) & ~(size_t)255;  // bf16 mirror of the factors
  if (c.O == REMIFA_FP8_TC32) {
    if (algo == REMIFA_LL)
      b += Fp8Mirror::bytes(m, n);
    else  // quantized Schur operands (m + n rows of K = nb), taken by the GEMM
      b += (size_t)(m + n) * (size_t)(Fp8Mirror::ld_codes(nb) * 33 / 32) + 4096;
  }
  bool need_cpy = c.W == REMIFA_R_32F && (c.I != REMIFA_FP32 || c.O != REMIFA_FP32);
  if (need_cpy && !have_cpy) b += ((size_t)m * n * 2 + 255) & ~(size_t)255;
  if (algo == REMIFA_LL && c.W == REMIFA_R_16F && c.IP == REMIFA_R_32F) {
    b += ((size_t)m * nb * 4 + 255) & ~(size_t)255;  // block column buffer
    b += ((size_t)nb * m * 4 + 255) & ~(size_t)255;  // block row buffer
    if (blockll_eligible((int)c.ib, c.I, nb, nb, m, 4))
      b += blockll_workspace_bytes(m, nb);  // the block kernels' operands and tables
  }
  if (lookahead_inplace(algo, c, n, nb)) b += gemm_reserve(m, nb);  // side stream's split-K
  if (lookahead_layout(algo, c, n, nb)) {
    // the next block's buffers and the side stream's split-K region
    b += ((size_t)m * nb * 4 + 255) & ~(size_t)255;
    b += ((size_t)nb * m * 4 + 255) & ~(size_t)255;
    b += gemm_reserve(m, nb);
  }
  return b + gemm_reserve(m, nb);
}

int lu_nopiv_impl(remifa_handle_t h, int algo, const remifa_lu_config_t* cfg, int64_t m, int64_t n,
                  int64_t nb, void* d_a, int64_t ldda, int* d_info, void* d_cpy, int64_t ldcpy) {
  if (!h) return -1;
  if (algo != REMIFA_LL && algo != REMIFA_RL) return -2;
  if (!cfg || validate(*cfg, algo)) return -3;
  if (m < 0) return -4;
  if (n != m) return -5;  // the reference drivers assume square matrices
  if (nb <= 0) return -6;
  if (cfg && cfg->O == REMIFA_FP8_TC32 && nb % 32) return -6;  // whole 32-blocks per outer block
  if (!d_a && m > 0) return -7;
  if (ldda < std::max<int64_t>(m, 1)) return -8;
  if (d_cpy && ldcpy < std::max<int64_t>(m, 1)) return -11;
  const remifa_lu_config_t c = *cfg;
  const int W = c.W;
  if (m == 0) {
    if (d_info) zero_int_kernel<<<1, 256, 0, h->stream>>>(d_info, 1);
    hipStreamCaptureStatus cap0 = hipStreamCaptureStatusNone;
    RF_HIP(hipStreamIsCapturing(h->stream, &cap0));
    if (cap0 == hipStreamCaptureStatusNone) RF_HIP(hipStreamSynchronize(h->stream));
    return 0;
  }
  const bool need_cpy = W == REMIFA_R_32F && (c.I != REMIFA_FP32 || c.O != REMIFA_FP32);
  const size_t ws = lu_workspace_bytes(algo, c, m, n, nb, d_cpy != nullptr);
  RF_CHECK(ensure_workspace(h, ws));
  Arena ar{(char*)h->ws};
  // slice width: a multiple of ib, at most 64 (handle knob, default 64)
  int pw = h->panel_width > 0 ? h->panel_width : 64;
  pw = std::max<int>((int)c.ib, std::min(64, pw / (int)c.ib * (int)c.ib));
  Ctx x{h, c, c.ib, pw, nullptr};
  x.tab = (float*)ar.take(subpanel_table_bytes());
  void* cpy = d_cpy;
  int64_t ldc = ldcpy;
  if (need_cpy && !cpy) {
    cpy = ar.take((size_t)m * n * 2);
    ldc = m;
  }
  const bool buffered = algo == REMIFA_LL && W == REMIFA_R_16F && c.IP == REMIFA_R_32F;
  void *pl = nullptr, *pu = nullptr;
  int64_t ldpl = m, ldpu = nb;
  if (buffered) {
    pl = ar.take((size_t)m * nb * 4);
    pu = ar.take((size_t)nb * m * 4);
  }
  // the inner factorization of each outer block in two launches
  // (blockll.hip) where it applies; panel_width set on the handle selects
  // the slice kernels instead
  void* blkws = buffered && h->panel_width == 0 && blockll_eligible((int)c.ib, c.I, nb, nb, m, 4)
                    ? ar.take(blockll_workspace_bytes(m, nb))
                    : nullptr;
  int blk_seq[2] = {0, 0};  // one-launch block kernels: launches, tickets
  // lookahead: block kk is factored in buffers plb/pub[kk % 2] while the
  // side stream prepares block kk + 1 in the other pair
  const bool la_layout = lookahead_layout(algo, c, n, nb);
  void* plb[2] = {pl, la_layout ? ar.take((size_t)m * nb * 4) : nullptr};
  void* pub[2] = {pu, la_layout ? ar.take((size_t)nb * m * 4) : nullptr};
  const bool la_inplace = lookahead_inplace(algo, c, n, nb);
  bool ahead = (la_layout || la_inplace) && h->lookahead;
  if (ahead) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(h->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
      ahead = false;  // one stream under graph capture
  }
  if (ahead) RF_CHECK(side_stream(h));
  // BF16 outer updates with fp32 inner buffers: a bf16 mirror of the
  // finished factors (each block rounded once) feeds the outer GEMMs
  const bool bf16m = buffered && c.O == REMIFA_BF16_TC32;
  void* b16 = bf16m ? ar.take((size_t)m * n * 2) : nullptr;
  const int64_t ldb16 = m;
  Fp8Mirror f8;
  const bool fp8 = algo == REMIFA_LL && c.O == REMIFA_FP8_TC32;
  if (fp8) {
    f8.ld8 = Fp8Mirror::ld_codes(n);
    f8.lds8 = f8.ld8 / 32;
    f8.l8 = (uint8_t*)ar.take((size_t)m * f8.ld8);
    f8.sl = (uint8_t*)ar.take((size_t)m * f8.lds8);
    f8.u8 = (uint8_t*)ar.take((size_t)n * f8.ld8);
    f8.su = (uint8_t*)ar.take((size_t)n * f8.lds8);
  }
  h->ws_used = ar.off;
  // split-K regions: [ar.off, +R) for the handle's stream, [ar.off + R,
  // +R) for the side stream (the arena's end is lowered while they run)
  const size_t ws_bytes0 = h->ws_bytes;
  const size_t side_lo = ar.off + gemm_reserve(m, nb), side_hi = side_lo + gemm_reserve(m, nb);
  if (la_layout || la_inplace) h->ws_bytes = side_lo;
  const int64_t nc = (n - 1) / nb + 1;
  // outer update from the mirrors: C[rows r0.., cols c0..] (M x N) -= L[r0.., 0:K] U[0:K, c0..]
  // (K range [k0, k0 + K), k0 a multiple of 128; cin: C = cin - L U, cin
  // in W precision with leading dimension ldda)
  auto gemm8 = [&](int out_t, int64_t M, int64_t N, int64_t K, int64_t r0, int64_t c0, void* cc,
                   int64_t ldcc, int64_t k0 = 0, const void* cin = nullptr) -> int {
    if (M <= 0 || N <= 0 || K <= 0) return 0;
    return fp8_gemm_launch(h, out_t, M, N, K, -1.0f, f8.l8 + r0 * f8.ld8 + k0, f8.ld8,
                           f8.sl + r0 * f8.lds8 + k0 / 32, f8.lds8, f8.u8 + c0 * f8.ld8 + k0,
                           f8.ld8, f8.su + c0 * f8.lds8 + k0 / 32, f8.lds8, 1.0f, cc, ldcc, cin,
                           W, ldda);
  };
  // quantize the finished block kk: L rows below its diagonal block, U
  // columns right of it (from the stored factors, in W precision)
  auto mirror8 = [&](int64_t ofs, int64_t in) -> int {
    const int64_t r0 = ofs + in;
    if (m > r0)
      RF_CHECK(fp8_quantize_launch(h, W, true, m - r0, in, at(d_a, W, r0, ofs, ldda), ldda,
                                   f8.l8 + r0 * f8.ld8 + ofs, f8.ld8, f8.sl + r0 * f8.lds8 + ofs / 32,
                                   f8.lds8));
    if (n > r0)
      RF_CHECK(fp8_quantize_launch(h, W, false, n - r0, in, at(d_a, W, ofs, r0, ldda), ldda,
                                   f8.u8 + r0 * f8.ld8 + ofs, f8.ld8, f8.su + r0 * f8.lds8 + ofs / 32,
                                   f8.lds8));
    return 0;
  };

  // the 16-bit operands of the buffered outer updates: A itself, or its
  // bf16 mirror
  void* const obase = bf16m ? b16 : d_a;
  const int otyp = bf16m ? REMIFA_R_16BF : W;
  const int64_t old_ = bf16m ? ldb16 : ldda;
  // side stream, while block k1 - 1 is factored: buffers of block k1 =
  // A - (blocks 0 .. k1 - 2) (a plain fp32 copy of A for k1 = 1). Waits for
  // the handle stream's last ev_main: block k1 - 1's own update part done,
  // which follows block k1 - 2's factorization and mirrors.
  auto side_prep = [&](int64_t k1) -> int {
    const int64_t o1 = k1 * nb, in1 = std::min(n - o1, nb), updm1 = m - o1, K1 = o1 - nb;
    void* blk1 = at(d_a, W, o1, o1, ldda);
    void *npl = plb[k1 & 1], *npu = pub[k1 & 1];
    RF_HIP(hipStreamWaitEvent(h->side, h->ev_main, 0));
    hipStream_t s0 = h->stream;
    const size_t u0 = h->ws_used, e0 = h->ws_bytes;
    h->stream = h->side;
    h->ws_used = side_lo;
    h->ws_bytes = side_hi;
    int st = 0;
    if (!buffered) {
      // in place: block column and block row of block k1 minus blocks
      // 0 .. k1-2 (nothing for k1 = 1)
      const bool oc = W == REMIFA_R_32F && c.O != REMIFA_FP32;
      const int64_t lo1 = oc ? ldc : ldda;
      auto opnd = [&](int64_t i, int64_t j) {
        return oc ? (const void*)at(cpy, REMIFA_R_16F, i, j, ldc) : (const void*)at(d_a, W, i, j, ldda);
      };
      if (K1 > 0) {
        st = gemm(x, c.O, W, updm1, in1, K1, opnd(o1, 0), lo1, opnd(0, o1), lo1, blk1, ldda);
        if (!st && updm1 > in1)
          st = gemm(x, c.O, W, in1, updm1 - in1, K1, opnd(o1, 0), lo1, opnd(0, o1 + in1), lo1,
                    at(d_a, W, o1, o1 + in1, ldda), ldda);
      }
    } else if (K1 == 0) {
      st = convert_launch(h, W, REMIFA_R_32F, updm1, in1, blk1, ldda, npl, ldpl);
      if (!st) st = convert_launch(h, W, REMIFA_R_32F, in1, updm1, blk1, ldda, npu, ldpu);
    } else {
      // the block column's update starts with block k1-1's inner
      // factorization on the handle's stream: its split-K rounds are
      // planned on the CUs the block kernels leave (the block row's update
      // follows, mostly after it)
      const int64_t ok0 = o1 - nb;
      if (blkws && blockll_eligible((int)c.ib, c.I, std::min(n - ok0, nb), nb, m, ldda))
        h->cu_busy = blockll_cus(h, m - ok0, std::min(n - ok0, nb));
      if (fp8) {
        // buffer = A (fp16, read by the epilogue) - L U from the e4m3 mirrors
        st = gemm8(REMIFA_R_32F, updm1, in1, K1, o1, o1, npl, ldpl, 0, blk1);
        h->cu_busy = 0;
        if (!st) st = gemm8(REMIFA_R_32F, in1, updm1, K1, o1, o1, npu, ldpu, 0, blk1);
      } else {
        const int oin = operand_type(c.O);
        const void* l1 = at(obase, otyp, o1, 0, old_);
        const void* u1 = at(obase, otyp, 0, o1, old_);
        st = gemm_launch_cin(h, c.O, bf16m ? REMIFA_R_16BF : oin, REMIFA_R_32F, updm1, in1, K1,
                             -1.0f, l1, old_, u1, old_, 1.0f, blk1, W, ldda, npl, ldpl);
        h->cu_busy = 0;
        if (!st)
          st = gemm_launch_cin(h, c.O, bf16m ? REMIFA_R_16BF : oin, REMIFA_R_32F, in1, updm1, K1,
                               -1.0f, l1, old_, u1, old_, 1.0f, blk1, W, ldda, npu, ldpu);
      }
    }
    if (!st && hipEventRecord(h->ev_side, h->side) != hipSuccess) st = REMIFA_ERR_HIP;
    h->stream = s0;
    h->ws_used = u0;
    h->ws_bytes = e0;
    return st;
  };

  auto run = [&]() -> int {
    if (ahead) RF_HIP(hipEventRecord(h->ev_main, h->stream));  // inputs ready
    if (algo == REMIFA_LL) {
      for (int64_t kk = 0; kk < nc; ++kk) {
        const int64_t ofs = kk * nb;
        const int64_t in = std::min(n - ofs, nb), updm = m - ofs;
        void* blk = at(d_a, W, ofs, ofs, ldda);
        // operands of the outer update: fp16 copy (W = fp32, TC) or A itself
        const bool o_copy = W == REMIFA_R_32F && c.O != REMIFA_FP32;
        const void* la = o_copy ? at(cpy, REMIFA_R_16F, ofs, 0, ldc) : at(d_a, W, ofs, 0, ldda);
        const void* ua = o_copy ? at(cpy, REMIFA_R_16F, 0, ofs, ldc) : at(d_a, W, 0, ofs, ldda);
        const void* ua2 = o_copy ? at(cpy, REMIFA_R_16F, 0, ofs + in, ldc)
                                 : at(d_a, W, 0, ofs + in, ldda);
        const int64_t lo = o_copy ? ldc : ldda;
        // operands of the buffered outer update: A itself, or its bf16 mirror
        const int oin = bf16m ? REMIFA_R_16BF : operand_type(c.O);
        const void* lb = bf16m ? at(b16, REMIFA_R_16BF, ofs, 0, ldb16) : la;
        const void* ub = bf16m ? at(b16, REMIFA_R_16BF, 0, ofs, ldb16) : ua;
        const int64_t lbo = bf16m ? ldb16 : lo;
        if (!buffered) {
          if (ofs > 0 && fp8) {
            RF_CHECK(gemm8(W, updm, in, ofs, ofs, ofs, blk, ldda));
            RF_CHECK(gemm8(W, in, updm - in, ofs, ofs, ofs + in, at(d_a, W, ofs, ofs + in, ldda),
                           ldda));
          } else if (ahead) {
            // the side stream subtracted blocks 0 .. kk-2 in place; block
            // kk-1 (K = nb) follows here, then the side stream starts on
            // block kk+1 while this block is factored
            if (kk > 0) {
              RF_HIP(hipStreamWaitEvent(h->stream, h->ev_side, 0));
              const int64_t op = ofs - nb;
              const void* lk = o_copy ? at(cpy, REMIFA_R_16F, ofs, op, ldc) : at(d_a, W, ofs, op, ldda);
              const void* uk = o_copy ? at(cpy, REMIFA_R_16F, op, ofs, ldc) : at(d_a, W, op, ofs, ldda);
              const void* uk2 = o_copy ? at(cpy, REMIFA_R_16F, op, ofs + in, ldc)
                                       : at(d_a, W, op, ofs + in, ldda);
              RF_CHECK(gemm(x, c.O, W, updm, in, nb, lk, lo, uk, lo, blk, ldda));
              if (updm > in)
                RF_CHECK(gemm(x, c.O, W, in, updm - in, nb, lk, lo, uk2, lo,
                              at(d_a, W, ofs, ofs + in, ldda), ldda));
              RF_HIP(hipEventRecord(h->ev_main, h->stream));
            }
            if (kk + 1 < nc) RF_CHECK(side_prep(kk + 1));
          } else if (ofs > 0) {
            // block column (:239-304) and block row (:311-410), in place
            RF_CHECK(gemm(x, c.O, W, updm, in, ofs, la, lo, ua, lo, blk, ldda));
            if (updm > in)
              RF_CHECK(gemm(x, c.O, W, in, updm - in, ofs, la, lo, ua2, lo,
                            at(d_a, W, ofs, ofs + in, ldda), ldda));
          }
          // inner factorization in place; with an fp16 copy (W = fp32 and a
          // TC update) the slice kernels keep it current (:554-563)
          void* cblk = need_cpy ? at(cpy, REMIFA_R_16F, ofs, ofs, ldc) : nullptr;
          RF_CHECK(inner_ll(x, W, updm, in, blk, ldda, blk, ldda, cblk, ldc, need_cpy));
          if (fp8) RF_CHECK(mirror8(ofs, in));
        } else {
          // fp16 storage, fp32 inner buffers: pl = block column (updm x in),
          // pu = block row incl. the diagonal block (in x updm)
          if (c.OP == REMIFA_R_16F && ofs > 0 && fp8) {
            RF_CHECK(gemm8(W, updm, in, ofs, ofs, ofs, blk, ldda));
            RF_CHECK(gemm8(W, in, updm - in, ofs, ofs, ofs + in, at(d_a, W, ofs, ofs + in, ldda),
                           ldda));
          } else if (c.OP == REMIFA_R_16F && ofs > 0) {
            RF_CHECK(gemm(x, c.O, W, updm, in, ofs, la, lo, ua, lo, blk, ldda));
            if (updm > in)
              RF_CHECK(gemm(x, c.O, W, in, updm - in, ofs, la, lo, ua2, lo,
                            at(d_a, W, ofs, ofs + in, ldda), ldda));
          }
          // (the diagonal block is in both buffers, as in the reference's
          // d_l_tmp / d_u_tmp: L and the diagonal tiles are read from pl,
          // the U part right of each tile from pu)
          void* const cpl = ahead ? plb[kk & 1] : pl;
          void* const cpu = ahead ? pub[kk & 1] : pu;
          if (ahead) {
            if (kk == 0) {
              RF_CHECK(convert_launch(h, W, REMIFA_R_32F, updm, in, blk, ldda, cpl, ldpl));
              RF_CHECK(convert_launch(h, W, REMIFA_R_32F, in, updm, blk, ldda, cpu, ldpu));
            } else {
              // the side stream left A - (blocks 0 .. kk-2) here; block
              // kk-1 (nb wide) follows: same products, summed in two parts
              RF_HIP(hipStreamWaitEvent(h->stream, h->ev_side, 0));
              const int64_t op = ofs - nb;
              if (fp8) {
                RF_CHECK(gemm8(REMIFA_R_32F, updm, in, nb, ofs, ofs, cpl, ldpl, op));
                RF_CHECK(gemm8(REMIFA_R_32F, in, updm, nb, ofs, ofs, cpu, ldpu, op));
              } else {
                const void* lr = at(obase, otyp, ofs, op, old_);
                const void* ur = at(obase, otyp, op, ofs, old_);
                RF_CHECK(gemm_dual_launch(h, c.O, oin, updm, in, nb, lr, old_, ur, old_, cpl, ldpl,
                                          in, updm, lr, old_, ur, old_, cpu, ldpu));
              }
            }
            // the second stream's next update starts after this block's
            // own part (which then has the CUs to itself); that point also
            // follows block kk-1's factorization, which the update reads
            if (kk > 0) RF_HIP(hipEventRecord(h->ev_main, h->stream));
            if (kk + 1 < nc) RF_CHECK(side_prep(kk + 1));
          } else if (c.OP == REMIFA_R_32F && ofs > 0 && !fp8) {
            // block column and block row straight into the fp32 buffers:
            // buffer = A (fp16, read by the epilogue) - L U (no separate
            // fp16 -> fp32 copy of the block first)
            RF_CHECK(gemm_launch_cin(h, c.O, oin, REMIFA_R_32F, updm, in, ofs, -1.0f, lb, lbo, ub,
                                     lbo, 1.0f, blk, W, ldda, pl, ldpl));
            RF_CHECK(gemm_launch_cin(h, c.O, oin, REMIFA_R_32F, in, updm, ofs, -1.0f, lb, lbo, ub,
                                     lbo, 1.0f, blk, W, ldda, pu, ldpu));
          } else if (c.OP == REMIFA_R_32F && ofs > 0 && fp8) {
            // likewise from the e4m3 mirrors
            RF_CHECK(gemm8(REMIFA_R_32F, updm, in, ofs, ofs, ofs, pl, ldpl, 0, blk));
            RF_CHECK(gemm8(REMIFA_R_32F, in, updm, ofs, ofs, ofs, pu, ldpu, 0, blk));
          } else {
            RF_CHECK(convert_launch(h, W, REMIFA_R_32F, updm, in, blk, ldda, pl, ldpl));
            RF_CHECK(convert_launch(h, W, REMIFA_R_32F, in, updm, blk, ldda, pu, ldpu));
          }
          // inner factorization in fp32; its fp16 mirror is A itself (the
          // reference passes d_a as the fp16 copy, src/lu_nopiv_ll.cxx:632),
          // so the rounded factors land in A as they are produced
          if (blkws && blockll_eligible((int)c.ib, c.I, in, nb, m, ldda))
            RF_CHECK(blockll_launch(h, blkws, m, nb, updm, in, (const float*)cpl, ldpl,
                                    (const float*)cpu, ldpu, (half_t*)blk, ldda, blk_seq));
          else
            RF_CHECK(inner_ll(x, REMIFA_R_32F, updm, in, cpl, ldpl, cpu, ldpu, blk, ldda, true));
          if (fp8) RF_CHECK(mirror8(ofs, in));
          if (bf16m) {
            // round the finished block column (below the diagonal block) and
            // block row (right of it) to bf16 once
            const int64_t r0 = ofs + in;
            if (m > r0)
              RF_CHECK(convert_launch(h, W, REMIFA_R_16BF, m - r0, in, at(d_a, W, r0, ofs, ldda),
                                      ldda, at(b16, REMIFA_R_16BF, r0, ofs, ldb16), ldb16));
            if (n > r0)
              RF_CHECK(convert_launch(h, W, REMIFA_R_16BF, in, n - r0, at(d_a, W, ofs, r0, ldda),
                                      ldda, at(b16, REMIFA_R_16BF, ofs, r0, ldb16), ldb16));
          }
          if (ahead) RF_HIP(hipEventRecord(h->ev_main, h->stream));  // block kk finished
        }
      }
    } else {
      // right-looking
      const bool tc_in = W == REMIFA_R_32F && c.I != REMIFA_FP32;
      const bool tc_out = W == REMIFA_R_32F && c.O != REMIFA_FP32;
      for (int64_t kk = 0; kk < nc; ++kk) {
        const int64_t ofs = kk * nb, ofst = (kk + 1) * nb;
        const int64_t in = std::min(n - ofs, nb), updm = m - ofs;
        void* p = at(d_a, W, ofs, ofs, ldda);
        for (int64_t so = 0; so < in; so += x.pw) {
          const int sw = (int)std::min<int64_t>(x.pw, in - so);
          const int64_t cm = updm - so, tso = so + sw;
          RF_CHECK(subpanel_launch(
              h, W, c.I, true, (int)x.ib, cm, sw, at(p, W, so, so, ldda), ldda,
              at(p, W, so, so, ldda), ldda,
              (tc_in || tc_out) ? (half_t*)at(cpy, REMIFA_R_16F, ofs + so, ofs + so, ldc) : nullptr,
              ldc, x.tab));
          if (in - tso > 0) {
            const void* lop = tc_in ? at(cpy, REMIFA_R_16F, ofs + tso, ofs + so, ldc)
                                    : at(p, W, tso, so, ldda);
            const void* uop = tc_in ? at(cpy, REMIFA_R_16F, ofs + so, ofs + tso, ldc)
                                    : at(p, W, so, tso, ldda);
            const int64_t lop_ld = tc_in ? ldc : ldda;
            // rank-sw update of the outer block column (:139-163)
            RF_CHECK(gemm(x, c.I, W, updm - tso, in - tso, sw, lop, lop_ld, uop, lop_ld,
                          at(p, W, tso, tso, ldda), ldda));
            // and of the outer block row beyond the outer block (:167-193)
            if (updm > in) {
              const void* uop2 = tc_in ? at(cpy, REMIFA_R_16F, ofs + so, ofs + in, ldc)
                                       : at(p, W, so, in, ldda);
              RF_CHECK(gemm(x, c.I, W, in - tso, updm - in, sw, lop, lop_ld, uop2, lop_ld,
                            at(p, W, tso, in, ldda), ldda));
            }
          }
        }
        // Schur complement (:170-195)
        const int64_t tm = m - ofst, tn = n - ofst;
        if (tn > 0) {
          const void* lop = tc_out ? at(cpy, REMIFA_R_16F, ofst, ofs, ldc) : at(d_a, W, ofst, ofs, ldda);
          const void* uop = tc_out ? at(cpy, REMIFA_R_16F, ofs, ofst, ldc) : at(d_a, W, ofs, ofst, ldda);
          const int64_t ldo = tc_out ? ldc : ldda;
          RF_CHECK(gemm(x, c.O, W, tm, tn, nb, lop, ldo, uop, ldo,
                        at(d_a, W, ofst, ofst, ldda), ldda));
        }
      }
    }
    return 0;
  };
  int st = run();
  h->ws_used = 0;
  h->ws_bytes = ws_bytes0;
  if (ahead) {
    // join the side stream (also after a failed launch)
    if (hipEventRecord(h->ev_side, h->side) != hipSuccess ||
        hipStreamWaitEvent(h->stream, h->ev_side, 0) != hipSuccess)
      st = st ? st : REMIFA_ERR_HIP;
  }
  if (st) return st;
  if (d_info) {
    info_kernel<<<1, 256, 0, h->stream>>>(n, d_a, ldda, W, d_info);
    RF_HIP(hipGetLastError());
  }
  // the reference drivers synchronise before returning (:567-569). Under
  // HIP graph capture the call only records its kernels: no readback and no
  // synchronisation (a replay's block-kernel wait flag is then not checked)
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  RF_HIP(hipStreamIsCapturing(h->stream, &cap));
  if (cap != hipStreamCaptureStatusNone) return 0;
  int waited_out = 0;
  if (blk_seq[0] > 0)
    RF_HIP(hipMemcpyAsync(&waited_out, blockll_sync_words(blkws, m, nb) + 2, sizeof(int),
                          hipMemcpyDeviceToHost, h->stream));
  RF_HIP(hipStreamSynchronize(h->stream));
  if (waited_out) {
    set_last_error("block kernels: wait for the diagonal role timed out");
    return REMIFA_ERR_HIP;
  }
  return 0;
}

}  // namespace remifa
