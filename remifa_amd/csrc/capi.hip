// extern "C" entry points: argument checking and dispatch. The twelve
// lu_nopiv_ll_* helpers fix ib = 8 and one precision tuple each, exactly as
// src/lu_nopiv_ll.cxx does.
#include "common.hpp"

namespace remifa {
size_t lu_workspace_bytes(int algo, const remifa_lu_config_t& c, int64_t m, int64_t n, int64_t nb,
                          bool have_cpy);
int lu_nopiv_impl(remifa_handle_t h, int algo, const remifa_lu_config_t* cfg, int64_t m, int64_t n,
                  int64_t nb, void* d_a, int64_t ldda, int* d_info, void* d_cpy, int64_t ldcpy);
}  // namespace remifa

using namespace remifa;

namespace {
constexpr int F32 = REMIFA_R_32F, F16 = REMIFA_R_16F;

remifa_lu_config_t tuple(int W, int I, int O, int OP, int IP, int OF, int IF) {
  remifa_lu_config_t c;
  c.W = W;
  c.I = I;
  c.O = O;
  c.OP = OP;
  c.IP = IP;
  c.OF = OF;
  c.IF = IF;
  c.ib = 8;
  return c;
}

int ll(remifa_handle_t h, const remifa_lu_config_t& c, int m, int n, int nb, void* d_a, int ldda,
       int* d_info, void* cpy = nullptr, int ldcpy = 0) {
  int st = lu_nopiv_impl(h, REMIFA_LL, &c, m, n, nb, d_a, ldda, d_info, cpy, ldcpy);
  // map the generic driver's argument positions onto the helper signature
  // (h, m, n, nb, d_a, ldda, d_info, d_a_f16, ld_a_f16)
  switch (st) {
    case -4: return -2;
    case -5: return -3;
    case -6: return -4;
    case -7: return -5;
    case -8: return -6;
    case -11: return -9;
    default: return st;
  }
}
}  // namespace

extern "C" {

int remifa_lu_workspace_size(remifa_handle_t h, int algo, const remifa_lu_config_t* cfg, int64_t m,
                             int64_t n, int64_t nb, size_t* bytes) {
  if (!h) return -1;
  if (algo != REMIFA_LL && algo != REMIFA_RL) return -2;
  if (!cfg) return -3;
  if (!bytes) return -7;
  *bytes = lu_workspace_bytes(algo, *cfg, m, n, nb, false);
  return 0;
}

int remifa_lu_nopiv(remifa_handle_t h, int algo, const remifa_lu_config_t* cfg, int64_t m,
                    int64_t n, int64_t nb, void* d_a, int64_t ldda, int* d_info, void* d_a_cpy,
                    int64_t ld_a_cpy) {
  return lu_nopiv_impl(h, algo, cfg, m, n, nb, d_a, ldda, d_info, d_a_cpy, ld_a_cpy);
}

int remifa_lu_nopiv_ll_f32(remifa_handle_t h, int m, int n, int nb, float* d_a, int ldda,
                           int* d_info) {
  return ll(h, tuple(F32, REMIFA_FP32, REMIFA_FP32, F32, F32, F32, F32), m, n, nb, d_a, ldda,
            d_info);
}
int remifa_lu_nopiv_ll_f32_fp32xtc32(remifa_handle_t h, int m, int n, int nb, float* d_a, int ldda,
                                     int* d_info, void* d_a_f16, int ld_a_f16) {
  // the reference throws without the fp16 buffer (src/lu_nopiv_ll.hxx:132-135)
  if (!d_a_f16) return -8;
  return ll(h, tuple(F32, REMIFA_FP32, REMIFA_TC32, F32, F32, F32, F32), m, n, nb, d_a, ldda,
            d_info, d_a_f16, ld_a_f16);
}
int remifa_lu_nopiv_ll_f32_tc32xtc32(remifa_handle_t h, int m, int n, int nb, float* d_a, int ldda,
                                     int* d_info, void* d_a_f16, int ld_a_f16) {
  if (!d_a_f16) return -8;
  return ll(h, tuple(F32, REMIFA_TC32, REMIFA_TC32, F32, F32, F32, F32), m, n, nb, d_a, ldda,
            d_info, d_a_f16, ld_a_f16);
}
int remifa_lu_nopiv_ll_f16(remifa_handle_t h, int m, int n, int nb, void* d_a, int ldda,
                           int* d_info) {
  return ll(h, tuple(F16, REMIFA_FP16, REMIFA_FP16, F16, F16, F16, F16), m, n, nb, d_a, ldda,
            d_info);
}
int remifa_lu_nopiv_ll_f16_tc16xtc16(remifa_handle_t h, int m, int n, int nb, void* d_a, int ldda,
                                     int* d_info) {
  return ll(h, tuple(F16, REMIFA_TC16, REMIFA_TC16, F16, F16, F16, F16), m, n, nb, d_a, ldda,
            d_info);
}
int remifa_lu_nopiv_ll_f16_tc32xtc32(remifa_handle_t h, int m, int n, int nb, void* d_a, int ldda,
                                     int* d_info) {
  return ll(h, tuple(F16, REMIFA_TC32, REMIFA_TC32, F16, F16, F16, F16), m, n, nb, d_a, ldda,
            d_info);
}
int remifa_lu_nopiv_ll_f16_f16xtc32(remifa_handle_t h, int m, int n, int nb, void* d_a, int ldda,
                                    int* d_info) {
  return ll(h, tuple(F16, REMIFA_FP16, REMIFA_TC32, F16, F16, F16, F16), m, n, nb, d_a, ldda,
            d_info);
}
int remifa_lu_nopiv_ll_f16_f16xtc32_fp32(remifa_handle_t h, int m, int n, int nb, void* d_a,
                                         int ldda, int* d_info) {
  // OP = fp32 with OF = fp16: the fp32-accumulated outer update is rounded
  // once to fp16 before the fp16 inner factorization, which is what the TC32
  // GEMM with fp16 output does in place (the reference reinterprets its fp32
  // buffer as fp16 here, src/lu_nopiv_ll.hxx:481-489).
  return ll(h, tuple(F16, REMIFA_FP16, REMIFA_TC32, F16, F16, F16, F16), m, n, nb, d_a, ldda,
            d_info);
}
int remifa_lu_nopiv_ll_f16_f32xt32(remifa_handle_t h, int m, int n, int nb, void* d_a, int ldda,
                                   int* d_info) {
  return ll(h, tuple(F16, REMIFA_FP32, REMIFA_TC32, F32, F32, F32, F32), m, n, nb, d_a, ldda,
            d_info);
}
int remifa_lu_nopiv_ll_f16_f16xtc16(remifa_handle_t h, int m, int n, int nb, void* d_a, int ldda,
                                    int* d_info) {
  return ll(h, tuple(F16, REMIFA_FP16, REMIFA_TC16, F16, F16, F16, F16), m, n, nb, d_a, ldda,
            d_info);
}
int remifa_lu_nopiv_ll_f16_tc32xtc32_fp16xfp32_fp16xfp32(remifa_handle_t h, int m, int n, int nb,
                                                         void* d_a, int ldda, int* d_info) {
  return ll(h, tuple(F16, REMIFA_TC32, REMIFA_TC32, F16, F32, F16, F32), m, n, nb, d_a, ldda,
            d_info);
}
int remifa_lu_nopiv_ll_f16_tc32xtc32_fp32xfp32_fp32xfp32(remifa_handle_t h, int m, int n, int nb,
                                                         void* d_a, int ldda, int* d_info) {
  return ll(h, tuple(F16, REMIFA_TC32, REMIFA_TC32, F32, F32, F32, F32), m, n, nb, d_a, ldda,
            d_info);
}

int remifa_lu_nopiv_panel(remifa_handle_t h, int t, int ib, int m, int n, const void* d, int ldd,
                          void* l, int ldl, void* u, int ldu, int* stat) {
  if (!h) return -1;
  if (t != REMIFA_R_32F && t != REMIFA_R_16F && t != REMIFA_R_64F) return -2;
  if (ib < 1 || ib > 32) return -3;
  if (m < 0) return -4;
  if (n < 0 || n > ib || n > m) return -5;
  if (!d && n > 0) return -6;
  if (ldd < (n > 1 ? n : 1)) return -7;
  if (!l && n > 0) return -8;
  if (ldl < (m > 1 ? m : 1)) return -9;
  if (!u && m > n) return -10;
  if (ldu < (n > 1 ? n : 1)) return -11;
  return panel_launch(h, t, ib, m, n, d, ldd, l, ldl, u, ldu, stat);
}

int remifa_convert(remifa_handle_t h, int src_t, int dst_t, int64_t m, int64_t n, const void* a,
                   int64_t lda, void* aout, int64_t ldaout) {
  if (!h) return -1;
  if (!dtype_size(src_t)) return -2;
  if (!dtype_size(dst_t)) return -3;
  if (m < 0) return -4;
  if (n < 0) return -5;
  if (!a && m > 0 && n > 0) return -6;
  if (lda < (m > 1 ? m : 1)) return -7;
  if (!aout && m > 0 && n > 0) return -8;
  if (ldaout < (m > 1 ? m : 1)) return -9;
  return convert_launch(h, src_t, dst_t, m, n, a, lda, aout, ldaout);
}

int remifa_gemmex(remifa_handle_t h, int ctype, int in_t, int out_t, int m, int n, int k,
                  float alpha, const void* a, int lda, const void* b, int ldb, float beta, void* c,
                  int ldc) {
  if (!h) return -1;
  if (ctype < REMIFA_FP32 || ctype > REMIFA_FP8_TC32) return -2;
  if (in_t != REMIFA_R_32F && in_t != REMIFA_R_16F &&
      !(in_t == REMIFA_R_16BF && ctype == REMIFA_BF16_TC32))
    return -3;
  if (out_t != REMIFA_R_32F && out_t != REMIFA_R_16F) return -4;
  if (m < 0) return -5;
  if (n < 0) return -6;
  if (k < 0) return -7;
  if (!a && m > 0 && k > 0) return -9;
  if (lda < (m > 1 ? m : 1)) return -10;
  if (!b && k > 0 && n > 0) return -11;
  if (ldb < (k > 1 ? k : 1)) return -12;
  if (!c && m > 0 && n > 0) return -14;
  if (ldc < (m > 1 ? m : 1)) return -15;
  int st = gemm_launch(h, ctype, in_t, out_t, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
  // gemm_launch numbers (ctype, in_t, out_t) as 2, 4, 5
  if (st == -4) return -3;
  if (st == -5) return -4;
  return st;
}

int remifa_quantize_fp8(remifa_handle_t h, int in_t, int rows, int64_t m, int64_t k,
                        const void* a, int64_t lda, uint8_t* q, int64_t ldq, uint8_t* s,
                        int64_t lds) {
  if (!h) return -1;
  if (in_t != REMIFA_R_32F && in_t != REMIFA_R_16F) return -2;
  if (m < 0) return -4;
  if (k < 0) return -5;
  if (m == 0 || k == 0) return 0;
  if (!a) return -6;
  if (lda < (rows ? m : k)) return -7;
  if (!q || (uintptr_t)q % 16) return -8;
  if (ldq % 16 || ldq < (k + 31) / 32 * 32) return -9;
  if (!s) return -10;
  if (lds < (k + 31) / 32) return -11;
  return fp8_quantize_launch(h, in_t, rows != 0, m, k, a, lda, q, ldq, s, lds);
}

int remifa_gemm_fp8(remifa_handle_t h, int out_t, int64_t m, int64_t n, int64_t k, float alpha,
                    const uint8_t* qa, int64_t ldqa, const uint8_t* sa, int64_t ldsa,
                    const uint8_t* qb, int64_t ldqb, const uint8_t* sb, int64_t ldsb, float beta,
                    void* c, int64_t ldc) {
  if (!h) return -1;
  if (out_t != REMIFA_R_32F && out_t != REMIFA_R_16F) return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (k < 0 || k % 32) return -5;
  if (m == 0 || n == 0) return 0;
  const int64_t nkb4 = (k + 127) / 128 * 4;
  if (k > 0) {
    if (!qa || (uintptr_t)qa % 16) return -7;
    if (ldqa % 16 || ldqa < k) return -8;
    if (!sa || (uintptr_t)sa % 4) return -9;
    if (ldsa % 4 || ldsa < nkb4) return -10;
    if (!qb || (uintptr_t)qb % 16) return -11;
    if (ldqb % 16 || ldqb < k) return -12;
    if (!sb || (uintptr_t)sb % 4) return -13;
    if (ldsb % 4 || ldsb < nkb4) return -14;
  }
  if (!c) return -16;
  if (ldc < m) return -17;
  return fp8_gemm_launch(h, out_t, m, n, k, alpha, qa, ldqa, sa, ldsa, qb, ldqb, sb, ldsb, beta, c,
                         ldc);
}

}  // extern "C"
