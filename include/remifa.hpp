// C++ layer over remifa.h with the reference's entry points: the same names
// and arguments as src/lu_nopiv_ll.hxx:588-793 (namespace remifa), a Handle
// in place of the cublasHandle_t (one stream per handle, as
// cublasGetStream, src/lu_nopiv_ll.hxx:139), and the reference's error
// convention: std::invalid_argument for bad arguments, std::runtime_error for
// runtime failures (src/errchk.hxx). Every call synchronises its stream
// before returning, as the reference drivers do (src/lu_nopiv_ll.hxx:568).
//
// The CUTLASS and CUTLASS split-K variants of the reference are the same
// factorizations with another GEMM backend; here they forward to the same
// matrix-core GEMM (slices sets the split-K factor, workspace is ignored: the
// handle owns its workspace).
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <stdexcept>
#include <string>

#include "remifa.h"

namespace remifa {

// src/utils.hxx:18-24 (+ the bf16 / fp8 update types of this library)
enum compute_type {
  FP32 = REMIFA_FP32,
  FP16 = REMIFA_FP16,
  TC32 = REMIFA_TC32,
  TC16 = REMIFA_TC16,
  BF16_TC32 = REMIFA_BF16_TC32,
  FP8_TC32 = REMIFA_FP8_TC32
};

inline void check(int status, const char* what) {
  if (status == 0) return;
  std::string msg = std::string(what) + ": ";
  if (status < 0)
    throw std::invalid_argument(msg + "argument " + std::to_string(-status) + " is invalid");
  const char* e = remifa_last_error();
  throw std::runtime_error(msg + (e && *e ? e : "error " + std::to_string(status)));
}

class Handle {
 public:
  explicit Handle(hipStream_t stream = nullptr) { check(remifa_create(&h_, stream), "remifa_create"); }
  ~Handle() { remifa_destroy(h_); }
  Handle(const Handle&) = delete;
  Handle& operator=(const Handle&) = delete;
  operator remifa_handle_t() const { return h_; }
  void set_stream(hipStream_t s) { check(remifa_set_stream(h_, s), "remifa_set_stream"); }

 private:
  remifa_handle_t h_ = nullptr;
};

// ---------------------------------------------------------------- templates
// lu_nopiv_ll<T, I, O, ib, OP, IP, OF, IF> (src/lu_nopiv_ll.hxx:41-66) and
// lu_nopiv_rl<T, I, O, ib> (src/lu_nopiv.hxx:44-63); element types are given
// as remifa_dtype_t values.
inline void lu_nopiv(Handle& h, int algo, int W, int I, int O, int ib, int OP, int IP, int OF,
                     int IF, int64_t m, int64_t n, int64_t nb, void* d_a, int64_t ldda,
                     int* d_info, void* d_a_cpy = nullptr, int64_t ld_a_cpy = 0) {
  remifa_lu_config_t c{W, I, O, OP, IP, OF, IF, ib};
  check(remifa_lu_nopiv(h, algo, &c, m, n, nb, d_a, ldda, d_info, d_a_cpy,
                        d_a_cpy ? ld_a_cpy : 0),
        "lu_nopiv");
}

// ---------------------------------------------------------------- helpers
#define REMIFA_HELPER_F32(name)                                                              \
  inline void name(Handle& h, int m, int n, int nb, float* const d_a, int ldda, int* d_info) { \
    check(remifa_##name(h, m, n, nb, d_a, ldda, d_info), #name);                              \
  }
#define REMIFA_HELPER_F32_CPY(name)                                                         \
  inline void name(Handle& h, int m, int n, int nb, float* const d_a, int ldda, int* d_info, \
                   __half* d_a_f16, int ld_a_f16) {                                          \
    check(remifa_##name(h, m, n, nb, d_a, ldda, d_info, d_a_f16, ld_a_f16), #name);          \
  }
#define REMIFA_HELPER_F16(name)                                                               \
  inline void name(Handle& h, int m, int n, int nb, __half* const d_a, int ldda, int* d_info) { \
    check(remifa_##name(h, m, n, nb, d_a, ldda, d_info), #name);                               \
  }

REMIFA_HELPER_F32(lu_nopiv_ll_f32)                        // src/lu_nopiv_ll.hxx:588
REMIFA_HELPER_F32_CPY(lu_nopiv_ll_f32_fp32xtc32)          // :603
REMIFA_HELPER_F32_CPY(lu_nopiv_ll_f32_tc32xtc32)          // :620
REMIFA_HELPER_F16(lu_nopiv_ll_f16)                        // :638
REMIFA_HELPER_F16(lu_nopiv_ll_f16_f16xtc16)               // :654
REMIFA_HELPER_F16(lu_nopiv_ll_f16_tc16xtc16)              // :670
REMIFA_HELPER_F16(lu_nopiv_ll_f16_tc32xtc32)              // :680
REMIFA_HELPER_F16(lu_nopiv_ll_f16_f16xtc32)               // :698
REMIFA_HELPER_F16(lu_nopiv_ll_f16_f16xtc32_fp32)          // :716
REMIFA_HELPER_F16(lu_nopiv_ll_f16_f32xt32)                // :734
REMIFA_HELPER_F16(lu_nopiv_ll_f16_tc32xtc32_fp16xfp32_fp16xfp32)  // :753
REMIFA_HELPER_F16(lu_nopiv_ll_f16_tc32xtc32_fp32xfp32_fp32xfp32)  // :775

#undef REMIFA_HELPER_F32
#undef REMIFA_HELPER_F32_CPY
#undef REMIFA_HELPER_F16

// CUTLASS / CUTLASS split-K spellings (same semantics, this library's GEMM)
inline void lu_nopiv_ll_cutlass_f32(Handle& h, int m, int n, int nb, float* const d_a, int ldda,
                                    int* d_info) {
  lu_nopiv_ll_f32(h, m, n, nb, d_a, ldda, d_info);
}
inline void lu_nopiv_ll_cutlass_splitk_f32(Handle& h, int m, int n, int nb, float* const d_a,
                                           int ldda, int* d_info, int slices, uint8_t*) {
  check(remifa_set_tuning(h, 0, slices), "remifa_set_tuning");
  lu_nopiv_ll_f32(h, m, n, nb, d_a, ldda, d_info);
  check(remifa_set_tuning(h, 0, 0), "remifa_set_tuning");
}
inline void lu_nopiv_ll_cutlass_f32_tc32xtc32(Handle& h, int m, int n, int nb, float* const d_a,
                                              int ldda, int* d_info, __half* d_a_f16,
                                              int ld_a_f16) {
  lu_nopiv_ll_f32_tc32xtc32(h, m, n, nb, d_a, ldda, d_info, d_a_f16, ld_a_f16);
}
inline void lu_nopiv_ll_cutlass_f16_tc32xtc32(Handle& h, int m, int n, int nb, __half* const d_a,
                                              int ldda, int* d_info) {
  lu_nopiv_ll_f16_tc32xtc32(h, m, n, nb, d_a, ldda, d_info);
}

// ---------------------------------------------------------------- kernels
// lu_nopiv_panel<T, ib> (src/lu_nopiv.cuh:12-20), T = float or __half
template <typename T, int ib>
void lu_nopiv_panel(Handle& h, int m, int n, T const* d, int ldd, T* l, int ldl, T* u, int ldu,
                    int* stat) {
  constexpr int t = sizeof(T) == 4 ? REMIFA_R_32F : REMIFA_R_16F;
  check(remifa_lu_nopiv_panel(h, t, ib, m, n, d, ldd, l, ldl, u, ldu, stat), "lu_nopiv_panel");
}

// convert<TA, TAO> (src/convert.cuh:17-18)
template <typename TA, typename TAO>
void convert(Handle& h, int m, int n, TA* a, int lda, TAO* aout, int ldaout) {
  constexpr int s = sizeof(TA) == 4 ? REMIFA_R_32F : REMIFA_R_16F;
  constexpr int d = sizeof(TAO) == 4 ? REMIFA_R_32F : REMIFA_R_16F;
  check(remifa_convert(h, s, d, m, n, a, lda, aout, ldaout), "convert");
}

// gemmex (src/gemm.cuh:24-40): C = alpha A B + beta C, column-major
inline void gemmex(Handle& h, compute_type ctype, int in_t, int out_t, int m, int n, int k,
                   float alpha, const void* a, int lda, const void* b, int ldb, float beta,
                   void* c, int ldc) {
  check(remifa_gemmex(h, ctype, in_t, out_t, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc),
        "gemmex");
}

}  // namespace remifa
