// Shared definitions of libremifa_hip: element types, device conversions,
// the handle, error plumbing and launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "remifa.h"

namespace remifa {

using half_t = _Float16;
using bf16_t = uint16_t;  // raw bits; converted explicitly
using fp8_t = uint8_t;    // OCP e4m3fn bits

// ---------------------------------------------------------------- errors
struct Error {
  int status;
  std::string what;
};

#define RF_HIP(call)                                                        \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess) {                                                 \
      remifa::set_last_error(std::string(#call) + ": " + hipGetErrorString(e_)); \
      return REMIFA_ERR_HIP;                                                \
    }                                                                       \
  } while (0)

#define RF_CHECK(call)      \
  do {                      \
    int s_ = (call);        \
    if (s_ != 0) return s_; \
  } while (0)

void set_last_error(const std::string& msg);
const std::string& last_error();

// ---------------------------------------------------------------- handle
struct KernelTimer {
  // Per-kernel-class accumulated time (ms) and launch counts, measured with
  // hipEvents recorded on the handle's stream around each launch.
  bool enabled = false;
  std::vector<hipEvent_t> pool;
  struct Rec {
    int cls;
    int ev0, ev1;
    double work;  // algorithmic flops or bytes of this launch
    int variant;  // kernel instantiation (tile shape, operand kind, ...)
  };
  std::vector<Rec> recs;
  size_t next = 0;
};

}  // namespace remifa

struct remifa_handle_st {
  hipStream_t stream = nullptr;
  int device = 0;
  int num_cu = 256;
  // workspace arena (device memory), grown on demand unless user-supplied
  void* ws = nullptr;
  size_t ws_bytes = 0;
  size_t ws_used = 0;  // prefix reserved by a running driver; GEMMs use the rest
  bool ws_user = false;
  remifa::KernelTimer timer;
  // tuning knobs (0 = automatic)
  int panel_width = 0;
  int split_k = 0;
  int gemm_tile = -1;  // -1: chosen per shape
  // block kernels of the LL inner factorization: one launch (1) or the
  // diagonal and far kernels one after the other (0); far units per
  // helper wavefront of the one-launch form (0 = automatic)
  int blk_fused = 1;
  int blk_upw = 0;
  long long blk_wait_polls = 0;  // polls per wait of the one-launch form (0 = 2^21)
  // CUs that concurrent work of the same driver holds while the next update
  // GEMMs run (the split-K model plans their rounds on the rest)
  int cu_busy = 0;
  // LL lookahead: the next outer block's update runs on a second stream
  // while the current block is factored (1 = on, 0 = off)
  int lookahead = 1;
  int side_reserve = 0;        // CUs the side stream leaves to the handle's stream
  int side_reserve_made = -1;  // ... as created
  hipStream_t side = nullptr;  // created on first use, default priority
  hipEvent_t ev_main = nullptr, ev_side = nullptr;
  // autotuned GEMM rules, first match wins: {opk (-1 any), m_lo, m_hi, n_lo,
  // n_hi, tile (-1 heuristic), split (0 heuristic)}, bounds inclusive
  std::vector<std::array<int64_t, 7>> gemm_rules;
  // batched factorization shape per size bucket: {bound, threads, per CU}
  std::vector<std::array<int, 3>> batched_rules;
  // device address of this device's static scratch array of the
  // no-allocation entry point (solve.hip), looked up at creation
  void* scratch = nullptr;
};

namespace remifa {

int ensure_workspace(remifa_handle_t h, size_t bytes);

// kernel classes for the timer
enum KernelClass {
  KC_GEMM = 0,
  KC_PANEL = 1,
  KC_CONVERT = 2,
  KC_TRSM = 3,
  KC_RESID = 4,
  KC_OTHER = 5,
  KC_COUNT = 6
};
void timer_begin(remifa_handle_t h, int cls, double work, int variant = 0);
void timer_end(remifa_handle_t h);

// ---------------------------------------------------------------- device conversions
__device__ __forceinline__ float to_f32(float x) { return x; }
__device__ __forceinline__ float to_f32(half_t x) { return (float)x; }
__device__ __forceinline__ float to_f32(double x) { return (float)x; }

__host__ __device__ __forceinline__ uint16_t f32_to_bf16_bits(float f) {
  uint32_t x = __builtin_bit_cast(uint32_t, f);
  if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40);
  uint32_t r = x + 0x7fffu + ((x >> 16) & 1u);
  return (uint16_t)(r >> 16);
}
__host__ __device__ __forceinline__ float bf16_bits_to_f32(uint16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

// OCP e4m3fn, round to nearest even, finite overflow saturates to +-448.
__host__ __device__ __forceinline__ uint8_t f32_to_e4m3_bits(float f) {
  uint32_t x = __builtin_bit_cast(uint32_t, f);
  uint8_t sign = (uint8_t)((x >> 24) & 0x80);
  uint32_t ax = x & 0x7fffffffu;
  if (ax > 0x7f800000u) return sign | 0x7f;
  if (ax >= 0x43e00000u) return sign | 0x7e;
  int e = (int)(ax >> 23);
  if (e == 0) return sign;
  uint32_t mant = (ax & 0x7fffff) | 0x800000;
  int exp_unb = e - 127;
  int s;
  int E;
  if (exp_unb >= -6) {
    s = 20;
    E = exp_unb + 7;
  } else {
    s = 141 - e;
    E = 0;
  }
  uint32_t q, rem, halfv;
  if (s >= 32) {
    q = 0;
    rem = mant;
    halfv = 0xffffffffu;  // mant < 2^24 < half: round down
    if (s == 32) halfv = 0x80000000u;
  } else {
    q = mant >> s;
    rem = mant & ((1u << s) - 1);
    halfv = 1u << (s - 1);
  }
  if (rem > halfv || (rem == halfv && (q & 1))) q += 1;
  if (E > 0) {
    if (q == 16) {
      q = 8;
      E += 1;
    }
    if (E > 15 || (E == 15 && (q & 7) == 7)) return sign | 0x7e;
    return sign | (uint8_t)((E << 3) | (q & 7));
  }
  return sign | (uint8_t)q;  // subnormal (q == 8 encodes the smallest normal)
}
__host__ __device__ __forceinline__ float e4m3_bits_to_f32(uint8_t b) {
  int e = (b >> 3) & 0xf, m = b & 7;
  float v;
  if (e == 0xf && m == 7) return __builtin_nanf("");
  if (e == 0)
    v = (float)m * 0.001953125f;  // 2^-9
  else
    v = __builtin_bit_cast(float, (uint32_t)(e - 7 + 127) << 23) * (1.0f + (float)m * 0.125f);
  return (b & 0x80) ? -v : v;
}

// element type traits ------------------------------------------------------
template <typename T>
struct dtype_of;
template <>
struct dtype_of<float> {
  static constexpr int value = REMIFA_R_32F;
};
template <>
struct dtype_of<half_t> {
  static constexpr int value = REMIFA_R_16F;
};
template <>
struct dtype_of<double> {
  static constexpr int value = REMIFA_R_64F;
};

inline size_t dtype_size(int t) {
  switch (t) {
    case REMIFA_R_32F: return 4;
    case REMIFA_R_16F: return 2;
    case REMIFA_R_64F: return 8;
    case REMIFA_R_16BF: return 2;
    case REMIFA_R_8F_E4M3: return 1;
    default: return 0;
  }
}

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------- internal entry points
// (implemented in the .hip files, called by the drivers and the C-ABI)
int convert_launch(remifa_handle_t h, int src_t, int dst_t, int64_t m, int64_t n, const void* a,
                   int64_t lda, void* b, int64_t ldb);
// the handle's second stream (h->side) and its two events, created on
// first use (drivers.hip)
int side_stream(remifa_handle_t h);
int panel_launch(remifa_handle_t h, int t, int ib, int64_t m, int64_t n, const void* d, int64_t ldd,
                 void* l, int64_t ldl, void* u, int64_t ldu, int* stat);

// C = alpha * op(A) op(B) + beta * C, column-major, no transposes. The compute
// type fixes the arithmetic (see remifa.h); in_t is the storage type of A and
// B, out_t the storage type of C.
int gemm_launch(remifa_handle_t h, int ctype, int in_t, int out_t, int64_t m, int64_t n, int64_t k,
                float alpha, const void* a, int64_t lda, const void* b, int64_t ldb, float beta,
                void* c, int64_t ldc);

// C = alpha A B + beta Cin with the beta term read from another matrix
// (fp16 or fp32, ld ldcin) and the result written to C (matrix-core
// compute types FP32 / TC32 / BF16_TC32; cin = null is gemm_launch)
int gemm_launch_cin(remifa_handle_t h, int ctype, int in_t, int out_t, int64_t m, int64_t n,
                    int64_t k, float alpha, const void* a, int64_t lda, const void* b, int64_t ldb,
                    float beta, const void* cin, int cin_t, int64_t ldcin, void* c, int64_t ldc);

// C0 -= A0 B0 (m0 x n0) and C1 -= A1 B1 (m1 x n1), both with inner size k:
// one launch when the narrow tiles apply, else two gemm_launch calls
int gemm_pair_launch(remifa_handle_t h, int ctype, int in_t, int out_t, int64_t m0, int64_t n0,
                     int64_t k, const void* a0, int64_t lda0, const void* b0, int64_t ldb0,
                     void* c0, int64_t ldc0, int64_t m1, int64_t n1, const void* a1, int64_t lda1,
                     const void* b1, int64_t ldb1, void* c1, int64_t ldc1);

// C0 -= A0 B0 (m0 x n0) and C1 -= A1 B1 (m1 x n1), fp32 C, inner size k:
// one launch when both take unsplit 128 x 128 tiles, else two gemm_launch calls
int gemm_dual_launch(remifa_handle_t h, int ctype, int in_t, int64_t m0, int64_t n0, int64_t k,
                     const void* a0, int64_t lda0, const void* b0, int64_t ldb0, void* c0,
                     int64_t ldc0, int64_t m1, int64_t n1, const void* a1, int64_t lda1,
                     const void* b1, int64_t ldb1, void* c1, int64_t ldc1);

// FP8 operands: per-32-block scaled e4m3 codes of the rows (rows = true:
// element (r, k) at src[r + k * ld]) or columns (element (k, r) at
// src[k + r * ld]) of a column-major fp16 / fp32 matrix; codes q[r * ldq + k],
// E8M0 scale bytes s[r * lds + k / 32].
int fp8_quantize_launch(remifa_handle_t h, int in_t, bool rows, int64_t R, int64_t K,
                        const void* src, int64_t ld, uint8_t* q, int64_t ldq, uint8_t* s,
                        int64_t lds);
// C = alpha A B + beta C from FP8 operands in that format (k a multiple of
// 32); with cin, the beta term reads Cin (fp16 or fp32, cin_t) instead of C
int fp8_gemm_launch(remifa_handle_t h, int out_t, int64_t m, int64_t n, int64_t k, float alpha,
                    const uint8_t* qa, int64_t ldqa, const uint8_t* sa, int64_t ldsa,
                    const uint8_t* qb, int64_t ldqb, const uint8_t* sb, int64_t ldsb, float beta,
                    void* c, int64_t ldc, const void* cin = nullptr, int cin_t = 0,
                    int64_t ldcin = 0);

int copy2d_launch(remifa_handle_t h, size_t elt, int64_t m, int64_t n, const void* a, int64_t lda,
                  void* b, int64_t ldb);

// device address of the static scratch array of remifa_getrs_nopiv_noalloc
// on `device` (current), looked up once per device (solve.hip)
int noalloc_scratch(int device, void** ptr);

}  // namespace remifa
