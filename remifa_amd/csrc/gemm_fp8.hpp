// FP8 update GEMM (included by gemm.hip inside its anonymous namespace,
// after gemm_mfma.hpp). C = alpha A B + beta C with OCP e4m3 operands and
// one power-of-two (E8M0) scale per 32-long K block of every row of A and
// every column of B, on the gfx950 scaled matrix cores
// (v_mfma_scale_f32_32x32x64_f8f6f4), fp32 accumulation.
//
// Operand formats (both K-contiguous, so both fragments are plain 16-byte
// LDS reads):
//   qa[m * ldqa + k]  codes of A (row m, column k), sa[m * ldsa + k / 32]
//   qb[n * ldqb + k]  codes of B (row k, column n), sb[n * ldsb + k / 32]
// A scale byte s stands for 2^(s - 127). The quantizer below produces them
// from column-major fp16 / fp32 matrices: per 32-block, scale 2^(floor(log2
// amax) - 8) (the OCP MX rule for e4m3, emax = 8), codes rounded to nearest
// even with saturation at +-448.
//
// Matrix-core operand layout (32x32x64, 8-bit): lane l holds row l % 32 of
// its operand; bytes 0-15 are K 16 h .. 16 h + 15 and bytes 16-31 are
// K 32 + 16 h .. 32 + 16 h + 15 (h = l / 32), and lane (r + 32 b) supplies
// the scale of row r for the 32-block b. The MFMA computes D^T = B^T A^T as
// the 16-bit kernel does, so the accumulator layout and epilogue match it.

struct Fp8Args {
  int64_t M, N, K;
  const uint8_t* A;
  int64_t lda;
  const uint8_t* SA;
  int64_t ldsa;
  const uint8_t* B;
  int64_t ldb;
  const uint8_t* SB;
  int64_t ldsb;
  float alpha, beta;
  void* C;
  int64_t ldc;
  const void* Cin;  // optional source of the beta term (else C): fp16 or fp32
  int64_t ldcin;
  int cin_half;
  float* W;        // split-K partials [split][N][ldw]
  int64_t ldw;
  int64_t kchunk;  // K range per split (multiple of 128)
  int tiles_m, tiles_n;
};

typedef int v8i32 __attribute__((ext_vector_type(8)));

// v if ok, else zeros (component selects: a select of whole structs would
// turn into an address select and push the staging registers to memory)
__device__ __forceinline__ uint4 keep16(bool ok, uint4 v) {
  return make_uint4(ok ? v.x : 0u, ok ? v.y : 0u, ok ? v.z : 0u, ok ? v.w : 0u);
}

__device__ __forceinline__ v8i32 join16(uint4 lo, uint4 hi) {
  return (v8i32){(int)lo.x, (int)lo.y, (int)lo.z, (int)lo.w,
                 (int)hi.x, (int)hi.y, (int)hi.z, (int)hi.w};
}

// scale bytes of a 128-long K slab (four 32-blocks) of row r, in two steps
// so that nothing waits for the load before the matrix-core work: the word
// (from a clamped, in-bounds address when out of range) and the number of
// its valid blocks; blocks past the K range, and rows past the matrix, then
// read as 2^0 (their codes are 0)
__device__ __forceinline__ uint32_t scale_word(const uint8_t* s, int64_t lds, int64_t r,
                                               int64_t nrows, int64_t kb0, int64_t nkb,
                                               int& valid) {
  const bool ok = r < nrows && kb0 < nkb;
  valid = ok ? (int)(nkb - kb0 < 4 ? nkb - kb0 : 4) : 0;
  return *(const uint32_t*)(s + (ok ? r * lds + kb0 : 0));
}
__device__ __forceinline__ uint32_t scale_mask(uint32_t w, int valid) {
  // bytes >= valid -> 127
  const uint32_t keep = valid >= 4 ? 0xFFFFFFFFu : ((1u << (8 * valid)) - 1u);
  return (w & keep) | (0x7F7F7F7Fu & ~keep);
}

template <typename OutT>
__global__ void __launch_bounds__(256, 2) gemm_fp8_kernel(Fp8Args g) {
  constexpr int BM = 128, BN = 128, BK = 128;  // BK: K per slab (bytes per row)
  constexpr int P = BK + 16;  // LDS row pitch: 16 rows of a fragment read hit distinct banks
  __shared__ __attribute__((aligned(16))) uint8_t As[2][BM * P];
  __shared__ __attribute__((aligned(16))) uint8_t Bs[2][BN * P];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % 2, wn = wave / 2;  // 2 x 2 waves of 64 x 64
  const int l32 = lane & 31, lh = lane >> 5;

  const int nt = g.tiles_m * g.tiles_n;
  const int tile = xcd_tile(blockIdx.x, nt);
  constexpr int GM = 8;  // grouped raster (as the 16-bit kernel)
  const int gsz = GM * g.tiles_n;
  const int grp = tile / gsz, first = grp * GM;
  const int rows = g.tiles_m - first < GM ? g.tiles_m - first : GM;
  const int tm = first + (tile % gsz) % rows, tn = (tile % gsz) / rows;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t kbeg = (int64_t)blockIdx.z * g.kchunk;
  const int64_t kend = kbeg + g.kchunk < g.K ? kbeg + g.kchunk : g.K;
  const int64_t klen = kend > kbeg ? kend - kbeg : 0;
  const int nk = (int)((klen + BK - 1) / BK);
  const bool split = gridDim.z > 1;
  const int64_t nkb = kend / 32;  // 32-blocks below kend (K is a multiple of 32)

  uint4 ra[4], rb[4];
  bool oka[4], okb[4];
  uint32_t nsa[2], nsb[2];  // scale words of the slab being loaded
  int nva[2], nvb[2];       // and their valid block counts
  uint32_t csa[2], csb[2];  // scale words of the slab being computed

  auto load_slab = [&](int kt) {
    const int64_t k0 = kbeg + (int64_t)kt * BK;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int c = tid + 256 * p;
      const int row = c >> 3, kq = c & 7;
      const int64_t gk = k0 + 16 * kq;
      const int64_t gm = m0 + row, gn = n0 + row;
      const bool ka = gm < g.M && gk < kend, kb = gn < g.N && gk < kend;
      ra[p] = *(const uint4*)(g.A + (ka ? gm * g.lda + gk : 0));
      rb[p] = *(const uint4*)(g.B + (kb ? gn * g.ldb + gk : 0));
      oka[p] = ka;
      okb[p] = kb;
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      nsa[i] = scale_word(g.SA, g.ldsa, m0 + wm * 64 + i * 32 + l32, g.M, k0 / 32, nkb, nva[i]);
      nsb[i] = scale_word(g.SB, g.ldsb, n0 + wn * 64 + i * 32 + l32, g.N, k0 / 32, nkb, nvb[i]);
    }
  };
  auto store_tiles = [&](int buf) {
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int c = tid + 256 * p;
      const int row = c >> 3, kq = c & 7;
      *(uint4*)&As[buf][row * P + 16 * kq] = keep16(oka[p], ra[p]);
      *(uint4*)&Bs[buf][row * P + 16 * kq] = keep16(okb[p], rb[p]);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      csa[i] = scale_mask(nsa[i], nva[i]);
      csb[i] = scale_mask(nsb[i], nvb[i]);
    }
  };

  f16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;

  // one 64-long K step (two 32-blocks): opsel picks scale byte 2 s (+ lh via the shift)
  auto kstep = [&](const uint8_t* as, const uint8_t* bs, int s, auto opsel_tag) {
    constexpr int OPS = decltype(opsel_tag)::value;
    v8i32 xb[2], ya[2];
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const uint8_t* p = &bs[(wn * 64 + j * 32 + l32) * P + 64 * s + 16 * lh];
      xb[j] = join16(*(const uint4*)p, *(const uint4*)(p + 32));
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const uint8_t* p = &as[(wm * 64 + i * 32 + l32) * P + 64 * s + 16 * lh];
      ya[i] = join16(*(const uint4*)p, *(const uint4*)(p + 32));
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
            xb[j], ya[i], acc[i][j], 0, 0, OPS, (int)(csb[j] >> (8 * lh)), OPS,
            (int)(csa[i] >> (8 * lh)));
  };
  auto compute = [&](int buf) {
    kstep(As[buf], Bs[buf], 0, std::integral_constant<int, 0>{});
    kstep(As[buf], Bs[buf], 1, std::integral_constant<int, 2>{});
  };

  if (nk > 0) {
    load_slab(0);
    store_tiles(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int buf = kt & 1;
    // straight-line body (after the last slab the loads are masked, clamped
    // to in-bounds addresses, and land in the buffer nobody reads again): a
    // branch around the stores lets the compiler sink the matrix-core work
    // below them, behind the waits for the loads just issued
    load_slab(kt + 1);
    __builtin_amdgcn_sched_barrier(0);  // the loads go out first
    compute(buf);
    __builtin_amdgcn_sched_barrier(0);
    store_tiles(buf ^ 1);
    __syncthreads();
  }

  // epilogue: lane holds C[m = .. + l32][n = .. + 8*(q/4) + 4*lh + q%4];
  // the beta term's source (C, or Cin in fp16 / fp32) is chosen once per
  // workgroup, outside the loops
  auto epilogue = [&](const auto* __restrict__ src, int64_t lds_) {
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int64_t m = m0 + wm * 64 + i * 32 + l32;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int64_t nb = n0 + wn * 64 + j * 32 + 4 * lh;
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int64_t n = nb + 8 * (q >> 2) + (q & 3);
          if (m >= g.M || n >= g.N) continue;
          if (split) {
            g.W[((int64_t)blockIdx.z * g.N + n) * g.ldw + m] = g.alpha * acc[i][j][q];
          } else {
            const float cv = g.beta != 0.0f ? (float)src[m + n * lds_] : 0.0f;
            ((OutT*)g.C)[m + n * g.ldc] =
                (OutT)__builtin_fmaf(g.alpha, acc[i][j][q], g.beta * cv);
          }
        }
      }
    }
  };
  if (split || !g.Cin)
    epilogue((const OutT*)g.C, g.ldc);
  else if (g.cin_half)
    epilogue((const half_t*)g.Cin, g.ldcin);
  else
    epilogue((const float*)g.Cin, g.ldcin);
}

// ---------------------------------------------------------------- quantizer
// One lane per (row, 32-block): amax, the block's scale exponent and the 32
// codes. ROWS: the block runs along row r of a column-major matrix (element
// (r, k) at src[r + k * ld]: A operands); else along column r (element
// (k, r) at src[k + r * ld]: B operands). Codes past K are zero.
template <typename T, bool ROWS>
__global__ void __launch_bounds__(256) quant_fp8_kernel(int64_t R, int64_t K, const T* __restrict__ src,
                                                        int64_t ld, uint8_t* __restrict__ q,
                                                        int64_t ldq, uint8_t* __restrict__ s,
                                                        int64_t lds) {
  const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int64_t k0 = (int64_t)blockIdx.y * 32;
  if (r >= R) return;
  const int kn = K - k0 < 32 ? (int)(K - k0) : 32;
  float v[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) {
    const int jj = j < kn ? j : kn - 1;
    const float x = ROWS ? (float)src[r + (k0 + jj) * ld] : (float)src[k0 + jj + r * ld];
    v[j] = j < kn ? x : 0.0f;
  }
  float amax = 0.0f;
#pragma unroll
  for (int j = 0; j < 32; ++j) amax = fmaxf(amax, fabsf(v[j]));
  // scale 2^(floor(log2 amax) - 8); 2^0 for an all-zero (or non-finite) block
  int se = 0;
  if (amax > 0.0f && amax <= 3.4028235e38f) {
    int e;
    frexpf(amax, &e);
    se = e - 1 - 8;
  }
  uint32_t w[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    uint32_t x = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e)
      x |= (uint32_t)f32_to_e4m3_bits(ldexpf(v[4 * b + e], -se)) << (8 * e);
    w[b] = x;
  }
  uint4* out = (uint4*)(q + r * ldq + k0);
  out[0] = make_uint4(w[0], w[1], w[2], w[3]);
  out[1] = make_uint4(w[4], w[5], w[6], w[7]);
  s[r * lds + k0 / 32] = (uint8_t)(se + 127);
}
