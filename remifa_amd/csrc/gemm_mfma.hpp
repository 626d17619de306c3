// Matrix-core update GEMM kernel (included by gemm.hip inside its anonymous
// namespace). C = alpha A B + beta C, column-major, fp32 accumulation.
//
// Structure per 256-thread workgroup (4 wave64s, one BM x BN tile of C):
//  * The MFMA computes D^T = B^T A^T: its first operand is a B^T fragment
//    (lane = column n of C), its second an A^T fragment (lane = row m of C),
//    so a lane's accumulators cover 32 consecutive rows m across the wave and
//    the C traffic coalesces along the column-major storage.
//  * When alpha = +-1 the accumulators start at (beta/alpha) C, so the
//    epilogue only stores.
//  * K slabs of BK go HBM -> registers -> LDS with double buffering and one
//    barrier per slab; the loads of slab k+1 are in flight while the matrix
//    cores work on slab k. VEC kernels (16-byte aligned operands, M and K
//    multiples of the 16-byte chunk) load whole chunks: a chunk is either in
//    range or replaced by zeros through an address clamp and a select, so
//    the loop body has no branches and the prefetch stays asynchronous.
//    Other shapes use the element-wise VEC = false instantiation.
//  * A slabs are stored m-contiguous ([k][m]) and B slabs k-contiguous
//    ([n][k]), both with 16-byte stores; a B fragment is one 16-byte LDS
//    read, an A fragment (fp16/bf16) two LDS transpose reads.
//  * Workgroup ids are remapped so that every XCD (private L2) processes a
//    contiguous range of tiles.

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short v4i16 __attribute__((ext_vector_type(4)));
typedef uint16_t v4u16 __attribute__((ext_vector_type(4)));

// gfx950 LDS transpose read (ds_read_b64_tr_b16). Within each group of 16
// lanes, source lane L supplies an 8-byte address; result lane j receives
// element j % 4 of the addresses of lanes j / 4, j / 4 + 4, j / 4 + 8 and
// j / 4 + 12 (in that order).
__device__ __forceinline__ v4u16 ds_read_tr16(const uint16_t* p) {
  const v4i16 r = __builtin_amdgcn_ds_read_tr16_b64_v4i16(
      (__attribute__((address_space(3))) v4i16*)(p));
  return __builtin_bit_cast(v4u16, r);
}

// operand kinds: fp16, bf16 rounded from fp16 storage, fp32, bf16 storage
enum { OPK_F16 = 0, OPK_BF16 = 1, OPK_F32 = 2, OPK_BF16S = 4 };

struct GemmArgs {
  int64_t M, N, K;
  int64_t lda, ldb, ldc;
  float alpha, beta;
  const void* A;
  const void* B;
  void* C;
  float* W;        // split-K partials [split][N][ldw]
  int64_t ldw;
  int64_t kchunk;  // K range per split (multiple of BK)
  int tiles_m, tiles_n;
  // the beta term from another matrix (C = alpha A B + beta Cin), fp16 or
  // fp32; null: Cin = C
  const void* Cin;
  int64_t ldcin;
  int cin_half;
};

// 16 elements (m, n0 + 8 (q / 4) + q % 4) of the beta operand, as floats; the
// source is chosen once, so the 16 loads of a branch are issued together
template <typename OutT>
__device__ __forceinline__ void c_in16(const GemmArgs& g, int64_t m, int64_t n0, float* cv) {
  auto idx = [&](int q, int64_t ld) {
    const int64_t n = n0 + 8 * (q >> 2) + (q & 3);
    return (m < g.M && n < g.N) ? m + n * ld : 0;
  };
  if (g.Cin == nullptr) {
#pragma unroll
    for (int q = 0; q < 16; ++q) cv[q] = (float)((const OutT*)g.C)[idx(q, g.ldc)];
  } else if (g.cin_half) {
#pragma unroll
    for (int q = 0; q < 16; ++q) cv[q] = (float)((const half_t*)g.Cin)[idx(q, g.ldcin)];
  } else {
#pragma unroll
    for (int q = 0; q < 16; ++q) cv[q] = ((const float*)g.Cin)[idx(q, g.ldcin)];
  }
}

template <int OPK>
struct OpTraits;
template <>
struct OpTraits<OPK_F16> {
  using Store = half_t;  // storage type in HBM
  using Lds = uint16_t;  // element in LDS
  static constexpr int CH = 8;  // elements per 16-byte chunk
  using Chunk = u16x8;
  static __device__ __forceinline__ Chunk cvt(Chunk c) { return c; }
};
template <>
struct OpTraits<OPK_BF16> {
  using Store = half_t;
  using Lds = uint16_t;
  static constexpr int CH = 8;
  using Chunk = u16x8;
  // fp16 storage -> bf16 operand, round to nearest even
  static __device__ __forceinline__ Chunk cvt(Chunk c) {
    Chunk r;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      float f = (float)__builtin_bit_cast(half_t, (uint16_t)c[v]);
      r[v] = f32_to_bf16_bits(f);
    }
    return r;
  }
};
template <>
struct OpTraits<OPK_BF16S> {
  using Store = uint16_t;  // bf16 bits in HBM
  using Lds = uint16_t;
  static constexpr int CH = 8;
  using Chunk = u16x8;
  static __device__ __forceinline__ Chunk cvt(Chunk c) { return c; }
};
template <>
struct OpTraits<OPK_F32> {
  using Store = float;
  using Lds = float;
  static constexpr int CH = 4;
  using Chunk = f32x4;
  static __device__ __forceinline__ Chunk cvt(Chunk c) { return c; }
};

// XCD-aware tile index: workgroup w runs on XCD w % 8; give every XCD a
// contiguous range of tiles (a bijection for any tile count).
__device__ __forceinline__ int xcd_tile(int w, int nt) {
  int q = nt >> 3, r = nt & 7;
  int x = w & 7, s = w >> 3;
  return x * q + (x < r ? x : r) + s;
}

// LDS geometry of a tile configuration (elements of the LDS type)
template <int OPK, int BM, int BN, int BK>
struct TileLdsGeom {
  // A rows 64 B apart modulo the 256-byte bank window (16-bit): the four rows
  // and two m-halves of a transpose read hit all 64 banks once
  static constexpr int PADA = 32;
  static constexpr int PADB = (OPK == OPK_F32) ? 4 : 8;
  static constexpr int SA = BM + PADA;  // As[k][m]
  static constexpr int SB = BK + PADB;  // Bs[n][k]
  static constexpr int A_ELEMS = BK * SA, B_ELEMS = BN * SB;  // per buffer
  static constexpr size_t BYTES =
      2 * (size_t)(A_ELEMS + B_ELEMS) * sizeof(typename OpTraits<OPK>::Lds);
};

// One workgroup's tile: `bid` is its tile index (remapped below), `bz` its
// K slice; As0 / Bs0: two LDS buffers each of A and B slabs.
template <int OPK, typename OutT, int BM, int BN, int WM, int WN, int BK, bool VEC, bool PRELOAD>
__device__ __forceinline__ void gemm_mfma_body(const GemmArgs& g, int bid, int bz, bool split,
                                               typename OpTraits<OPK>::Lds* As0,
                                               typename OpTraits<OPK>::Lds* Bs0) {
  using TR = OpTraits<OPK>;
  using S = typename TR::Store;
  using L = typename TR::Lds;
  using Chunk = typename TR::Chunk;
  constexpr int CH = TR::CH;
  constexpr int WTM = BM / WM, WTN = BN / WN;
  constexpr int FM = WTM / 32, FN = WTN / 32;
  using GEO = TileLdsGeom<OPK, BM, BN, BK>;
  constexpr int SA = GEO::SA, SB = GEO::SB;
  constexpr int TA = BK * BM / CH, TB = BN * BK / CH;  // 16-byte chunks per slab
  constexpr int NA = (TA + 255) / 256;
  constexpr int NB = (TB + 255) / 256;
  static_assert(TA % 64 == 0 && TB % 64 == 0, "slab chunks must fill whole waves");
  static_assert(FM >= 1 && FN >= 1, "wave tile below 32x32");

  auto As = [&](int buf) { return As0 + buf * GEO::A_ELEMS; };
  auto Bs = [&](int buf) { return Bs0 + buf * GEO::B_ELEMS; };

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WM, wn = wave / WM;
  const int l32 = lane & 31, lh = lane >> 5;

  const int nt = g.tiles_m * g.tiles_n;
  // grouped raster: GM tile rows at a time, all their tile columns, so the
  // tiles sharing an A row block (and a B column block) run together on one
  // XCD and meet in its L2
  const int tile = xcd_tile(bid, nt);
  constexpr int GM = 8;
  const int gsz = GM * g.tiles_n;
  const int grp = tile / gsz, first = grp * GM;
  const int rows = g.tiles_m - first < GM ? g.tiles_m - first : GM;
  const int tm = first + (tile % gsz) % rows, tn = (tile % gsz) / rows;
  const int64_t m0 = (int64_t)tm * BM, n0 = (int64_t)tn * BN;
  const int64_t kbeg = (int64_t)bz * g.kchunk;
  const int64_t kend = kbeg + g.kchunk < g.K ? kbeg + g.kchunk : g.K;
  const int64_t klen = kend > kbeg ? kend - kbeg : 0;
  const int nk = (int)((klen + BK - 1) / BK);

  const S* __restrict__ A = (const S*)g.A;
  const S* __restrict__ B = (const S*)g.B;

  Chunk ra[NA], rb[NB];
  // VEC: a chunk out of range was loaded from a clamped address; it is
  // zeroed when stored to LDS, after the compute step, so that the loads stay
  // in flight while the matrix cores work
  bool oka[NA], okb[NB];

  // FULL path addressing: a uniform slab base (advanced by the scalar unit)
  // plus per-lane 32-bit element offsets fixed for the whole loop
  const S* const At = A + m0 + kbeg * g.lda;
  const S* const Bt = B + kbeg + n0 * g.ldb;
  uint32_t offa[NA], offb[NB];
#pragma unroll
  for (int p = 0; p < NA; ++p) {
    const int c = tid + 256 * p;
    offa[p] = (uint32_t)((c % (BM / CH)) * CH + (int64_t)(c / (BM / CH)) * g.lda);
  }
#pragma unroll
  for (int p = 0; p < NB; ++p) {
    const int c = tid + 256 * p;
    offb[p] = (uint32_t)((c % (BK / CH)) * CH + (int64_t)(c / (BK / CH)) * g.ldb);
  }

  // FULL (a tag type): every slab of this tile is in range, so the loads
  // need no address clamps and the stores no zero-selects
  auto load_slab = [&](int kt, auto full_tag) {
    constexpr bool FULL = decltype(full_tag)::value;
    const int64_t kk0 = kbeg + (int64_t)kt * BK;
    if constexpr (VEC && FULL) {
      const S* const as = At + (int64_t)kt * BK * g.lda;
      const S* const bs = Bt + (int64_t)kt * BK;
#pragma unroll
      for (int p = 0; p < NA; ++p) {
        if (TA % 256 && tid + 256 * p >= TA) break;
        ra[p] = *(const Chunk*)(as + offa[p]);
      }
#pragma unroll
      for (int p = 0; p < NB; ++p) {
        if (TB % 256 && tid + 256 * p >= TB) break;
        rb[p] = *(const Chunk*)(bs + offb[p]);
      }
      return;
    }
#pragma unroll
    for (int p = 0; p < NA; ++p) {
      int c = tid + 256 * p;
      if (TA % 256 && c >= TA) break;
      int mq = c % (BM / CH), k = c / (BM / CH);
      int64_t gm = m0 + mq * CH, gk = kk0 + k;
      if constexpr (VEC) {
        const bool ok = gm < g.M && gk < kend;
        ra[p] = *(const Chunk*)(A + (ok ? gm + gk * g.lda : 0));
        oka[p] = ok;
      } else {
        Chunk v;
#pragma unroll
        for (int e = 0; e < CH; ++e) {
          bool ok = gk < kend && gm + e < g.M;
          S x = A[ok ? gm + e + gk * g.lda : 0];
          if constexpr (OPK == OPK_F32)
            v[e] = ok ? x : 0.0f;
          else
            v[e] = ok ? __builtin_bit_cast(uint16_t, x) : (uint16_t)0;
        }
        ra[p] = v;
      }
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      int c = tid + 256 * p;
      if (TB % 256 && c >= TB) break;
      int kq = c % (BK / CH), n = c / (BK / CH);
      int64_t gk = kk0 + kq * CH, gn = n0 + n;
      if constexpr (VEC) {
        const bool ok = gn < g.N && gk < kend;
        rb[p] = *(const Chunk*)(B + (ok ? gk + gn * g.ldb : 0));
        okb[p] = ok;
      } else {
        Chunk v;
#pragma unroll
        for (int e = 0; e < CH; ++e) {
          bool ok = gn < g.N && gk + e < kend;
          S x = B[ok ? gk + e + gn * g.ldb : 0];
          if constexpr (OPK == OPK_F32)
            v[e] = ok ? x : 0.0f;
          else
            v[e] = ok ? __builtin_bit_cast(uint16_t, x) : (uint16_t)0;
        }
        rb[p] = v;
      }
    }
  };
  auto store_tiles = [&](int buf, auto full_tag) {
    constexpr bool FULL = decltype(full_tag)::value;
#pragma unroll
    for (int p = 0; p < NA; ++p) {
      int c = tid + 256 * p;
      if (TA % 256 && c >= TA) break;
      int mq = c % (BM / CH), k = c / (BM / CH);
      Chunk v = ra[p];
      if constexpr (VEC && !FULL) v = oka[p] ? v : (Chunk)0;
      *(Chunk*)&As(buf)[k * SA + mq * CH] = TR::cvt(v);
    }
#pragma unroll
    for (int p = 0; p < NB; ++p) {
      int c = tid + 256 * p;
      if (TB % 256 && c >= TB) break;
      int kq = c % (BK / CH), n = c / (BK / CH);
      Chunk v = rb[p];
      if constexpr (VEC && !FULL) v = okb[p] ? v : (Chunk)0;
      *(Chunk*)&Bs(buf)[n * SB + kq * CH] = TR::cvt(v);
    }
  };

  f16v acc[FM][FN];
  // accumulator start: (beta/alpha) C when that is exact, else 0
  const bool from_c = PRELOAD && !split && g.beta != 0.0f && (g.alpha == 1.0f || g.alpha == -1.0f);
  const float cscale = g.beta * g.alpha;  // beta/alpha for alpha = +-1
  const OutT* __restrict__ Cr = (const OutT*)g.C;
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int64_t m = m0 + wm * WTM + i * 32 + l32;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int64_t nb = n0 + wn * WTN + j * 32 + 4 * lh;
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int64_t n = nb + 8 * (q >> 2) + (q & 3);
        bool ok = from_c && m < g.M && n < g.N;
        float x = (float)Cr[ok ? m + n * g.ldc : 0];
        acc[i][j][q] = ok ? cscale * x : 0.0f;
      }
    }
  }

  auto compute = [&](int buf) {
    const L* as = As(buf);
    const L* bs = Bs(buf);
    if constexpr (OPK == OPK_F32) {
#pragma unroll
      for (int ks = 0; ks < BK; ks += 2) {
        float xb[FN], ya[FM];
#pragma unroll
        for (int j = 0; j < FN; ++j) xb[j] = bs[(wn * WTN + j * 32 + l32) * SB + ks + lh];
#pragma unroll
        for (int i = 0; i < FM; ++i) ya[i] = as[(ks + lh) * SA + wm * WTM + i * 32 + l32];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(xb[j], ya[i], acc[i][j], 0, 0, 0);
      }
    } else {
#pragma unroll
      for (int ks = 0; ks < BK; ks += 16) {
        u16x8 xb[FN], ya[FM];
#pragma unroll
        for (int j = 0; j < FN; ++j)
          xb[j] = *(const u16x8*)&bs[(wn * WTN + j * 32 + l32) * SB + ks + 8 * lh];
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          // A^T fragment (lane: m = l32, k = 8 lh .. 8 lh + 7) from the
          // m-contiguous slab with two LDS transpose reads: in each 16-lane
          // group, source lane L points at 4 m-values of row k = kb + L / 4,
          // and result lane j receives A[mb' + j][kb .. kb + 3]
          const int grp = lane >> 4, sl = lane & 15;
          const int mb = wm * WTM + i * 32 + 16 * (grp & 1) + 4 * (sl & 3);
          const int kb = ks + 8 * (grp >> 1) + (sl >> 2);
          const v4u16 lo = ds_read_tr16(&as[kb * SA + mb]);
          const v4u16 hi = ds_read_tr16(&as[(kb + 4) * SA + mb]);
          ya[i] = (u16x8){lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
        }
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            if constexpr (OPK == OPK_F16)
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                  __builtin_bit_cast(h8, xb[j]), __builtin_bit_cast(h8, ya[i]), acc[i][j], 0, 0, 0);
            else
              acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                  __builtin_bit_cast(bf8, xb[j]), __builtin_bit_cast(bf8, ya[i]), acc[i][j], 0, 0,
                  0);
          }
      }
    }
  };

  auto run = [&](auto full_tag) {
    if (nk > 0) {
      load_slab(0, full_tag);
      store_tiles(0, full_tag);
    }
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int buf = kt & 1;
      const bool more = kt + 1 < nk;
      if (more) load_slab(kt + 1, full_tag);
      compute(buf);
      if (more) store_tiles(buf ^ 1, full_tag);
      __syncthreads();
    }
  };
  // interior tiles (the bulk of the LU's updates) take the unmasked loop
  // (when the lane offsets above fit in 32 bits)
  if (VEC && m0 + BM <= g.M && n0 + BN <= g.N && klen % BK == 0 &&
      (int64_t)BK * g.lda < (1LL << 31) && (int64_t)BN * g.ldb < (1LL << 31))
    run(std::integral_constant<bool, true>{});
  else
    run(std::integral_constant<bool, false>{});

  // epilogue: lane holds C[m = .. + l32][n = .. + 8*(q/4) + 4*lh + q%4]
#pragma unroll
  for (int i = 0; i < FM; ++i) {
    const int64_t m = m0 + wm * WTM + i * 32 + l32;
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int64_t nb = n0 + wn * WTN + j * 32 + 4 * lh;
      if (split) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int64_t n = nb + 8 * (q >> 2) + (q & 3);
          if (m < g.M && n < g.N)
            g.W[((int64_t)bz * g.N + n) * g.ldw + m] = g.alpha * acc[i][j][q];
        }
      } else if (from_c || g.beta == 0.0f) {
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int64_t n = nb + 8 * (q >> 2) + (q & 3);
          if (m < g.M && n < g.N) ((OutT*)g.C)[m + n * g.ldc] = (OutT)(g.alpha * acc[i][j][q]);
        }
      } else {
        // general alpha / beta: read this fragment of C, then write it
        float cv[16];
        c_in16<OutT>(g, m, nb, cv);
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const int64_t n = nb + 8 * (q >> 2) + (q & 3);
          float r = __builtin_fmaf(g.alpha, acc[i][j][q], g.beta * cv[q]);
          if (m < g.M && n < g.N) ((OutT*)g.C)[m + n * g.ldc] = (OutT)r;
        }
      }
    }
  }
}

template <int OPK, typename OutT, int BM, int BN, int WM, int WN, int BK, bool VEC, bool PRELOAD>
__global__ void __launch_bounds__(256) gemm_mfma_kernel(GemmArgs g) {
  using L = typename OpTraits<OPK>::Lds;
  using GEO = TileLdsGeom<OPK, BM, BN, BK>;
  __shared__ __attribute__((aligned(16))) L As[2 * GEO::A_ELEMS];
  __shared__ __attribute__((aligned(16))) L Bs[2 * GEO::B_ELEMS];
  gemm_mfma_body<OPK, OutT, BM, BN, WM, WN, BK, VEC, PRELOAD>(g, blockIdx.x, blockIdx.z,
                                                              gridDim.z > 1, As, Bs);
}

// Two independent GEMMs in one launch (no split-K): workgroups [0, nt0) run
// g0 with a 256x32 tile, the rest g1 with a 32x256 tile. The slice updates
// of the LL inner factorization come in such pairs (the slice's columns,
// M x 64, and its rows, 64 x N, with the same K); one launch lets them fill
// the chip together instead of one after the other.
template <int OPK, typename OutT, int BK, bool VEC>
__global__ void __launch_bounds__(256) gemm_pair_kernel(GemmArgs g0, GemmArgs g1, int nt0) {
  using L = typename OpTraits<OPK>::Lds;
  using G0 = TileLdsGeom<OPK, 256, 32, BK>;
  using G1 = TileLdsGeom<OPK, 32, 256, BK>;
  constexpr int N0 = 2 * (G0::A_ELEMS + G0::B_ELEMS), N1 = 2 * (G1::A_ELEMS + G1::B_ELEMS);
  __shared__ __attribute__((aligned(16))) L smem[N0 > N1 ? N0 : N1];  // one tile's buffers
  const int b = blockIdx.x;
  if (b < nt0)
    gemm_mfma_body<OPK, OutT, 256, 32, 4, 1, BK, VEC, false>(g0, b, 0, false, smem,
                                                             smem + 2 * G0::A_ELEMS);
  else
    gemm_mfma_body<OPK, OutT, 32, 256, 1, 4, BK, VEC, false>(g1, b - nt0, 0, false, smem,
                                                             smem + 2 * G1::A_ELEMS);
}

// Two independent GEMMs on 128 x 128 tiles in one launch (no split-K):
// workgroups [0, nt0) run g0, the rest g1. The LL driver's K = nb parts of
// an outer block's update (the block column, M x nb, and the block row,
// nb x M) run on the handle's stream while the second stream waits for
// them; one launch drops the gap between two.
template <int OPK, typename OutT, int BK, bool VEC>
__global__ void __launch_bounds__(256) gemm_dual_kernel(GemmArgs g0, GemmArgs g1, int nt0) {
  using L = typename OpTraits<OPK>::Lds;
  using GEO = TileLdsGeom<OPK, 128, 128, BK>;
  __shared__ __attribute__((aligned(16))) L As[2 * GEO::A_ELEMS];
  __shared__ __attribute__((aligned(16))) L Bs[2 * GEO::B_ELEMS];
  const int b = blockIdx.x;
  if (b < nt0)
    gemm_mfma_body<OPK, OutT, 128, 128, 2, 2, BK, VEC, false>(g0, b, 0, false, As, Bs);
  else
    gemm_mfma_body<OPK, OutT, 128, 128, 2, 2, BK, VEC, false>(g1, b - nt0, 0, false, As, Bs);
}
