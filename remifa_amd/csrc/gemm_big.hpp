// Large-tile update GEMM for 16-bit operands (fp16, or bf16 storage) with
// fp32 accumulation; included by gemm.hip inside its anonymous namespace,
// after gemm_mfma.hpp (GemmArgs, xcd_tile, c_in16, ds_read_tr16).
//
// One 256-thread workgroup (four wave64s in 2 x 2) owns a 256 x 256 tile of
// C, each wavefront a 128 x 128 sub-tile of 4 x 4 v_mfma_f32_32x32x16 blocks
// (256 fp32 accumulators per lane, one workgroup per CU). Against the
// 128 x 128 tile of gemm_mfma_kernel (64 x 64 per wavefront) this halves both
// the operand bytes a CU pulls from L2 per flop and the LDS fragment reads
// per matrix-core instruction (8 fragments per 16 instructions).
//
// Operand slabs go from global memory straight into LDS, without staging
// registers: buffer_load_dwordx4 ... lds writes each lane's 16 bytes at the
// instruction's LDS base (M0) + 16 x lane, so the lane -> global address map
// chooses the layout. NS slabs of BK k-values are in flight or resident:
// while slab k is multiplied, slabs k+1 .. k+NS-2 are arriving and slab
// k+NS-1 is requested into the buffer slab k-1 used. Per slab one
// `s_waitcnt vmcnt(..)` (this wavefront's requests for slab k have landed)
// and one s_barrier (so have every wavefront's, and nobody still reads the
// buffer about to be refilled).
//
// LDS layouts, in 16-byte units (8 elements), XOR-swizzled so that each
// fragment read touches every bank once:
//  * A slab [k][m]: row k holds 256 m (32 units); unit u of row k sits at
//    position u ^ (4 (k % 4)). An A^T fragment is two transpose reads
//    (ds_read_b64_tr_b16) of four consecutive rows x 32 m: the four rows land
//    in four different 64-byte bank groups.
//  * B slab [n][k]: column n holds BK k (UC = BK / 8 units); unit c of column
//    n sits at position n UC + (c ^ ((n / (16 / UC)) % UC)). A B^T fragment
//    is one 16-byte read per lane; the 16 lanes of a read phase (16
//    consecutive n, one c) hit 16 different 16-byte bank groups.
// The global side of a DMA instruction stays coalesced: it reads whole
// 512-byte rows of A (two per instruction) or whole BK-element column
// segments of B, with the 16-byte pieces permuted among the lanes.
//
// Shapes: M and N multiples of 256, every K range (K, and the split-K chunk)
// a multiple of BK, 16-byte aligned operands with leading dimensions that are
// multiples of 8 (gemm.hip checks, and uses gemm_mfma_kernel otherwise).

template <int BK, int WAVES>
struct BigGeom {
  static_assert(BK == 32 || BK == 64, "slab depth");
  static_assert(WAVES == 4 || WAVES == 8, "wavefronts per workgroup");
  static constexpr int UC = BK / 8;            // 16-byte units per B column
  static constexpr int A_ST = BK * 256;        // elements per A slab
  static constexpr int B_ST = 256 * BK;        // elements per B slab
  static constexpr int ST = A_ST + B_ST;       // elements per stage
  static constexpr int LPW = BK / (2 * WAVES);  // DMA instructions per wave per operand slab
  static constexpr int CPS = 16 / UC;          // B columns per 256-byte bank window
  static constexpr int WN = WAVES / 2;         // wavefronts along n (two along m)
  static constexpr int WTN = 256 / WN;         // columns of a wavefront's sub-tile
  static constexpr int FN = WTN / 32;          // 32-column MFMA blocks per wavefront
};

typedef __attribute__((address_space(3))) void* lds_vptr;

// s_waitcnt vmcnt(N) then s_barrier; the memory clobber keeps the compiler's
// LDS reads of the next slab below it
template <int N>
__device__ __forceinline__ void vm_wait_barrier() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  asm volatile("s_waitcnt vmcnt(%0)\n\ts_barrier" ::"n"(N) : "memory");
}

constexpr int kBigBufWord3 = 0x00020000;  // buffer resource word 3 (raw 32-bit data)

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// buffer resource of a raw byte range starting at p (wave-uniform p)
__device__ __forceinline__ u32x4 big_rsrc(const void* p) {
  const uint64_t a = (uint64_t)p;
  u32x4 r;
  r[0] = __builtin_amdgcn_readfirstlane((uint32_t)a);
  r[1] = __builtin_amdgcn_readfirstlane((uint32_t)(a >> 32) & 0xffffu);
  r[2] = 0x7fffffffu;
  r[3] = (uint32_t)kBigBufWord3;
  return r;
}

// 16 bytes per lane from rsrc + voff into LDS at lds_base + 16 lane. Written
// as one asm block (M0 set and used inside it): the compiler does not track
// this request, so it adds no wait of its own before later LDS reads; the
// kernel's explicit vmcnt waits order them. Nothing is written to registers.
__device__ __forceinline__ void dma16(const u32x4& rsrc, uint32_t voff, uint32_t lds_base) {
  asm volatile("s_mov_b32 m0, %2\n\tbuffer_load_dwordx4 %0, %1, 0 offen lds"
               :
               : "v"(voff), "s"(rsrc), "s"(__builtin_amdgcn_readfirstlane(lds_base))
               : "memory", "m0");
}

template <int OPK, typename OutT, int BK, int NS, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) gemm_big_kernel(GemmArgs g) {
  using GEO = BigGeom<BK, WAVES>;
  constexpr int UC = GEO::UC, LPW = GEO::LPW, FN = GEO::FN, WTN = GEO::WTN;
  static_assert(NS >= 2 && NS * GEO::ST * 2 <= 160 * 1024, "LDS stages");
  static_assert(OPK == OPK_F16 || OPK == OPK_BF16S, "16-bit operands in memory");
  __shared__ __attribute__((aligned(1024))) uint16_t lds[NS * GEO::ST];

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int l32 = lane & 31, lh = lane >> 5;

  // grouped, XCD-aware raster (as gemm_mfma_kernel)
  const int nt = g.tiles_m * g.tiles_n;
  const int tile = xcd_tile(blockIdx.x, nt);
  constexpr int GM = 8;
  const int gsz = GM * g.tiles_n;
  const int grp = tile / gsz, first = grp * GM;
  const int rows = g.tiles_m - first < GM ? g.tiles_m - first : GM;
  const int tm = first + (tile % gsz) % rows, tn = (tile % gsz) / rows;
  const int64_t m0 = (int64_t)tm * 256, n0 = (int64_t)tn * 256;
  const int64_t kbeg = (int64_t)blockIdx.z * g.kchunk;
  const int64_t kend = kbeg + g.kchunk < g.K ? kbeg + g.kchunk : g.K;
  const int nk = kend > kbeg ? (int)((kend - kbeg) / BK) : 0;

  const uint16_t* const A = (const uint16_t*)g.A;
  const uint16_t* const B = (const uint16_t*)g.B;

  // DMA lane offsets in bytes from the slab's corner (fixed for the loop)
  uint32_t offa[LPW], offb[LPW];
#pragma unroll
  for (int t = 0; t < LPW; ++t) {
    const int r = wave * (BK / WAVES) + 2 * t + lh;  // A row (k) of this lane
    const int u = l32 ^ (4 * (r & 3));           // unit loaded into position l32
    offa[t] = (uint32_t)(((int64_t)r * g.lda + 8 * u) * 2);
    const int p = (wave * LPW + t) * 64 + lane;  // B position
    const int n = p / UC, c = (p % UC) ^ ((n / GEO::CPS) % UC);
    offb[t] = (uint32_t)(((int64_t)n * g.ldb + 8 * c) * 2);
  }
  // LDS byte address of the stage area (wave-uniform)
  const uint32_t lds0 = __builtin_amdgcn_readfirstlane((uint32_t)(uintptr_t)(lds_vptr)lds);
  auto issue = [&](int k, int st) {
    const int64_t kk = kbeg + (int64_t)k * BK;
    const u32x4 ra = big_rsrc(A + m0 + kk * g.lda);
    const u32x4 rb = big_rsrc(B + kk + n0 * g.ldb);
    const uint32_t as = lds0 + (uint32_t)(st * GEO::ST * 2);
    const uint32_t bs = as + GEO::A_ST * 2;
#pragma unroll
    for (int t = 0; t < LPW; ++t)
      dma16(ra, offa[t], as + (uint32_t)((wave * (BK / WAVES) + 2 * t) * 512));
#pragma unroll
    for (int t = 0; t < LPW; ++t) dma16(rb, offb[t], bs + (uint32_t)((wave * LPW + t) * 1024));
  };

  // fragment read offsets (elements from the stage base)
  const int gq = lane >> 4, sl = lane & 15;
  const int xr = sl >> 2;  // row % 4 of this lane's transpose reads
  int aoff[4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
    aoff[i] = (8 * (gq >> 1) + xr) * 256 +
              (16 * wm + 4 * (i ^ xr) + 2 * (gq & 1) + ((sl >> 1) & 1)) * 8 + 4 * (sl & 1);
  const int nb0 = WTN * wn + l32;
  const int fb = (nb0 / GEO::CPS) % UC;  // B swizzle of this lane's columns (same for every j)
  int boff[BK / 16];
#pragma unroll
  for (int s = 0; s < BK / 16; ++s) boff[s] = GEO::A_ST + (nb0 * UC + ((2 * s + lh) ^ fb)) * 8;

  f16v acc[4][FN];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j)
#pragma unroll
      for (int q = 0; q < 16; ++q) acc[i][j][q] = 0.0f;

  auto compute = [&](int st) {
    const uint16_t* const base = lds + st * GEO::ST;
#pragma unroll
    for (int s = 0; s < BK / 16; ++s) {
      u16x8 xb[FN], ya[4];
#pragma unroll
      for (int j = 0; j < FN; ++j) xb[j] = *(const u16x8*)&base[boff[s] + j * 32 * UC * 8];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const v4u16 lo = ds_read_tr16(&base[aoff[i] + s * 16 * 256]);
        const v4u16 hi = ds_read_tr16(&base[aoff[i] + s * 16 * 256 + 4 * 256]);
        ya[i] = (u16x8){lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          if constexpr (OPK == OPK_F16)
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                __builtin_bit_cast(h8, xb[j]), __builtin_bit_cast(h8, ya[i]), acc[i][j], 0, 0, 0);
          else
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                __builtin_bit_cast(bf8, xb[j]), __builtin_bit_cast(bf8, ya[i]), acc[i][j], 0, 0, 0);
        }
    }
  };

#pragma unroll
  for (int p = 0; p < NS - 1; ++p)
    if (p < nk) issue(p, p);
  // steady state: slab k landed while NS - 2 later slabs may be in flight;
  // after slab k is multiplied, slab k + NS - 1 goes into the buffer of slab
  // k - 1 (every wavefront finished that one before the barrier). Requesting
  // it after the compute, right before the next wait, also keeps the
  // compiler from placing a full vmcnt wait ahead of the fragment reads (it
  // cannot tell the DMA's LDS target from the buffer being read).
  int k = 0, st = 0;
  for (; k + NS - 1 < nk; ++k) {
    vm_wait_barrier<(NS - 2) * 2 * LPW>();
    compute(st);
    issue(k + NS - 1, st == 0 ? NS - 1 : st - 1);
    st = st == NS - 1 ? 0 : st + 1;
  }
  // the last NS - 1 slabs: everything requested has to land
  for (; k < nk; ++k) {
    vm_wait_barrier<0>();
    compute(st);
    st = st == NS - 1 ? 0 : st + 1;
  }

  // epilogue: lane holds C[m = .. + l32][n = .. + 8 (q / 4) + 4 lh + q % 4].
  // Buffer accesses from the tile's corner: a per-lane 32-bit offset (row
  // and lane half) plus a uniform one per (i, j, q), so no 64-bit address
  // arithmetic per element.
  const int ml = 128 * wm + l32, nl = WTN * wn + 4 * lh;
  auto store_all = [&](auto out_tag, void* base, int64_t ld, auto value) {
    using T = decltype(out_tag);
    const __amdgpu_buffer_rsrc_t rc =
        __builtin_amdgcn_make_buffer_rsrc(base, 0, 0x7fffffff, kBigBufWord3);
    const uint32_t vo = (uint32_t)((ml + (int64_t)nl * ld) * sizeof(T));
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int q = 0; q < 16; ++q) {
          const uint32_t so =
              (uint32_t)((32 * i + (int64_t)(32 * j + 8 * (q >> 2) + (q & 3)) * ld) * sizeof(T));
          const float v = value(i, j, q, so);
          if constexpr (sizeof(T) == 4)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), rc, vo, so, 0);
          else
            __builtin_amdgcn_raw_buffer_store_b16(
                __builtin_bit_cast(uint16_t, (half_t)v), rc, vo, so, 0);
        }
  };
  if (gridDim.z > 1) {
    // split-K partial alpha * acc of slice z
    float* const wb = g.W + ((int64_t)blockIdx.z * g.N + n0) * g.ldw + m0;
    store_all(0.0f, (void*)wb, g.ldw,
              [&](int i, int j, int q, uint32_t) { return g.alpha * acc[i][j][q]; });
  } else if (g.beta == 0.0f) {
    store_all(OutT{}, (void*)((OutT*)g.C + m0 + n0 * g.ldc), g.ldc,
              [&](int i, int j, int q, uint32_t) { return g.alpha * acc[i][j][q]; });
  } else {
    // beta term from C itself or from the separate input (fp16 or fp32)
    const bool half_in = g.Cin ? g.cin_half != 0 : sizeof(OutT) == 2;
    const void* const cb = g.Cin ? g.Cin : (const void*)g.C;
    const int64_t ldi = g.Cin ? g.ldcin : g.ldc;
    const int esz = half_in ? 2 : 4;
    const __amdgpu_buffer_rsrc_t ri = __builtin_amdgcn_make_buffer_rsrc(
        (void*)((const char*)cb + (m0 + n0 * ldi) * esz), 0, 0x7fffffff, kBigBufWord3);
    const uint32_t vi = (uint32_t)((ml + (int64_t)nl * ldi) * esz);
    store_all(OutT{}, (void*)((OutT*)g.C + m0 + n0 * g.ldc), g.ldc,
              [&](int i, int j, int q, uint32_t) {
                const uint32_t si =
                    (uint32_t)((32 * i + (int64_t)(32 * j + 8 * (q >> 2) + (q & 3)) * ldi) * esz);
                float c;
                if (half_in)
                  c = (float)__builtin_bit_cast(
                      half_t, (uint16_t)__builtin_amdgcn_raw_buffer_load_b16(ri, vi, si, 0));
                else
                  c = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(ri, vi, si, 0));
                return __builtin_fmaf(g.alpha, acc[i][j][q], g.beta * c);
              });
  }
}
