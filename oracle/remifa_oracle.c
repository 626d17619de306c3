/*
 * remifa_oracle.c -- CPU restatement of flipflapflop/remifa's no-pivot LU
 * hot path (see remifa_oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and the
 * cpu_baseline leg of bench.py. Never linked into the product.
 *
 * Arithmetic conventions (the reference leaves them to nvcc / cuBLAS):
 *  - fp32 panel arithmetic: reciprocal 1/u_kk correctly rounded, multiply,
 *    then a -= l*u as one fused multiply-add (nvcc contracts by default).
 *  - fp16 panel arithmetic: every operation is computed exactly and rounded
 *    once to fp16 (a half fma), the reciprocal is RN16(RN32(1/x)).
 *  - GEMM C -= A*B: the product sum is accumulated from zero in k order
 *    (fp32 fma for FP32/TC32/BF16/FP8, exact-then-RN16 for FP16/TC16) and
 *    then subtracted from C once, rounded to C's precision (cuBLAS computes
 *    alpha*(A*B) + beta*C). Operands are rounded to the compute type first
 *    (fp16 for TC32/TC16/FP16, bf16 for BF16TC32, scaled e4m3 for FP8TC32).
 * Built with -ffp-contract=off so that only the explicit fmaf() calls fuse.
 */
#include "remifa_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* threads                                                              */
/* ------------------------------------------------------------------ */
static int g_threads = 0;
void orc_set_threads(int nthreads) { g_threads = nthreads; }
int orc_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}
#define OMP_THREADS num_threads(orc_get_threads())

/* ------------------------------------------------------------------ */
/* minstd_rand0 (std::default_random_engine in libstdc++):             */
/* x <- 16807 x mod (2^31-1); min()=1, max()=2^31-2.                   */
/* ------------------------------------------------------------------ */
#define MINSTD_A 16807ULL
#define MINSTD_M 2147483647ULL

uint64_t orc_minstd_seed(uint64_t seed)
{
    /* linear_congruential_engine::seed: s mod m, and 1 if that is 0
     * (increment c = 0) */
    uint64_t x = seed % MINSTD_M;
    return x == 0 ? 1 : x;
}

static uint64_t mulmod(uint64_t a, uint64_t b) { return (a * b) % MINSTD_M; }

uint64_t orc_minstd_jump(uint64_t state, uint64_t k)
{
    /* state after k more draws: state * A^k mod M */
    uint64_t r = 1, p = MINSTD_A;
    while (k) {
        if (k & 1) r = mulmod(r, p);
        p = mulmod(p, p);
        k >>= 1;
    }
    return mulmod(state, r);
}

float orc_minstd_uniform(uint64_t x)
{
    /* uniform_real_distribution<float>(-1,1)(g) in libstdc++:
     * generate_canonical<float,24>: one draw (log2 r = 30 >= 24),
     * sum = float(x - min), tmp = float((long double)(max-min+1)) = 2^31,
     * ret = sum / tmp, clamped below 1; then ret*(b-a) + a. */
    float sum = (float)(x - 1);
    float tmp = 2147483648.0f;
    float ret = sum / tmp;
    if (ret >= 1.0f) ret = nextafterf(1.0f, 0.0f);
    return ret * 2.0f + (-1.0f);
}

void orc_minstd_stream(uint64_t seed, int64_t count, float *out)
{
    uint64_t x = orc_minstd_seed(seed);
    for (int64_t i = 0; i < count; ++i) {
        x = (MINSTD_A * x) % MINSTD_M;
        out[i] = orc_minstd_uniform(x);
    }
}

/* tests/genmat.hxx:68-102: default-seeded engine, entries drawn row by row
 * (i outer, j inner), then a_ii = m. */
void orc_gen_diagdom(int64_t m, int64_t n, float *a, int64_t lda)
{
    uint64_t x = 1;
    for (int64_t i = 0; i < m; ++i) {
        for (int64_t j = 0; j < n; ++j) {
            x = (MINSTD_A * x) % MINSTD_M;
            a[j * lda + i] = orc_minstd_uniform(x);
        }
        if (i < n) a[i * lda + i] = (float)m;
    }
}

/* The alpha generator of the benchmark configurations: off-diagonal
 * entries as gen_diagdom (row-major draws from an engine seeded with
 * `seed`), then a_ii = fp32(alpha * sum_{j != i} |a_ij|) with the row sum in
 * FP64. Every |a_ij| is a multiple of 2^-30 below 1, so the FP64 row sum is
 * exact for n < 2^23 whatever the summation order. */
void orc_gen_diagdom_alpha(int64_t n, uint64_t seed, double alpha, float *a,
                           int64_t lda)
{
    uint64_t s0 = orc_minstd_seed(seed);
#pragma omp parallel for OMP_THREADS schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        uint64_t x = orc_minstd_jump(s0, (uint64_t)(i * n));
        double s = 0.0;
        for (int64_t j = 0; j < n; ++j) {
            x = (MINSTD_A * x) % MINSTD_M;
            float v = orc_minstd_uniform(x);
            a[j * lda + i] = v;
            if (j != i) s += fabs((double)v);
        }
        a[i * lda + i] = (float)(alpha * s);
    }
}

/* libstdc++ uniform_real_distribution<float>(lo, hi) for one engine state:
 * generate_canonical<float> * (hi - lo) + lo, all in float. */
static float minstd_uniform_ab(uint64_t x, float lo, float hi)
{
    float sum = (float)(x - 1);
    float ret = sum / 2147483648.0f;
    if (ret >= 1.0f) ret = nextafterf(1.0f, 0.0f);
    return ret * (hi - lo) + lo;
}

/* tests/common.hxx:298-319 gen_randmat: engine seeded with `seed`,
 * uniform(-1,1) drawn column by column (j outer, i inner). */
void orc_gen_randmat(int64_t m, int64_t n, uint64_t seed, float *a, int64_t lda)
{
    uint64_t x = orc_minstd_seed(seed);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            x = (MINSTD_A * x) % MINSTD_M;
            a[j * lda + i] = minstd_uniform_ab(x, -1.0f, 1.0f);
        }
}

/* tests/common.hxx:391-455 gen_arrowhead with k = 256: zero matrix; in each
 * k x k diagonal block, for j then i over the block, a(i,j) and a(j,i) take
 * the next two draws of a default-seeded engine (uniform(lo, hi)); then for
 * every column j and rows i in [max(m - k, j), m): a(i,j) and a(j,i) take
 * the next two draws (the arrow); finally a_jj = 2k. */
void orc_gen_arrowhead(int64_t m, int64_t n, float lo, float hi, float *a, int64_t lda)
{
    const int64_t k = 256;
    uint64_t x = 1;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) a[j * lda + i] = 0.0f;
    int64_t nc = 1 + (n - 1) / k;
    for (int64_t c = 0; c < nc; ++c) {
        int64_t je = n < (c + 1) * k ? n : (c + 1) * k;
        int64_t ie = m < (c + 1) * k ? m : (c + 1) * k;
        for (int64_t j = c * k; j < je; ++j)
            for (int64_t i = c * k; i < ie; ++i) {
                x = (MINSTD_A * x) % MINSTD_M;
                a[j * lda + i] = minstd_uniform_ab(x, lo, hi);
                x = (MINSTD_A * x) % MINSTD_M;
                a[i * lda + j] = minstd_uniform_ab(x, lo, hi);
            }
    }
    for (int64_t j = 0; j < n; ++j) {
        int64_t i0 = m - k > j ? m - k : j;
        for (int64_t i = i0; i < m; ++i) {
            x = (MINSTD_A * x) % MINSTD_M;
            a[j * lda + i] = minstd_uniform_ab(x, lo, hi);
            x = (MINSTD_A * x) % MINSTD_M;
            a[i * lda + j] = minstd_uniform_ab(x, lo, hi);
        }
        a[j * lda + j] = (float)(2 * k);
    }
}

/* Seeded uniform(-1,1) block (right-hand sides), row-major draw order. */
void orc_gen_uniform(int64_t m, int64_t n, uint64_t seed, float *b,
                     int64_t ldb)
{
    uint64_t x = orc_minstd_seed(seed);
    for (int64_t i = 0; i < m; ++i)
        for (int64_t j = 0; j < n; ++j) {
            x = (MINSTD_A * x) % MINSTD_M;
            b[j * ldb + i] = orc_minstd_uniform(x);
        }
}

/* ------------------------------------------------------------------ */
/* conversions                                                         */
/* ------------------------------------------------------------------ */
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* round a non-negative integer mantissa `v` right by `s` bits, RNE */
static uint64_t rne_shift(uint64_t v, int s)
{
    if (s <= 0) return v << (-s);
    if (s >= 64) return 0;
    uint64_t q = v >> s;
    uint64_t rem = v & ((1ULL << s) - 1);
    uint64_t half = 1ULL << (s - 1);
    if (rem > half || (rem == half && (q & 1))) q += 1;
    return q;
}

uint16_t orc_f32_to_f16(float f)
{
    uint32_t x = f2u(f);
    uint16_t sign = (uint16_t)((x >> 16) & 0x8000);
    uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) /* NaN: quiet, keep top payload bits */
        return sign | 0x7e00 | (uint16_t)((ax >> 13) & 0x3ff);
    if (ax == 0x7f800000u) return sign | 0x7c00;
    int e = (int)(ax >> 23);
    uint64_t mant = ax & 0x7fffff;
    if (e == 0) return sign; /* fp32 subnormals are < 2^-126: round to 0 */
    mant |= 0x800000;        /* value = mant * 2^(e-150) */
    /* fp16 value = q * 2^(E-25) for normals (E = biased exp, q in
     * [2^10, 2^11)), q * 2^-24 for subnormals. Choose the quantum. */
    int exp_unb = e - 127;
    if (exp_unb >= -14) {
        /* normal range: quantum 2^(exp_unb-10); shift = 13 */
        uint64_t q = rne_shift(mant, 13); /* q in [2^10, 2^11] */
        int E = exp_unb + 15;
        if (q == (1ULL << 11)) { q >>= 1; E += 1; }
        if (E >= 31) return sign | 0x7c00; /* overflow to inf */
        return sign | (uint16_t)((E << 10) | (q & 0x3ff));
    }
    /* subnormal range: quantum 2^-24; value/2^-24 = mant * 2^(e-126) */
    uint64_t q = rne_shift(mant, 126 - e);
    /* q may round up to 1024 = smallest normal: encoding works directly */
    return sign | (uint16_t)q;
}

float orc_f16_to_f32(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000) << 16;
    uint32_t e = (h >> 10) & 0x1f;
    uint32_t m = h & 0x3ff;
    if (e == 0x1f) return u2f(sign | 0x7f800000u | (m << 13));
    if (e == 0) {
        float v = (float)m * 0x1p-24f; /* exact */
        return sign ? -v : v;
    }
    return u2f(sign | ((e - 15 + 127) << 23) | (m << 13));
}

uint16_t orc_f64_to_f16(double d)
{
    uint64_t x;
    memcpy(&x, &d, 8);
    uint16_t sign = (uint16_t)((x >> 48) & 0x8000);
    uint64_t ax = x & 0x7fffffffffffffffULL;
    if (ax > 0x7ff0000000000000ULL)
        return sign | 0x7e00 | (uint16_t)((ax >> 42) & 0x3ff);
    if (ax == 0x7ff0000000000000ULL) return sign | 0x7c00;
    int e = (int)(ax >> 52);
    if (e == 0) return sign;
    uint64_t mant = (ax & 0xfffffffffffffULL) | (1ULL << 52);
    int exp_unb = e - 1023;
    if (exp_unb >= -14) {
        uint64_t q = rne_shift(mant, 42);
        int E = exp_unb + 15;
        if (q == (1ULL << 11)) { q >>= 1; E += 1; }
        if (E >= 31) return sign | 0x7c00;
        return sign | (uint16_t)((E << 10) | (q & 0x3ff));
    }
    /* value/2^-24 = mant * 2^(e-1023-52+24) */
    int s = 1023 + 52 - 24 - e;
    uint64_t q = rne_shift(mant, s);
    return sign | (uint16_t)q;
}

uint16_t orc_f32_to_bf16(float f)
{
    uint32_t x = f2u(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40);
    uint32_t r = x + 0x7fffu + ((x >> 16) & 1u);
    return (uint16_t)(r >> 16);
}

float orc_bf16_to_f32(uint16_t h) { return u2f((uint32_t)h << 16); }

/* OCP e4m3fn: bias 7, no infinities, 0x7f/0xff = NaN, max 448.
 * Round to nearest even; finite overflow saturates to +-448. */
uint8_t orc_f32_to_e4m3(float f)
{
    uint32_t x = f2u(f);
    uint8_t sign = (uint8_t)((x >> 24) & 0x80);
    uint32_t ax = x & 0x7fffffffu;
    if (ax > 0x7f800000u) return sign | 0x7f;
    if (ax >= 0x43e00000u) return sign | 0x7e; /* >= 448 (incl. inf) */
    int e = (int)(ax >> 23);
    if (e == 0) return sign;
    uint64_t mant = (ax & 0x7fffff) | 0x800000;
    int exp_unb = e - 127;
    if (exp_unb >= -6) {
        uint64_t q = rne_shift(mant, 20); /* 3 mantissa bits */
        int E = exp_unb + 7;
        if (q == 16) { q = 8; E += 1; }
        if (E > 15 || (E == 15 && (q & 7) == 7)) return sign | 0x7e;
        return sign | (uint8_t)((E << 3) | (q & 7));
    }
    /* subnormal quantum 2^-9: value/2^-9 = mant * 2^(e-127-23+9) */
    uint64_t q = rne_shift(mant, 127 + 23 - 9 - e);
    return sign | (uint8_t)q;
}

float orc_e4m3_to_f32(uint8_t b)
{
    int sign = b & 0x80;
    int e = (b >> 3) & 0xf, m = b & 7;
    float v;
    if (e == 0xf && m == 7) return NAN;
    if (e == 0) v = (float)m * 0x1p-9f;
    else v = ldexpf(1.0f + (float)m / 8.0f, e - 7);
    return sign ? -v : v;
}

float orc_round(int prec, float x)
{
    switch (prec) {
    case ORC_F16: return orc_f16_to_f32(orc_f32_to_f16(x));
    case ORC_BF16: return orc_bf16_to_f32(orc_f32_to_bf16(x));
    case ORC_E4M3: return orc_e4m3_to_f32(orc_f32_to_e4m3(x));
    default: return x;
    }
}

static inline float r16(float x) { return orc_f16_to_f32(orc_f32_to_f16(x)); }
/* one rounding from a double value straight to fp16 */
static inline float r16d(double x) { return orc_f16_to_f32(orc_f64_to_f16(x)); }

static size_t elt_size(int p)
{
    switch (p) {
    case ORC_F32: return 4;
    case ORC_F64: return 8;
    case ORC_F16: case ORC_BF16: return 2;
    case ORC_E4M3: return 1;
    default: return 0;
    }
}

static double load_elt(int p, const void *a, size_t idx)
{
    switch (p) {
    case ORC_F32: return ((const float *)a)[idx];
    case ORC_F64: return ((const double *)a)[idx];
    case ORC_F16: return orc_f16_to_f32(((const uint16_t *)a)[idx]);
    case ORC_BF16: return orc_bf16_to_f32(((const uint16_t *)a)[idx]);
    case ORC_E4M3: return orc_e4m3_to_f32(((const uint8_t *)a)[idx]);
    }
    return 0;
}

static void store_elt(int p, void *a, size_t idx, double v)
{
    switch (p) {
    case ORC_F32: ((float *)a)[idx] = (float)v; break;
    case ORC_F64: ((double *)a)[idx] = v; break;
    case ORC_F16: ((uint16_t *)a)[idx] = orc_f64_to_f16(v); break;
    case ORC_BF16: ((uint16_t *)a)[idx] = orc_f32_to_bf16((float)v); break;
    case ORC_E4M3: ((uint8_t *)a)[idx] = orc_f32_to_e4m3((float)v); break;
    }
}

/* src/convert.cu:324-360. Narrowing is one RNE rounding from the source
 * value (f64->f16 is direct, like the reference's convert<double,half>).
 * Same-type conversion copies (the reference's convert<T,T> is a no-op,
 * marked FIXME at src/convert.cu:378-387). */
int orc_convert(int src, int dst, int64_t m, int64_t n, const void *a,
                int64_t lda, void *b, int64_t ldb)
{
    if (!elt_size(src)) return -1;
    if (!elt_size(dst)) return -2;
    if (m < 0) return -3;
    if (n < 0) return -4;
    if (lda < (m > 1 ? m : 1)) return -6;
    if (ldb < (m > 1 ? m : 1)) return -8;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) {
            size_t ia = (size_t)(j * lda + i), ib = (size_t)(j * ldb + i);
            if (src == ORC_F32 && dst == ORC_F16)
                ((uint16_t *)b)[ib] = orc_f32_to_f16(((const float *)a)[ia]);
            else
                store_elt(dst, b, ib, load_elt(src, a, ia));
        }
    return 0;
}

/* ------------------------------------------------------------------ */
/* precision-parametrised scalar ops for the panel                      */
/* ------------------------------------------------------------------ */
static inline float p_rcp(int prec, float x)
{
    float r = 1.0f / x;
    return prec == ORC_F16 ? r16(r) : r;
}
static inline float p_mul(int prec, float a, float b)
{
    float r = a * b; /* exact for fp16 operands */
    return prec == ORC_F16 ? r16(r) : r;
}
static inline float p_fnma(int prec, float a, float b, float c)
{
    /* c - a*b fused */
    if (prec == ORC_F16) return r16d((double)c - (double)a * (double)b);
    return fmaf(-a, b, c);
}

/* ------------------------------------------------------------------ */
/* panel: src/lu_nopiv.cu:100-358 on float arrays, precision `prec`    */
/* ------------------------------------------------------------------ */
/* d: n x n diagonal tile snapshot (ld ldd), l: m x n block column whose top
 * n x n tile receives the factored diagonal tile, u: n x m block row whose
 * columns n..m-1 are solved (U12 = L11^-1 A12). n <= ib. */
static void panel_f(int prec, int64_t m, int64_t n, const float *d,
                    int64_t ldd, float *l, int64_t ldl, float *u,
                    int64_t ldu)
{
    float t[32 * 32], r[32];
    /* factor the diagonal tile (every thread block does this redundantly,
     * src/lu_nopiv.cu:123-170 / :274-316) */
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) t[i + 32 * j] = d[i + ldd * j];
    for (int64_t k = 0; k < n; ++k) {
        r[k] = p_rcp(prec, t[k + 32 * k]);
        for (int64_t i = k + 1; i < n; ++i)
            t[i + 32 * k] = p_mul(prec, t[i + 32 * k], r[k]);
        for (int64_t j = k + 1; j < n; ++j)
            for (int64_t i = k + 1; i < n; ++i)
                t[i + 32 * j] = p_fnma(prec, t[i + 32 * k], t[k + 32 * j],
                                       t[i + 32 * j]);
    }
    /* L21 rows (blockIdx.y == 0): scale by 1/u_kk, then rank-1 update */
#pragma omp parallel for OMP_THREADS schedule(static) if (m > 4096)
    for (int64_t i = n; i < m; ++i) {
        for (int64_t k = 0; k < n; ++k) {
            float lik = p_mul(prec, l[i + ldl * k], r[k]);
            l[i + ldl * k] = lik;
            for (int64_t j = k + 1; j < n; ++j)
                l[i + ldl * j] = p_fnma(prec, lik, t[k + 32 * j], l[i + ldl * j]);
        }
    }
    /* U12 columns (blockIdx.y == 1, only when m > n): forward substitution
     * with the unit lower L11 */
#pragma omp parallel for OMP_THREADS schedule(static) if (m > 4096)
    for (int64_t j = n; j < m; ++j)
        for (int64_t k = 0; k < n; ++k) {
            float ukj = u[k + ldu * j];
            for (int64_t i = k + 1; i < n; ++i)
                u[i + ldu * j] = p_fnma(prec, t[i + 32 * k], ukj, u[i + ldu * j]);
        }
    /* diagonal tile stored by block (0,0) (src/lu_nopiv.cu:81-87) */
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) l[i + ldl * j] = t[i + 32 * j];
}

int orc_lu_nopiv_panel(int prec, int64_t ib, int64_t m, int64_t n,
                       const void *d, int64_t ldd, void *l, int64_t ldl,
                       void *u, int64_t ldu, int *stat)
{
    if (prec != ORC_F32 && prec != ORC_F16 && prec != ORC_F64) return -1;
    if (ib < 1 || ib > 32) return -2;
    if (m < 0) return -3;
    if (n < 0 || n > ib || n > m) return -4;
    if (ldd < (n > 1 ? n : 1)) return -6;
    if (ldl < (m > 1 ? m : 1)) return -8;
    if (ldu < (n > 1 ? n : 1)) return -10;
    if (n == 0) return 0;
    if (prec == ORC_F64) {
        /* double panel: same algorithm in double arithmetic */
        const double *dd = (const double *)d;
        double *ll = (double *)l, *uu = (double *)u;
        double t[32 * 32], r[32];
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) t[i + 32 * j] = dd[i + ldd * j];
        for (int64_t k = 0; k < n; ++k) {
            r[k] = 1.0 / t[k + 32 * k];
            for (int64_t i = k + 1; i < n; ++i) t[i + 32 * k] *= r[k];
            for (int64_t j = k + 1; j < n; ++j)
                for (int64_t i = k + 1; i < n; ++i)
                    t[i + 32 * j] = fma(-t[i + 32 * k], t[k + 32 * j], t[i + 32 * j]);
        }
        for (int64_t i = n; i < m; ++i)
            for (int64_t k = 0; k < n; ++k) {
                double lik = ll[i + ldl * k] * r[k];
                ll[i + ldl * k] = lik;
                for (int64_t j = k + 1; j < n; ++j)
                    ll[i + ldl * j] = fma(-lik, t[k + 32 * j], ll[i + ldl * j]);
            }
        for (int64_t j = n; j < m; ++j)
            for (int64_t k = 0; k < n; ++k)
                for (int64_t i = k + 1; i < n; ++i)
                    uu[i + ldu * j] = fma(-t[i + 32 * k], uu[k + ldu * j], uu[i + ldu * j]);
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < n; ++i) ll[i + ldl * j] = t[i + 32 * j];
        if (stat) *stat = (int)n;
        return 0;
    }
    /* fp32 / fp16: widen to float arrays, run, narrow back */
    int64_t mu = m; /* u columns span 0..m-1 */
    float *df = malloc(sizeof(float) * n * n);
    float *lf = malloc(sizeof(float) * m * n);
    float *uf = malloc(sizeof(float) * n * mu);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) df[i + n * j] = (float)load_elt(prec, d, j * ldd + i);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) lf[i + m * j] = (float)load_elt(prec, l, j * ldl + i);
    for (int64_t j = n; j < mu; ++j)
        for (int64_t i = 0; i < n; ++i) uf[i + n * j] = (float)load_elt(prec, u, j * ldu + i);
    panel_f(prec, m, n, df, n, lf, m, uf, n);
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) store_elt(prec, l, j * ldl + i, lf[i + m * j]);
    for (int64_t j = n; j < mu; ++j)
        for (int64_t i = 0; i < n; ++i) store_elt(prec, u, j * ldu + i, uf[i + n * j]);
    free(df); free(lf); free(uf);
    if (stat) *stat = (int)n; /* src/lu_nopiv.cu:173-174 */
    return 0;
}

/* ------------------------------------------------------------------ */
/* GEMM update C -= A*B (src/gemm.cu; alpha=-1, beta=1 at every call)   */
/* ------------------------------------------------------------------ */
static int op_prec(int ctype)
{
    switch (ctype) {
    case ORC_TC32: case ORC_TC16: case ORC_FP16: return ORC_F16;
    case ORC_BF16TC32: return ORC_BF16;
    case ORC_FP8TC32: return ORC_E4M3;
    default: return ORC_F32;
    }
}

/* power-of-two scale for an e4m3 block with max magnitude amax:
 * 2^(floor(log2 amax) - 8), so that amax / scale < 2^9 and the largest
 * element lands in [256, 512) before saturation at 448 (OCP MX rule with
 * emax(e4m3) = 8). Returns 1 for an all-zero block. */
static float e4m3_block_scale(float amax)
{
    if (!(amax > 0.0f) || !isfinite(amax)) return 1.0f;
    int e;
    frexpf(amax, &e); /* amax = f * 2^e, f in [0.5,1): floor(log2) = e-1 */
    return ldexpf(1.0f, e - 1 - 8);
}

/* codes and scale exponents of the e4m3 operand format: for row r (rows =
 * 1: a is m x k, the A operand) or column r (rows = 0: a is k x m, the B
 * operand) and each 32-long k block b starting at k = 0, sexp[r * nkb + b] =
 * log2 of the block scale and codes[r * k + kk] = e4m3(a / scale). */
void orc_quantise_e4m3_codes(int64_t m, int64_t k, const float *a, int64_t lda, uint8_t *codes,
                             int32_t *sexp, int rows)
{
    int64_t nkb = (k + 31) / 32;
    for (int64_t r = 0; r < m; ++r)
        for (int64_t k0 = 0; k0 < k; k0 += 32) {
            int64_t k1 = k0 + 32 < k ? k0 + 32 : k;
            float amax = 0.0f;
            for (int64_t kk = k0; kk < k1; ++kk) {
                float v = rows ? a[r + lda * kk] : a[kk + lda * r];
                amax = fmaxf(amax, fabsf(v));
            }
            float s = e4m3_block_scale(amax);
            int e;
            frexpf(s, &e);
            sexp[r * nkb + k0 / 32] = e - 1;
            for (int64_t kk = k0; kk < k1; ++kk) {
                float v = rows ? a[r + lda * kk] : a[kk + lda * r];
                codes[r * k + kk] = orc_f32_to_e4m3(v / s);
            }
        }
}

/* quantise operand panels. a_q: m x k, rows scaled per 32-long k block;
 * b_q: k x n, columns scaled per 32-long k block. Blocks start at k = 0
 * of the given panels (callers align k to the factor's column index). */
static void quantise_e4m3(int64_t m, int64_t k, const float *a, int64_t lda,
                          float *aq, int rows)
{
    /* rows = 1: blocks run along a row (a is m x k, scale per row+kblock)
     * rows = 0: a is k x m (B operand), scale per column+kblock */
    int64_t nkb = (k + 31) / 32;
    uint8_t *codes = malloc((size_t)(m * k > 0 ? m * k : 1));
    int32_t *sexp = malloc(sizeof(int32_t) * (size_t)(m * nkb > 0 ? m * nkb : 1));
    orc_quantise_e4m3_codes(m, k, a, lda, codes, sexp, rows);
    for (int64_t r = 0; r < m; ++r)
        for (int64_t kk = 0; kk < k; ++kk) {
            float q = ldexpf(orc_e4m3_to_f32(codes[r * k + kk]), sexp[r * nkb + kk / 32]);
            if (rows) aq[r + m * kk] = q; else aq[kk + k * r] = q;
        }
    free(codes);
    free(sexp);
}

/* order of the fp32-accumulated product sums of orc_gemm_sub: 0 = k
 * ascending (the restatement), 1 = k descending. The second order is a
 * checker's tool: two results with the same rounding points and different
 * summation orders show how far summation order alone moves the factors. */
static int g_sum_order = 0;
void orc_set_sum_order(int order) { g_sum_order = order; }

void orc_gemm_sub(int ctype, int64_t m, int64_t n, int64_t k,
                  const float *a, int64_t lda, const float *b, int64_t ldb,
                  float *c, int64_t ldc, int cprec)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    int opp = op_prec(ctype);
    const float *A = a, *B = b;
    int64_t LDA = lda, LDB = ldb;
    float *ta = NULL, *tb = NULL;
    if (opp != ORC_F32) {
        ta = malloc(sizeof(float) * m * k);
        tb = malloc(sizeof(float) * k * n);
        if (opp == ORC_E4M3) {
            quantise_e4m3(m, k, a, lda, ta, 1);
            quantise_e4m3(n, k, b, ldb, tb, 0);
        } else {
#pragma omp parallel for OMP_THREADS schedule(static) if (m * k > 65536)
            for (int64_t kk = 0; kk < k; ++kk)
                for (int64_t i = 0; i < m; ++i) ta[i + m * kk] = orc_round(opp, a[i + lda * kk]);
#pragma omp parallel for OMP_THREADS schedule(static) if (k * n > 65536)
            for (int64_t j = 0; j < n; ++j)
                for (int64_t kk = 0; kk < k; ++kk) tb[kk + k * j] = orc_round(opp, b[kk + ldb * j]);
        }
        A = ta; LDA = m; B = tb; LDB = k;
    }
    int half_acc = (ctype == ORC_FP16 || ctype == ORC_TC16);
    const int64_t MB = 128; /* row block kept in L1 */
    int64_t nib = (m + MB - 1) / MB;
#pragma omp parallel for OMP_THREADS schedule(dynamic) collapse(2) if (m * n * k > 262144)
    for (int64_t ibk = 0; ibk < nib; ++ibk)
        for (int64_t j0 = 0; j0 < n; j0 += 4) {
            int64_t i0 = ibk * MB, i1 = i0 + MB < m ? i0 + MB : m;
            int64_t jn = j0 + 4 < n ? 4 : n - j0;
            float acc[4][MB];
            memset(acc, 0, sizeof(acc));
            for (int64_t kq = 0; kq < k; ++kq) {
                const int64_t kk = (g_sum_order && !half_acc) ? k - 1 - kq : kq;
                const float *ac = A + LDA * kk;
                for (int64_t jj = 0; jj < jn; ++jj) {
                    float bv = B[kk + LDB * (j0 + jj)];
                    float *accj = acc[jj];
                    if (half_acc) {
                        for (int64_t i = i0; i < i1; ++i)
                            accj[i - i0] = r16d((double)ac[i] * bv + accj[i - i0]);
                    } else {
                        for (int64_t i = i0; i < i1; ++i)
                            accj[i - i0] = fmaf(ac[i], bv, accj[i - i0]);
                    }
                }
            }
            for (int64_t jj = 0; jj < jn; ++jj) {
                float *cc = c + ldc * (j0 + jj);
                for (int64_t i = i0; i < i1; ++i) {
                    float v = cc[i] - acc[jj][i - i0];
                    cc[i] = orc_round(cprec, v);
                }
            }
        }
    free(ta);
    free(tb);
}

/* round a rectangular block of a float array to precision p */
static void round_block(int p, int64_t m, int64_t n, float *a, int64_t lda)
{
    if (p == ORC_F32 || m <= 0 || n <= 0) return;
#pragma omp parallel for OMP_THREADS schedule(static) if (m * n > 65536)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) a[i + lda * j] = orc_round(p, a[i + lda * j]);
}

static int valid_ctype(int c) { return c >= ORC_FP32 && c <= ORC_FP8TC32; }
static int valid_prec(int p) { return p == ORC_F32 || p == ORC_F16; }

static int check_cfg(const orc_lu_cfg *cfg)
{
    if (!valid_prec(cfg->W)) return -1;
    if (!valid_ctype(cfg->I) || !valid_ctype(cfg->O)) return -1;
    if (!valid_prec(cfg->OP) || !valid_prec(cfg->IP)) return -1;
    if (!valid_prec(cfg->OF) || !valid_prec(cfg->IF)) return -1;
    if (cfg->ib < 1 || cfg->ib > 32) return -1;
    /* fp16-accumulating updates need fp16 output (src/lu_nopiv_ll.hxx:78-88) */
    if ((cfg->O == ORC_FP16 || cfg->O == ORC_TC16) && cfg->OP != ORC_F16) return -1;
    if ((cfg->I == ORC_FP16 || cfg->I == ORC_TC16) && cfg->IP != ORC_F16) return -1;
    /* fp32 storage needs fp32 panels and factors (src/lu_nopiv_inner_ll.hxx:74-100) */
    if (cfg->W == ORC_F32 && (cfg->OP != ORC_F32 || cfg->IP != ORC_F32 ||
                              cfg->OF != ORC_F32 || cfg->IF != ORC_F32))
        return -1;
    if (cfg->I == ORC_FP8TC32) return -1; /* fp8 only for the outer update */
    return 0;
}

/* ------------------------------------------------------------------ */
/* inner left-looking factorization of one outer block                 */
/* src/lu_nopiv_inner_ll.hxx:37-372                                     */
/* ------------------------------------------------------------------ */
/* p: points at the outer block's top-left corner inside the work array
 * (ld ldp). Block column: rows 0..m-1, cols 0..n-1; block row: rows
 * 0..n-1, cols n..m-1. The values in the block are in OF precision.
 * Operands of the inner updates are the factors of this outer block
 * rounded to the inner compute type. */
static void inner_ll(const orc_lu_cfg *cfg, int64_t m, int64_t n, float *p,
                     int64_t ldp)
{
    const int64_t ib = cfg->ib;
    const int64_t inc = (n - 1) / ib + 1;
    float dsnap[32 * 32];
    for (int64_t k = 0; k < inc; ++k) {
        int64_t iofs = k * ib, cblkm = m - iofs;
        int64_t cblkn = n - iofs < ib ? n - iofs : ib;
        if (iofs > 0) {
            /* L update: p[iofs:m, kb] -= L[iofs:m, 0:iofs] * U[0:iofs, kb] */
            orc_gemm_sub(cfg->I, cblkm, cblkn, iofs, p + iofs, ldp,
                         p + ldp * iofs, ldp, p + iofs + ldp * iofs, ldp,
                         cfg->IP);
            /* U update: p[kb, iofs+cblkn:m] -= L[kb, 0:iofs] * U[0:iofs, iofs+cblkn:m].
             * The reference reads B from column iofst = iofs+ib and writes C
             * from column iofs+cblkn (:256-282); they agree when nb % ib == 0,
             * the only case where its result is defined. */
            if (cblkm > cblkn)
                orc_gemm_sub(cfg->I, cblkn, cblkm - cblkn, iofs, p + iofs,
                             ldp, p + ldp * (iofs + cblkn), ldp,
                             p + iofs + ldp * (iofs + cblkn), ldp, cfg->IP);
        }
        /* panel in IF arithmetic on the IF-rounded values, factors stored
         * in OF (the inner working precision) */
        float *blk = p + iofs + ldp * iofs;
        round_block(cfg->IF, cblkm, cblkn, blk, ldp);
        if (cblkm > cblkn) round_block(cfg->IF, cblkn, cblkm - cblkn, blk + ldp * cblkn, ldp);
        for (int64_t j = 0; j < cblkn; ++j)
            for (int64_t i = 0; i < cblkn; ++i) dsnap[i + 32 * j] = blk[i + ldp * j];
        panel_f(cfg->IF, cblkm, cblkn, dsnap, 32, blk, ldp, blk, ldp);
        round_block(cfg->OF, cblkm, cblkn, blk, ldp);
        if (cblkm > cblkn) round_block(cfg->OF, cblkn, cblkm - cblkn, blk + ldp * cblkn, ldp);
    }
}

/* ------------------------------------------------------------------ */
/* left-looking driver: src/lu_nopiv_ll.hxx:55-582                      */
/* ------------------------------------------------------------------ */
int orc_lu_nopiv_ll(const orc_lu_cfg *cfg, int64_t m, int64_t n, int64_t nb,
                    float *a, int64_t lda)
{
    if (check_cfg(cfg)) return -1;
    if (m <= 0) return -2;
    if (n != m) return -3;
    if (nb <= 0) return -4;
    if (lda < m) return -6;
    if (cfg->O == ORC_FP8TC32 && nb % 32) return -4;
    const int64_t nc = (n - 1) / nb + 1;
    for (int64_t kk = 0; kk < nc; ++kk) {
        int64_t ofs = kk * nb;
        int64_t in = n - ofs < nb ? n - ofs : nb;
        int64_t updm = m - ofs;
        float *blk = a + ofs + lda * ofs;
        if (ofs > 0) {
            /* block column: A[ofs:m, ofs:ofs+in] -= L[ofs:m,0:ofs] U[0:ofs,ofs:ofs+in]
             * (:239-304), result in OP */
            orc_gemm_sub(cfg->O, updm, in, ofs, a + ofs, lda, a + lda * ofs,
                         lda, blk, lda, cfg->OP);
            /* block row: A[ofs:ofs+in, ofs+in:m] -= L[ofs:ofs+in,0:ofs] U[0:ofs,ofs+in:m]
             * (:311-410) */
            if (updm > in)
                orc_gemm_sub(cfg->O, in, updm - in, ofs, a + ofs, lda,
                             a + lda * (ofs + in), lda, blk + lda * in, lda,
                             cfg->OP);
        }
        /* the outer block is factored in OF (:451-489) */
        round_block(cfg->OF, updm, in, blk, lda);
        if (updm > in) round_block(cfg->OF, in, updm - in, blk + lda * in, lda);
        inner_ll(cfg, updm, in, blk, lda);
        /* factors stored back at the working precision (:498-519) */
        round_block(cfg->W, updm, in, blk, lda);
        if (updm > in) round_block(cfg->W, in, updm - in, blk + lda * in, lda);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* right-looking driver: src/lu_nopiv.hxx:53-208 and                    */
/* src/lu_nopiv_inner_rl.hxx:34-198 (everything in working precision)   */
/* ------------------------------------------------------------------ */
int orc_lu_nopiv_rl(const orc_lu_cfg *cfg, int64_t m, int64_t n, int64_t nb,
                    float *a, int64_t lda)
{
    if (check_cfg(cfg)) return -1;
    if (m <= 0) return -2;
    if (n != m) return -3;
    if (nb <= 0) return -4;
    if (lda < m) return -6;
    if (cfg->O == ORC_FP8TC32 && nb % 32) return -4;
    const int T = cfg->W, ib = (int)cfg->ib;
    const int64_t nc = (n - 1) / nb + 1;
    float dsnap[32 * 32];
    for (int64_t kk = 0; kk < nc; ++kk) {
        int64_t ofs = kk * nb, ofst = (kk + 1) * nb;
        int64_t in = n - ofs < nb ? n - ofs : nb;
        int64_t updm = m - ofs;
        float *p = a + ofs + lda * ofs;
        int64_t inc = (in - 1) / ib + 1;
        for (int64_t k = 0; k < inc; ++k) {
            int64_t iofs = k * ib, cblkm = updm - iofs;
            int64_t cblkn = in - iofs < ib ? in - iofs : ib;
            int64_t iofst = (k + 1) * ib;
            int64_t tblkm = updm - iofst, tblkn = in - iofst;
            float *b = p + iofs + lda * iofs;
            for (int64_t j = 0; j < cblkn; ++j)
                for (int64_t i = 0; i < cblkn; ++i) dsnap[i + 32 * j] = b[i + lda * j];
            panel_f(T, cblkm, cblkn, dsnap, 32, b, lda, b, lda);
            if (tblkn > 0) {
                /* rank-ib update of the outer block column (:139-163) */
                orc_gemm_sub(cfg->I, tblkm, tblkn, ib, p + iofst + lda * iofs,
                             lda, p + iofs + lda * iofst, lda,
                             p + iofst + lda * iofst, lda, T);
                /* and of the outer block row beyond the outer block (:167-193) */
                if (tblkm > tblkn)
                    orc_gemm_sub(cfg->I, tblkn, tblkm - tblkn, ib,
                                 p + iofst + lda * iofs, lda,
                                 p + iofs + lda * (ofst - ofs), lda,
                                 p + iofst + lda * (ofst - ofs), lda, T);
            }
        }
        /* Schur complement A[ofst:m, ofst:n] -= L[ofst:m, ofs:ofst] U[ofs:ofst, ofst:n]
         * (:170-195) */
        int64_t tm = m - ofst, tn = n - ofst;
        if (tn > 0)
            orc_gemm_sub(cfg->O, tm, tn, nb, a + ofst + lda * ofs, lda,
                         a + ofs + lda * ofst, lda, a + ofst + lda * ofst, lda,
                         T);
    }
    return 0;
}

static int typed_driver(int rl, const orc_lu_cfg *cfg, int64_t m, int64_t n,
                        int64_t nb, void *a, int64_t lda)
{
    if (cfg->W == ORC_F32)
        return rl ? orc_lu_nopiv_rl(cfg, m, n, nb, (float *)a, lda)
                  : orc_lu_nopiv_ll(cfg, m, n, nb, (float *)a, lda);
    if (cfg->W != ORC_F16) return -1;
    if (m <= 0 || lda < m) return -2;
    float *w = malloc(sizeof(float) * (size_t)lda * n);
    uint16_t *h = (uint16_t *)a;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < lda; ++i) w[i + lda * j] = orc_f16_to_f32(h[i + lda * j]);
    int st = rl ? orc_lu_nopiv_rl(cfg, m, n, nb, w, lda)
                : orc_lu_nopiv_ll(cfg, m, n, nb, w, lda);
    if (st == 0)
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) h[i + lda * j] = orc_f32_to_f16(w[i + lda * j]);
    free(w);
    return st;
}

int orc_lu_nopiv_ll_typed(const orc_lu_cfg *cfg, int64_t m, int64_t n,
                          int64_t nb, void *a, int64_t lda)
{
    return typed_driver(0, cfg, m, n, nb, a, lda);
}

int orc_lu_nopiv_rl_typed(const orc_lu_cfg *cfg, int64_t m, int64_t n,
                          int64_t nb, void *a, int64_t lda)
{
    return typed_driver(1, cfg, m, n, nb, a, lda);
}

/* ------------------------------------------------------------------ */
/* FP64 textbook Doolittle LU                                          */
/* ------------------------------------------------------------------ */
int orc_lu_f64(int64_t n, double *a, int64_t lda)
{
    for (int64_t k = 0; k < n; ++k) {
        double piv = a[k + lda * k];
        if (piv == 0.0 || !isfinite(piv)) return (int)(k + 1);
        double r = 1.0 / piv;
        for (int64_t i = k + 1; i < n; ++i) a[i + lda * k] *= r;
#pragma omp parallel for OMP_THREADS schedule(static) if (n - k > 256)
        for (int64_t j = k + 1; j < n; ++j) {
            double ukj = a[k + lda * j];
            for (int64_t i = k + 1; i < n; ++i) a[i + lda * j] -= a[i + lda * k] * ukj;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* solves                                                              */
/* ------------------------------------------------------------------ */
void orc_getrs_nopiv_f32(int64_t n, int64_t nrhs, const float *lu,
                         int64_t ldlu, float *b, int64_t ldb)
{
    for (int64_t r = 0; r < nrhs; ++r) {
        float *x = b + ldb * r;
        /* strsm Left/Lower/NoTrans/Unit, column oriented */
        for (int64_t k = 0; k < n; ++k) {
            float xk = x[k];
            if (xk != 0.0f)
                for (int64_t i = k + 1; i < n; ++i) x[i] -= xk * lu[i + ldlu * k];
        }
        /* strsm Left/Upper/NoTrans/NonUnit */
        for (int64_t k = n - 1; k >= 0; --k) {
            if (x[k] != 0.0f) {
                x[k] /= lu[k + ldlu * k];
                float xk = x[k];
                for (int64_t i = 0; i < k; ++i) x[i] -= xk * lu[i + ldlu * k];
            }
        }
    }
}

void orc_getrs_nopiv_f64(int64_t n, int64_t nrhs, const double *lu,
                         int64_t ldlu, double *b, int64_t ldb)
{
    for (int64_t r = 0; r < nrhs; ++r) {
        double *x = b + ldb * r;
        for (int64_t k = 0; k < n; ++k) {
            double xk = x[k];
            if (xk != 0.0)
                for (int64_t i = k + 1; i < n; ++i) x[i] -= xk * lu[i + ldlu * k];
        }
        for (int64_t k = n - 1; k >= 0; --k) {
            if (x[k] != 0.0) {
                x[k] /= lu[k + ldlu * k];
                double xk = x[k];
                for (int64_t i = 0; i < k; ++i) x[i] -= xk * lu[i + ldlu * k];
            }
        }
    }
}

void orc_getrs_nopiv_trans_f64(int64_t n, int64_t nrhs, const double *lu,
                               int64_t ldlu, double *b, int64_t ldb)
{
    for (int64_t r = 0; r < nrhs; ++r) {
        double *x = b + ldb * r;
        /* U^T y = b: row-oriented forward substitution on U's columns */
        for (int64_t k = 0; k < n; ++k) {
            double s = x[k];
            for (int64_t i = 0; i < k; ++i) s -= lu[i + ldlu * k] * x[i];
            x[k] = s / lu[k + ldlu * k];
        }
        /* L^T x = y, unit diagonal */
        for (int64_t k = n - 1; k >= 0; --k) {
            double s = x[k];
            for (int64_t i = k + 1; i < n; ++i) s -= lu[i + ldlu * k] * x[i];
            x[k] = s;
        }
    }
}

/* ------------------------------------------------------------------ */
/* error measures: tests/errors.hxx                                     */
/* ------------------------------------------------------------------ */
/* tests/errors.hxx:186-249: max over rhs of
 * ||b - A x||_inf / (||A||_1 ||x||_inf + ||b||_inf), NaN propagating. */
double orc_unsym_backward_error(int64_t m, int64_t n, const float *a,
                                int64_t lda, const double *rhs, int64_t nrhs,
                                int64_t ldrhs, const double *soln,
                                int64_t ldsoln)
{
    double anorm = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        double s = 0.0;
        for (int64_t i = 0; i < m; ++i) s += fabs((double)a[j * lda + i]);
        anorm = s > anorm ? s : anorm;
    }
    double worst = 0.0;
    double *resid = malloc(sizeof(double) * m);
    for (int64_t r = 0; r < nrhs; ++r) {
        for (int64_t i = 0; i < m; ++i) resid[i] = rhs[r * ldrhs + i];
        for (int64_t j = 0; j < n; ++j) {
            double xj = soln[r * ldsoln + j];
            for (int64_t i = 0; i < m; ++i) resid[i] -= (double)a[j * lda + i] * xj;
        }
        double rhsnorm = 0.0, residnorm = 0.0, solnnorm = 0.0;
        for (int64_t i = 0; i < m; ++i) {
            rhsnorm = fmax(rhsnorm, fabs(rhs[r * ldrhs + i]));
            residnorm = fmax(residnorm, fabs(resid[i]));
            if (isnan(resid[i])) residnorm = resid[i];
        }
        for (int64_t i = 0; i < n; ++i) solnnorm = fmax(solnnorm, fabs(soln[i + r * ldsoln]));
        double e = residnorm / (anorm * solnnorm + rhsnorm);
        worst = e > worst ? e : worst;
        if (isnan(residnorm)) worst = residnorm;
    }
    free(resid);
    return worst;
}

/* tests/errors.hxx:22-182 (one right-hand side): max_i |b - A x|_i /
 * (|A||x| + |L||U||x|)_i; cwerr_ax receives max_i |b - A x|_i / (|A||x|)_i */
double orc_unsym_compwise_error(int64_t m, int64_t n, const float *a,
                                int64_t lda, const double *rhs,
                                const double *soln, const float *lu,
                                int64_t ldlu, double *cwerr_ax)
{
    double *resid = malloc(sizeof(double) * m), *ax = calloc(m, sizeof(double));
    double *lux = malloc(sizeof(double) * n), *t = malloc(sizeof(double) * n);
    for (int64_t i = 0; i < m; ++i) resid[i] = -rhs[i]; /* A x - b */
    for (int64_t j = 0; j < n; ++j) {
        double xj = soln[j], axj = fabs(soln[j]);
        for (int64_t i = 0; i < m; ++i) {
            resid[i] += (double)a[j * lda + i] * xj;
            ax[i] += fabs((double)a[j * lda + i]) * axj;
        }
    }
    /* |U| |x| (upper incl. diagonal) then |L| (unit lower) on top */
    for (int64_t i = 0; i < n; ++i) {
        double s = 0.0;
        for (int64_t j = i; j < n; ++j) s += fabs((double)lu[j * ldlu + i]) * fabs(soln[j]);
        t[i] = s;
    }
    for (int64_t i = 0; i < n; ++i) {
        double s = t[i];
        for (int64_t j = 0; j < i; ++j) s += fabs((double)lu[j * ldlu + i]) * t[j];
        lux[i] = s;
    }
    double cw = 0.0, cwa = 0.0;
    for (int64_t i = 0; i < m; ++i) {
        double r = fabs(resid[i]);
        double d1 = 1.0 / ax[i];
        if (d1 > 0.0) cwa = fmax(cwa, r * d1);
        double d2 = 1.0 / (ax[i] + lux[i]);
        if (d2 > 0.0) r = r * d2;
        cw = fmax(cw, fabs(r));
    }
    if (cwerr_ax) *cwerr_ax = cwa;
    free(resid); free(ax); free(lux); free(t);
    return cw;
}

/* tests/errors.hxx:251-346: max_ij |A - L U|_ij / (|L||U|)_ij */
double orc_factor_backward_error(int64_t n, const float *a, int64_t lda,
                                 const float *lu, int64_t ldlu)
{
    double err = 0.0;
#pragma omp parallel for OMP_THREADS schedule(dynamic) reduction(max : err)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) {
            int64_t kmax = i < j ? i : j; /* sum over k <= min(i,j) */
            double s = 0.0, d = 0.0;
            for (int64_t k = 0; k <= kmax; ++k) {
                double l = (k == i) ? 1.0 : (double)lu[k * ldlu + i];
                double u = (double)lu[j * ldlu + k];
                s += l * u;
                d += fabs(l) * fabs(u);
            }
            double e = fabs((double)a[j * lda + i] - s);
            if (d > 0.0) e /= d;
            if (e > err) err = e;
        }
    return err;
}

void orc_residual_f64(int64_t n, int64_t nrhs, const float *a, int64_t lda,
                      const double *x, int64_t ldx, const double *b,
                      int64_t ldb, double *r, int64_t ldr)
{
#pragma omp parallel for OMP_THREADS schedule(static)
    for (int64_t i = 0; i < n; ++i)
        for (int64_t c = 0; c < nrhs; ++c) {
            double s = b[i + ldb * c];
            for (int64_t j = 0; j < n; ++j) s -= (double)a[i + lda * j] * x[j + ldx * c];
            r[i + ldr * c] = s;
        }
}

int orc_ir_solve(int64_t n, int64_t nrhs, const float *a, int64_t lda,
                 const float *lu, int64_t ldlu, const double *b, int64_t ldb,
                 double *x, int64_t ldx, double tol, int maxit, double *hist)
{
    float *w = malloc(sizeof(float) * n * nrhs);
    double *r = malloc(sizeof(double) * n * nrhs);
    for (int64_t c = 0; c < nrhs; ++c)
        for (int64_t i = 0; i < n; ++i) w[i + n * c] = (float)b[i + ldb * c];
    orc_getrs_nopiv_f32(n, nrhs, lu, ldlu, w, n);
    for (int64_t c = 0; c < nrhs; ++c)
        for (int64_t i = 0; i < n; ++i) x[i + ldx * c] = (double)w[i + n * c];
    int it = 0, done = -1;
    for (;; ++it) {
        orc_residual_f64(n, nrhs, a, lda, x, ldx, b, ldb, r, n);
        double e = orc_unsym_backward_error(n, n, a, lda, b, nrhs, ldb, x, ldx);
        if (hist) hist[it] = e;
        if (e <= tol) { done = it; break; }
        if (it >= maxit) break;
        for (int64_t c = 0; c < nrhs; ++c)
            for (int64_t i = 0; i < n; ++i) w[i + n * c] = (float)r[i + n * c];
        orc_getrs_nopiv_f32(n, nrhs, lu, ldlu, w, n);
        for (int64_t c = 0; c < nrhs; ++c)
            for (int64_t i = 0; i < n; ++i) x[i + ldx * c] += (double)w[i + n * c];
    }
    free(w);
    free(r);
    return done;
}

/* ------------------------------------------------------------------ */
/* equilibration (tests/testing_lu_gpu.hxx:288-311 calls LAPACK's      */
/* sgeequb then slaqge through src/wrappers.cxx:209-250; LAPACK is a    */
/* third-party dependency, restated here from its published routines)  */
/* ------------------------------------------------------------------ */

/* RADIX**INT(LOG(x)/LOG(RADIX)) for x > 0, radix 2, with the truncation
 * toward zero of log2(x) taken exactly from the exponent */
static float pow2_trunc_log2(float x)
{
    int e;
    float f = frexpf(x, &e); /* x = f 2^e, f in [0.5, 1) */
    int k = (x >= 1.0f || f == 0.5f) ? e - 1 : e;
    return ldexpf(1.0f, k);
}

int orc_geequb(int64_t m, int64_t n, const float *a, int64_t lda, float *r, float *c,
               float *rowcnd, float *colcnd, float *amax)
{
    if (m == 0 || n == 0) {
        *rowcnd = 1.0f; *colcnd = 1.0f; *amax = 0.0f;
        return 0;
    }
    const float smlnum = FLT_MIN, bignum = 1.0f / FLT_MIN;
    for (int64_t i = 0; i < m; ++i) r[i] = 0.0f;
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) r[i] = fmaxf(r[i], fabsf(a[j * lda + i]));
    for (int64_t i = 0; i < m; ++i)
        if (r[i] > 0.0f) r[i] = pow2_trunc_log2(r[i]);
    float rcmin = bignum, rcmax = 0.0f;
    for (int64_t i = 0; i < m; ++i) { rcmax = fmaxf(rcmax, r[i]); rcmin = fminf(rcmin, r[i]); }
    *amax = rcmax;
    if (rcmin == 0.0f) {
        for (int64_t i = 0; i < m; ++i)
            if (r[i] == 0.0f) return (int)(i + 1);
    }
    for (int64_t i = 0; i < m; ++i) r[i] = 1.0f / fminf(fmaxf(r[i], smlnum), bignum);
    *rowcnd = fmaxf(rcmin, smlnum) / fminf(rcmax, bignum);
    for (int64_t j = 0; j < n; ++j) {
        float cj = 0.0f;
        for (int64_t i = 0; i < m; ++i) cj = fmaxf(cj, fabsf(a[j * lda + i]) * r[i]);
        c[j] = cj > 0.0f ? pow2_trunc_log2(cj) : 0.0f;
    }
    rcmin = bignum; rcmax = 0.0f;
    for (int64_t j = 0; j < n; ++j) { rcmin = fminf(rcmin, c[j]); rcmax = fmaxf(rcmax, c[j]); }
    if (rcmin == 0.0f) {
        for (int64_t j = 0; j < n; ++j)
            if (c[j] == 0.0f) return (int)(m + j + 1);
    }
    for (int64_t j = 0; j < n; ++j) c[j] = 1.0f / fminf(fmaxf(c[j], smlnum), bignum);
    *colcnd = fmaxf(rcmin, smlnum) / fminf(rcmax, bignum);
    return 0;
}

char orc_laqge(int64_t m, int64_t n, float *a, int64_t lda, const float *r, const float *c,
               float rowcnd, float colcnd, float amax)
{
    const float thresh = 0.1f;
    /* SMALL = SLAMCH('S') / SLAMCH('P') = 2^-126 / 2^-23, LARGE = 1 / SMALL */
    const float small = ldexpf(1.0f, -103), large = 1.0f / small;
    if (m <= 0 || n <= 0) return 'N';
    if (rowcnd >= thresh && amax >= small && amax <= large) {
        if (colcnd >= thresh) return 'N';
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) a[j * lda + i] = c[j] * a[j * lda + i];
        return 'C';
    }
    if (colcnd >= thresh) {
        for (int64_t j = 0; j < n; ++j)
            for (int64_t i = 0; i < m; ++i) a[j * lda + i] = r[i] * a[j * lda + i];
        return 'R';
    }
    for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < m; ++i) a[j * lda + i] = (c[j] * r[i]) * a[j * lda + i];
    return 'B';
}
