/*
 * remifa_oracle.h -- CPU restatement of flipflapflop/remifa's no-pivot LU
 * hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 * The product (remifa_amd/, include/) never links or calls it.
 *
 * Every routine names the reference file:line it restates. Matrices are
 * column-major with a leading dimension, like the reference. Reduced
 * precision values (fp16, bf16, e4m3) are carried in float arrays whose
 * values are exactly representable in that precision; the typed wrappers
 * (orc_*_typed) convert from/to the packed 16-bit storage.
 */
#ifndef REMIFA_ORACLE_H
#define REMIFA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* storage / arithmetic precisions */
enum { ORC_F32 = 0, ORC_F16 = 1, ORC_F64 = 2, ORC_BF16 = 3, ORC_E4M3 = 4 };

/* update compute types: src/utils.hxx:18-24 (FP32, FP16, TC32, TC16) plus
 * the two MI355X additions (bf16 operands / fp32 accumulate, and e4m3
 * operands with one power-of-two scale per 32-long K block / fp32 acc). */
enum { ORC_FP32 = 0, ORC_FP16 = 1, ORC_TC32 = 2, ORC_TC16 = 3,
       ORC_BF16TC32 = 4, ORC_FP8TC32 = 5 };

/* ---- seeded generators (libstdc++ minstd_rand0 + uniform_real<float>) ---- */
uint64_t orc_minstd_seed(uint64_t seed);
uint64_t orc_minstd_jump(uint64_t state, uint64_t k);
float orc_minstd_uniform(uint64_t state);   /* value for one engine state */
void orc_minstd_stream(uint64_t seed, int64_t count, float *out);
void orc_gen_diagdom(int64_t m, int64_t n, float *a, int64_t lda);
void orc_gen_diagdom_alpha(int64_t n, uint64_t seed, double alpha,
                           float *a, int64_t lda);
void orc_gen_uniform(int64_t m, int64_t n, uint64_t seed, float *b,
                     int64_t ldb);
/* tests/common.hxx:298-319 gen_randmat (column-major draws, uniform(-1,1)) */
void orc_gen_randmat(int64_t m, int64_t n, uint64_t seed, float *a, int64_t lda);
/* tests/common.hxx:391-455 gen_arrowhead (k = 256, entries uniform(lo,hi),
 * a_jj = 2k) */
void orc_gen_arrowhead(int64_t m, int64_t n, float lo, float hi, float *a, int64_t lda);

/* ---- scalar conversions (IEEE round to nearest even) ---- */
uint16_t orc_f32_to_f16(float x);
float orc_f16_to_f32(uint16_t h);
uint16_t orc_f64_to_f16(double x);
uint16_t orc_f32_to_bf16(float x);
float orc_bf16_to_f32(uint16_t h);
uint8_t orc_f32_to_e4m3(float x);
float orc_e4m3_to_f32(uint8_t b);
float orc_round(int prec, float x);

/* src/convert.cu:324-360: cast an m x n block between precisions */
int orc_convert(int src, int dst, int64_t m, int64_t n, const void *a,
                int64_t lda, void *b, int64_t ldb);

/* ---- panel: src/lu_nopiv.cu:330-400 ---- */
int orc_lu_nopiv_panel(int prec, int64_t ib, int64_t m, int64_t n,
                       const void *d, int64_t ldd, void *l, int64_t ldl,
                       void *u, int64_t ldu, int *stat);

/* ---- GEMM update semantics: src/gemm.cu (C -= A*B, alpha=-1, beta=1) ---- */
/* e4m3 operand format of the FP8TC32 update (a MI355X addition, no
 * reference counterpart): for row r (rows = 1, a is m x k) or column r
 * (rows = 0, a is k x m) and each 32-long k block b from k = 0,
 * sexp[r * ceil(k/32) + b] = floor(log2 amax) - 8 (0 for an all-zero block)
 * and codes[r * k + kk] = e4m3(a / 2^sexp), nearest even, saturating. */
void orc_quantise_e4m3_codes(int64_t m, int64_t k, const float *a, int64_t lda, uint8_t *codes,
                             int32_t *sexp, int rows);
/* 0: k ascending (default), 1: k descending for fp32-accumulated sums */
void orc_set_sum_order(int order);
void orc_gemm_sub(int ctype, int64_t m, int64_t n, int64_t k,
                  const float *a, int64_t lda, const float *b, int64_t ldb,
                  float *c, int64_t ldc, int cprec);

/* ---- drivers ---- */
typedef struct {
    int W;        /* storage precision of A (ORC_F32 / ORC_F16) */
    int I, O;     /* inner / outer update compute types */
    int OP, IP;   /* outer / inner panel (accumulation buffer) precisions */
    int OF, IF;   /* outer / inner factor precisions */
    int64_t ib;   /* inner block size (reference helpers: 8) */
} orc_lu_cfg;

/* src/lu_nopiv_ll.hxx:55-582 + src/lu_nopiv_inner_ll.hxx:37-372.
 * a holds W-valued floats; m == n. Returns 0 or a negative argument index. */
int orc_lu_nopiv_ll(const orc_lu_cfg *cfg, int64_t m, int64_t n, int64_t nb,
                    float *a, int64_t lda);
/* src/lu_nopiv.hxx:53-208 + src/lu_nopiv_inner_rl.hxx:34-198 */
int orc_lu_nopiv_rl(const orc_lu_cfg *cfg, int64_t m, int64_t n, int64_t nb,
                    float *a, int64_t lda);
/* typed wrappers: a is float* (W=F32) or uint16_t* fp16 bits (W=F16) */
int orc_lu_nopiv_ll_typed(const orc_lu_cfg *cfg, int64_t m, int64_t n,
                          int64_t nb, void *a, int64_t lda);
int orc_lu_nopiv_rl_typed(const orc_lu_cfg *cfg, int64_t m, int64_t n,
                          int64_t nb, void *a, int64_t lda);

/* textbook FP64 Doolittle LU without pivoting (the FP64 oracle) */
int orc_lu_f64(int64_t n, double *a, int64_t lda);

/* ---- solves ---- */
/* src/wrappers.cxx:476-482 (sgetrs with ipiv = 1..n): unit-lower then
 * upper substitution, column oriented, in fp32 */
void orc_getrs_nopiv_f32(int64_t n, int64_t nrhs, const float *lu,
                         int64_t ldlu, float *b, int64_t ldb);
void orc_getrs_nopiv_f64(int64_t n, int64_t nrhs, const double *lu,
                         int64_t ldlu, double *b, int64_t ldb);
/* transposed solve A^T x = b with A = L U: U^T y = b, then L^T x = y */
void orc_getrs_nopiv_trans_f64(int64_t n, int64_t nrhs, const double *lu,
                               int64_t ldlu, double *b, int64_t ldb);

/* ---- error measures: tests/errors.hxx ---- */
double orc_unsym_backward_error(int64_t m, int64_t n, const float *a,
                                int64_t lda, const double *rhs, int64_t nrhs,
                                int64_t ldrhs, const double *soln,
                                int64_t ldsoln);
double orc_unsym_compwise_error(int64_t m, int64_t n, const float *a,
                                int64_t lda, const double *rhs,
                                const double *soln, const float *lu,
                                int64_t ldlu, double *cwerr_ax);
double orc_factor_backward_error(int64_t n, const float *a, int64_t lda,
                                 const float *lu, int64_t ldlu);

/* residual r = b - A x in FP64 (A fp32) */
void orc_residual_f64(int64_t n, int64_t nrhs, const float *a, int64_t lda,
                      const double *x, int64_t ldx, const double *b,
                      int64_t ldb, double *r, int64_t ldr);

/* iterative refinement with the FP64 residual: x0 = LU\b in fp32 then
 * x += LU\(b - A x) until nwerr <= tol. Returns iterations (or -1). */
int orc_ir_solve(int64_t n, int64_t nrhs, const float *a, int64_t lda,
                 const float *lu, int64_t ldlu, const double *b, int64_t ldb,
                 double *x, int64_t ldx, double tol, int maxit,
                 double *hist);

/* ---- equilibration: LAPACK sgeequb / slaqge (third party), called by the
 * reference harness at tests/testing_lu_gpu.hxx:288-311 ---- */
/* power-of-two row / column scalings; returns LAPACK's INFO (0, i for a
 * zero row i, m + j for a zero column j, 1-based) */
int orc_geequb(int64_t m, int64_t n, const float *a, int64_t lda, float *r, float *c,
               float *rowcnd, float *colcnd, float *amax);
/* scale a in place as slaqge does; returns EQUED ('N', 'R', 'C', 'B') */
char orc_laqge(int64_t m, int64_t n, float *a, int64_t lda, const float *r, const float *c,
               float rowcnd, float colcnd, float amax);

/* threads used by the OpenMP loops (0 = OpenMP default) */
void orc_set_threads(int nthreads);
int orc_get_threads(void);

#ifdef __cplusplus
}
#endif

#endif
