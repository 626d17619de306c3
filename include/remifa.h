/*
 * remifa.h -- C ABI of libremifa_hip, the MI355X (gfx950) no-pivot LU library.
 *
 * Plain C: pointers, sizes and int status codes; no C++ or torch types cross
 * this boundary. Matrices are column-major with a leading dimension, exactly
 * as in the reference (flipflapflop/remifa). All device pointers are HIP
 * device memory; every call enqueues work on the handle's stream and, like
 * the reference drivers, the LU drivers synchronise that stream before
 * returning (src/lu_nopiv_ll.hxx:567-569).
 *
 * Status codes: REMIFA_SUCCESS (0); -i when argument i (1-based, the handle
 * counts as argument 1) is invalid; REMIFA_ERR_* (> 0) for runtime errors.
 * remifa_last_error() returns a message for the last failure of the thread.
 */
#ifndef REMIFA_H
#define REMIFA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define REMIFA_SUCCESS 0
#define REMIFA_ERR_HIP 1000         /* a HIP runtime call failed */
#define REMIFA_ERR_ALLOC 1001       /* device allocation failed */
#define REMIFA_ERR_UNSUPPORTED 1002 /* valid but unsupported combination */
#define REMIFA_ERR_WORKSPACE 1003   /* user workspace too small */

typedef struct remifa_handle_st *remifa_handle_t;

/* element types */
typedef enum {
  REMIFA_R_32F = 0,
  REMIFA_R_16F = 1,
  REMIFA_R_64F = 2,
  REMIFA_R_16BF = 3,
  REMIFA_R_8F_E4M3 = 4
} remifa_dtype_t;

/* update compute types. The first four are the reference's
 * remifa::compute_type (src/utils.hxx:18-24):
 *   FP32: fp32 operands and accumulation (cublasSgemm, src/gemm.cu:31)
 *   FP16: fp16 operands and accumulation (cublasHgemm, src/gemm.cu:96)
 *   TC32: fp16 operands, fp32 accumulation (GemmEx 32F, src/gemm.cu:240,276)
 *   TC16: fp16 operands, fp16 accumulation (GemmEx 16F, src/gemm.cu:162)
 * The MI355X additions feed the same fp32-accumulating matrix cores with
 *   BF16_TC32: bf16 operands (rounded from the stored factors)
 *   FP8_TC32:  OCP e4m3 operands, one power-of-two scale per 32-long K block
 *              of every row of A and column of B (outer updates only). */
typedef enum {
  REMIFA_FP32 = 0,
  REMIFA_FP16 = 1,
  REMIFA_TC32 = 2,
  REMIFA_TC16 = 3,
  REMIFA_BF16_TC32 = 4,
  REMIFA_FP8_TC32 = 5
} remifa_compute_t;

typedef enum { REMIFA_LL = 0, REMIFA_RL = 1 } remifa_algo_t;

/* Precision tuple of the reference driver templates (src/lu_nopiv_ll.hxx:41-54):
 * W storage, I/O inner/outer update compute types, OP/IP outer/inner panel
 * (update accumulation) precisions, OF/IF outer/inner factor precisions, ib
 * the inner (diagonal tile) block size. Precisions are remifa_dtype_t
 * (REMIFA_R_32F or REMIFA_R_16F). */
typedef struct {
  int W;
  int I, O;
  int OP, IP;
  int OF, IF;
  int64_t ib;
} remifa_lu_config_t;

/* ------------------------------------------------------------ handle */
/* Replaces the cublasHandle_t the reference drivers take
 * (src/lu_nopiv_ll.hxx:56, :139 cublasGetStream). stream is a hipStream_t
 * (NULL = the default stream). */
int remifa_create(remifa_handle_t *handle, void *stream);
int remifa_destroy(remifa_handle_t handle);
int remifa_set_stream(remifa_handle_t handle, void *stream);
int remifa_get_stream(remifa_handle_t handle, void **stream);
/* Optional caller-owned workspace (device memory). When set, the library
 * never allocates; calls needing more return REMIFA_ERR_WORKSPACE. */
int remifa_set_workspace(remifa_handle_t handle, void *d_ws, size_t bytes);
/* Bytes of workspace remifa_lu_nopiv needs for this problem. */
int remifa_lu_workspace_size(remifa_handle_t handle, int algo, const remifa_lu_config_t *cfg,
                             int64_t m, int64_t n, int64_t nb, size_t *bytes);
/* Tuning knobs (0 = automatic): width of the fused panel kernels and the
 * split-K factor of the left-looking updates. */
int remifa_set_tuning(remifa_handle_t handle, int panel_width, int split_k);
/* Block kernels of the LL inner factorization (fp16 storage, fp32 buffers,
 * TC32, ib = 32): one_launch = 1 (the default) runs the diagonal tiles and
 * the rows / columns below / right of them in one launch, the latter slice
 * by slice as the diagonal role publishes each slice; 0 runs them as two
 * launches (the diagonal block, then the rest), one after the other.
 * units_per_wave (0 = automatic, 3): units per helper wavefront of the
 * one-launch form, which sets its grid. wait_polls (0 = automatic, 2^21):
 * polls of a published counter before a wait of the one-launch form gives
 * up; a wait that gives up makes remifa_lu_nopiv return REMIFA_ERR_HIP (a
 * tiny value forces that path in tests). Results are the same bit for bit
 * either way. No reference counterpart (the reference's ib loop is
 * sequential launches, src/lu_nopiv_inner_ll.hxx:197-370). */
int remifa_set_block_kernels(remifa_handle_t handle, int one_launch, int units_per_wave,
                             int64_t wait_polls);
/* Force the output tile of the update GEMMs (-1 = chosen per shape):
 * 0: 256x32, 1: 32x256, 2: 256x64, 3: 64x256, 4: 256x256, 5: 128x128,
 * 6: 256x128, 7: 128x256 (fp32 operands only take 0, 1, 5). Tile 4 runs
 * the LDS-DMA kernel for fp16 operands or bf16 storage when M and N are
 * multiples of 256, K of 32 and the operands 16-byte aligned (else 5); a
 * rule's tile 4 also falls back to 5 below K = 2048. Used by the autotuner. */
int remifa_set_gemm_tile(remifa_handle_t handle, int tile);
/* Autotuned GEMM table: count rules of 7 int64 each, {operand kind (0 fp16,
 * 1 bf16, 2 fp32; -1 any), m_lo, m_hi, n_lo, n_hi, tile (-1 = heuristic),
 * split-K (0 = heuristic)}; the first rule whose ranges (inclusive) hold the
 * GEMM's M and N decides. Written by remifa_amd/autotune.py. No reference
 * counterpart. */
int remifa_set_gemm_rules(remifa_handle_t handle, int count, const int64_t *rules);
/* Left-looking lookahead (1 = on, the default; 0 = off). With fp16 storage,
 * fp32 inner buffers and TC32 / BF16_TC32 outer updates (FP8_TC32 when nb
 * is a multiple of 128), the update of outer block
 * k+1 by the blocks before k runs on a second stream of the handle while
 * block k is factored; block k's own contribution follows on the handle's
 * stream. Both streams are joined before remifa_lu_nopiv returns.
 * reserve_cus (a multiple of 8 in [0, 64]; default 0): CUs masked off the
 * second stream, reserve_cus / 8 in each of the 8 XCDs, so that the
 * factorization of block k always finds free CUs (on MI355X any mask made
 * the n = 32768 factorization slower, 16 CUs: 1.6x). The mask is built for
 * the 256-CU MI355X layout: a nonzero reserve on a device with another CU
 * count is rejected (-3). A rejected call leaves both settings unchanged.
 * No reference counterpart (src/lu_nopiv_ll.hxx:208-562 runs the block
 * updates and factorizations in sequence on one stream). */
int remifa_set_lookahead(remifa_handle_t handle, int on, int reserve_cus);
/* The current lookahead settings (either pointer may be null). */
int remifa_get_lookahead(remifa_handle_t handle, int *on, int *reserve_cus);
const char *remifa_last_error(void);

/* Per-kernel timing with hipEvents on the handle's stream. While enabled,
 * every launch of the library is bracketed by two events. class: 0 update
 * GEMM, 1 panel, 2 convert, 3 triangular solve, 4 FP64 residual, 5 other. */
int remifa_timing_enable(remifa_handle_t handle, int enable);
/* Sums over the recorded launches of `cls`: total ms, launches and work
 * (flops for GEMM/TRSM/residual, bytes for panel/convert). Synchronises. */
int remifa_timing_query(remifa_handle_t handle, int cls, double *ms, int64_t *launches,
                        double *work);
int remifa_timing_reset(remifa_handle_t handle);
/* Every recorded launch (up to max): class, kernel variant (for GEMMs:
 * operand kind << 8 | fp16-output << 4 | tile shape, see remifa_amd/csrc/
 * gemm.hip), elapsed ms and work. *count receives the number recorded. */
int remifa_timing_records(remifa_handle_t handle, int64_t max, int *cls, int *variant, double *ms,
                          double *work, int64_t *count);

/* ------------------------------------------------------------ drivers */
/* Generic no-pivot LU of the m x n (m == n) column-major matrix d_a.
 * algo REMIFA_LL: lu_nopiv_ll<W,I,O,ib,OP,IP,OF,IF> (src/lu_nopiv_ll.hxx:55);
 * algo REMIFA_RL: lu_nopiv_rl<W,I,O,ib> (src/lu_nopiv.hxx:53; OP..IF = W).
 * d_a_cpy/ld_a_cpy: fp16 copy of the factors the reference requires when
 * W = fp32 and an update is TC32 (src/lu_nopiv_ll.hxx:132-135); pass NULL to
 * let the library use workspace instead. d_info (device int) receives 0, or
 * k > 0 when u_kk of column k (1-based) is zero or not finite. Returns with
 * the handle's stream synchronised, as the reference drivers do; while that
 * stream is being captured into a HIP graph the call only records its
 * kernels on it (one stream, no synchronisation or readback; size the
 * workspace with one eager call first). */
int remifa_lu_nopiv(remifa_handle_t handle, int algo, const remifa_lu_config_t *cfg, int64_t m,
                    int64_t n, int64_t nb, void *d_a, int64_t ldda, int *d_info, void *d_a_cpy,
                    int64_t ld_a_cpy);

/* The 12 cuBLAS-backed helpers of src/lu_nopiv_ll.cxx (ib = 8). Same names
 * (with the remifa_ prefix), arguments and precision tuples. */
/* src/lu_nopiv_ll.cxx:14   W=f32 I=O=FP32 */
int remifa_lu_nopiv_ll_f32(remifa_handle_t h, int m, int n, int nb, float *d_a, int ldda,
                           int *d_info);
/* src/lu_nopiv_ll.cxx:74   W=f32 I=FP32 O=TC32 */
int remifa_lu_nopiv_ll_f32_fp32xtc32(remifa_handle_t h, int m, int n, int nb, float *d_a,
                                     int ldda, int *d_info, void *d_a_f16, int ld_a_f16);
/* src/lu_nopiv_ll.cxx:131  W=f32 I=O=TC32 */
int remifa_lu_nopiv_ll_f32_tc32xtc32(remifa_handle_t h, int m, int n, int nb, float *d_a,
                                     int ldda, int *d_info, void *d_a_f16, int ld_a_f16);
/* src/lu_nopiv_ll.cxx:191  W=f16 I=O=FP16 */
int remifa_lu_nopiv_ll_f16(remifa_handle_t h, int m, int n, int nb, void *d_a, int ldda,
                           int *d_info);
/* src/lu_nopiv_ll.cxx:246  W=f16 I=O=TC16 */
int remifa_lu_nopiv_ll_f16_tc16xtc16(remifa_handle_t h, int m, int n, int nb, void *d_a,
                                     int ldda, int *d_info);
/* src/lu_nopiv_ll.cxx:282  W=f16 I=O=TC32 */
int remifa_lu_nopiv_ll_f16_tc32xtc32(remifa_handle_t h, int m, int n, int nb, void *d_a,
                                     int ldda, int *d_info);
/* src/lu_nopiv_ll.cxx:338  W=f16 I=FP16 O=TC32 */
int remifa_lu_nopiv_ll_f16_f16xtc32(remifa_handle_t h, int m, int n, int nb, void *d_a,
                                    int ldda, int *d_info);
/* src/lu_nopiv_ll.cxx:393  W=f16 I=FP16 O=TC32 OP=f32 */
int remifa_lu_nopiv_ll_f16_f16xtc32_fp32(remifa_handle_t h, int m, int n, int nb, void *d_a,
                                         int ldda, int *d_info);
/* src/lu_nopiv_ll.cxx:448  W=f16 I=FP32 O=TC32 OP=IP=OF=IF=f32 */
int remifa_lu_nopiv_ll_f16_f32xt32(remifa_handle_t h, int m, int n, int nb, void *d_a, int ldda,
                                   int *d_info);
/* src/lu_nopiv_ll.cxx:507  W=f16 I=FP16 O=TC16 */
int remifa_lu_nopiv_ll_f16_f16xtc16(remifa_handle_t h, int m, int n, int nb, void *d_a,
                                    int ldda, int *d_info);
/* src/lu_nopiv_ll.cxx:563  W=f16 I=O=TC32 OP=OF=f16 IP=IF=f32 */
int remifa_lu_nopiv_ll_f16_tc32xtc32_fp16xfp32_fp16xfp32(remifa_handle_t h, int m, int n, int nb,
                                                         void *d_a, int ldda, int *d_info);
/* src/lu_nopiv_ll.cxx:618  W=f16 I=O=TC32 OP=IP=OF=IF=f32 */
int remifa_lu_nopiv_ll_f16_tc32xtc32_fp32xfp32_fp32xfp32(remifa_handle_t h, int m, int n, int nb,
                                                         void *d_a, int ldda, int *d_info);

/* ------------------------------------------------------------ kernels */
/* src/lu_nopiv.cuh:12-20 / src/lu_nopiv.cu:364-400 lu_nopiv_panel<T,ib>:
 * factor the n x n diagonal tile d (n <= ib <= 32), write it to l, scale and
 * update the rows n..m-1 of the block column l (L21 = A21 U11^-1) and solve
 * the columns n..m-1 of the block row u (U12 = L11^-1 A12). t is
 * REMIFA_R_32F, REMIFA_R_16F or REMIFA_R_64F. stat[0] = n on return, as in
 * the reference (:173). */
int remifa_lu_nopiv_panel(remifa_handle_t h, int t, int ib, int m, int n, const void *d, int ldd,
                          void *l, int ldl, void *u, int ldu, int *stat);

/* src/convert.cuh:17-18 / src/convert.cu:324-360 convert<TA,TAO>: cast an
 * m x n block, round to nearest even when narrowing. Any pair of
 * remifa_dtype_t; src == dst copies (the reference's same-type convert is a
 * no-op). */
int remifa_convert(remifa_handle_t h, int src_t, int dst_t, int64_t m, int64_t n, const void *a,
                   int64_t lda, void *aout, int64_t ldaout);

/* src/gemm.cuh:24-40 gemmex<Impl,Op,AccType,InType,OutType>: C = alpha A B +
 * beta C (no transposes). ctype is a remifa_compute_t, in_t the storage type
 * of A and B (f32 for FP32; f16 for FP16/TC16/TC32/BF16_TC32; f16 or f32
 * for FP8_TC32, quantized per 32-long k block as remifa_quantize_fp8 does,
 * in the workspace), out_t the storage type of C (f32 or f16; f16 for
 * FP16/TC16). */
int remifa_gemmex(remifa_handle_t h, int ctype, int in_t, int out_t, int m, int n, int k,
                  float alpha, const void *a, int lda, const void *b, int ldb, float beta, void *c,
                  int ldc);

/* FP8 update operands (no reference counterpart: the MI355X FP8_TC32 update).
 * Quantize the m x k block a (column-major, ld lda, f32 or f16) by rows
 * (rows != 0: one scale per row and 32-long k block, the A operand) or its
 * k x m transpose view by columns (rows == 0: a is k x m, one scale per
 * column and 32-long k block, the B operand): codes q[r * ldq + kk] (OCP
 * e4m3, nearest even, saturating), scale bytes s[r * lds + kk / 32] meaning
 * 2^(s - 127), the scale being 2^(floor(log2 amax) - 8) of the block (2^0
 * for an all-zero block). Needs q 16-byte aligned, ldq a multiple of 16 and
 * >= k rounded up to 32, lds >= ceil(k / 32). */
int remifa_quantize_fp8(remifa_handle_t h, int in_t, int rows, int64_t m, int64_t k,
                        const void *a, int64_t lda, uint8_t *q, int64_t ldq, uint8_t *s,
                        int64_t lds);
/* C = alpha A B + beta C (C m x n column-major, f32 or f16) from quantized
 * operands: A rows (qa, sa) and B columns (qb, sb) as produced above, k a
 * multiple of 32; code buffers 16-byte aligned with ld multiples of 16,
 * scale buffers 4-byte aligned with ld multiples of 4 and >= ceil(k / 128)
 * * 4. Products of e4m3 codes on the scaled matrix cores, fp32 sums. */
int remifa_gemm_fp8(remifa_handle_t h, int out_t, int64_t m, int64_t n, int64_t k, float alpha,
                    const uint8_t *qa, int64_t ldqa, const uint8_t *sa, int64_t ldsa,
                    const uint8_t *qb, int64_t ldqb, const uint8_t *sb, int64_t ldsb, float beta,
                    void *c, int64_t ldc);

/* ------------------------------------------------------------ solves */
/* Identity-pivot getrs (the reference's check solve host_getrs with
 * ipiv = 1..n, tests/testing_lu_gpu.hxx:1309 -> src/wrappers.cxx:477), on the
 * GPU: d_b (n x nrhs, fp32) <- (L U)^-1 d_b, or (L U)^-T d_b when trans != 0.
 * The factors are lu_t (REMIFA_R_32F or REMIFA_R_16F); arithmetic is fp32. */
int remifa_getrs_nopiv(remifa_handle_t h, int lu_t, int trans, int64_t n, int64_t nrhs,
                       const void *d_lu, int64_t ldlu, float *d_b, int64_t ldb);

/* remifa_getrs_nopiv without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate (code under HIP graph capture
 * setup, serving loops, harnesses that forbid hipMalloc). Same arguments,
 * kernels and results as remifa_getrs_nopiv, but its scratch (sync words,
 * inverted diagonal blocks, tagged solution words: about 36 MiB) is a
 * static device array that every GPU the library runs on holds until the
 * process ends, instead of the handle's workspace. n is limited to 32768
 * (REMIFA_ERR_WORKSPACE above it). The array is shared by every call on a
 * GPU: calls that overlap on one GPU (other streams, other threads) race. */
int remifa_getrs_nopiv_noalloc(remifa_handle_t h, int lu_t, int trans, int64_t n, int64_t nrhs,
                               const void *d_lu, int64_t ldlu, float *d_b, int64_t ldb);

/* FP64 residual R = B - A X (A fp32 n x n, X/B/R fp64 n x nrhs). No
 * reference counterpart (the reference checks on the host in FP64,
 * tests/errors.hxx:186-249). */
int remifa_residual_f64(remifa_handle_t h, int64_t n, int64_t nrhs, const float *d_a,
                        int64_t lda, const double *d_x, int64_t ldx, const double *d_b,
                        int64_t ldb, double *d_r, int64_t ldr);

/* Iterative refinement: X0 = (LU)^-1 B in fp32, then X += (LU)^-1 (B - A X)
 * with the residual in FP64 until the normwise backward error of every
 * column (tests/errors.hxx:186-249) is <= tol or maxit corrections. Returns
 * the corrections applied in *iters (-1 if not converged) and the final
 * backward error in *bwd_err. No reference counterpart. */
int remifa_ir_solve(remifa_handle_t h, int lu_t, int64_t n, int64_t nrhs, const float *d_a,
                    int64_t lda, const void *d_lu, int64_t ldlu, const double *d_b, int64_t ldb,
                    double *d_x, int64_t ldx, double tol, int maxit, int *iters, double *bwd_err,
                    double *history);

/* The same refinement for the transposed system A^T X = B (transposed
 * solves with the stored factors, residual B - A^T X in FP64, backward error
 * with ||A^T||_1). No reference counterpart: it is the adjoint solve of the
 * differentiable X = A^-1 B (gB = A^-T gX). */
int remifa_ir_solve_trans(remifa_handle_t h, int lu_t, int64_t n, int64_t nrhs,
                          const float *d_a, int64_t lda, const void *d_lu, int64_t ldlu,
                          const double *d_b, int64_t ldb, double *d_x, int64_t ldx, double tol,
                          int maxit, int *iters, double *bwd_err, double *history);

/* ------------------------------------------------------------ error measures */
/* The reference's error measures (tests/errors.hxx) in FP64 on the GPU, for
 * A (fp32, n x n) and solutions d_x of A X = B (FP64, n x nrhs), with the
 * factors d_lu (lu_t REMIFA_R_32F / REMIFA_R_16F, unit L below the
 * diagonal, U on and above it; may be null). out[0] = the normwise backward
 * error ||b - A x||_inf / (||A||_1 ||x||_inf + ||b||_inf)
 * (unsym_backward_error, :186-249), out[1] = the componentwise error
 * max_i |b - A x|_i / (|A||x| + |L||U||x|)_i (unsym_compwise_error,
 * :22-182; NaN when d_lu is null), out[2] = the Oettli-Prager form
 * max_i |b - A x|_i / (|A||x|)_i (printed by the same function), each the
 * worst over the columns (the reference's compwise error takes one column).
 * out is host memory (3 doubles); the call synchronises the stream. */
int remifa_error_measures(remifa_handle_t h, int lu_t, int64_t n, int64_t nrhs, const float *d_a,
                          int64_t lda, const void *d_lu, int64_t ldlu, const double *d_b,
                          int64_t ldb, const double *d_x, int64_t ldx, double *out);
/* Factor error of tests/errors.hxx:251-346 (factor_backward_error) in FP64:
 * out[0] = max_ij |A - L U|_ij / (|L||U|)_ij (|A - L U|_ij where the
 * denominator is 0), out[1] = max_ij |A - L U|_ij, out[2] = ||A - L U||_F /
 * ||A||_F, out[3] = ||A - L U||_F. O(n^3) FP64 work. out is host memory (4
 * doubles); the call synchronises the stream. */
int remifa_factor_error(remifa_handle_t h, int lu_t, int64_t n, const float *d_a, int64_t lda,
                        const void *d_lu, int64_t ldlu, double *out);

/* ------------------------------------------------------------ generators */
/* The benchmark generator on the GPU: a_ij uniform(-1,1) drawn row by row
 * from libstdc++'s minstd_rand0 seeded with seed (tests/genmat.hxx:68-102
 * draws the same way), then a_ii = fp32(alpha * sum_{j!=i} |a_ij|) (FP64
 * sum). alpha <= 0 gives gen_diagdom itself: a_ii = n, default seed. */
int remifa_gen_diagdom(remifa_handle_t h, int64_t n, uint64_t seed, double alpha, float *d_a,
                       int64_t lda);
/* Seeded uniform(-1,1) m x n block, row-major draw order. */
int remifa_gen_uniform(remifa_handle_t h, int64_t m, int64_t n, uint64_t seed, float *d_b,
                       int64_t ldb);

/* gen_randmat (tests/common.hxx:298-319): a_ij uniform(-1,1) drawn column
 * by column (j outer, i inner) from minstd_rand0 seeded with seed (the
 * reference's default argument is 1). The random matrix behind the
 * condition-number generator gen_posdef_cond (tests/common.hxx:526-727). */
int remifa_gen_randmat(remifa_handle_t h, int64_t m, int64_t n, uint64_t seed, float *d_a,
                       int64_t lda);
/* gen_arrowhead (tests/common.hxx:391-455) for an n x n matrix: 256 x 256
 * diagonal blocks and the last 256 rows and columns (the arrow) filled with
 * uniform(lo, hi) draws of a default-seeded minstd_rand0 in the reference's
 * loop order (later writes win), zeros elsewhere, a_jj = 512. The harness
 * uses [lo, hi] = [0, 1] (--arrowhead, --arrowhead-pos) and [-1, 1]
 * (--arrowhead-pos-neg). */
int remifa_gen_arrowhead(remifa_handle_t h, int64_t n, float lo, float hi, float *d_a,
                         int64_t lda);

/* ------------------------------------------------------------ equilibration */
/* LAPACK sgeequb on the GPU (the reference harness's --scale,
 * tests/testing_lu_gpu.hxx:288-311 -> src/wrappers.cxx:244-250): row and
 * column scale factors d_r (m), d_c (n) that are powers of two,
 * r_i = 1 / 2^trunc(log2 max_j |a_ij|), c_j likewise for the row-scaled
 * columns; *rowcnd, *colcnd, *amax and *info (0; i for a zero row i; m + j
 * for a zero column j, 1-based) as LAPACK returns them (host memory). */
int remifa_geequb(remifa_handle_t h, int64_t m, int64_t n, const float *d_a, int64_t lda,
                  float *d_r, float *d_c, float *rowcnd, float *colcnd, float *amax, int *info);
/* LAPACK slaqge on the GPU (src/wrappers.cxx:209-213): scale d_a in place by
 * rows, columns, both or neither from the geequb results (THRESH = 0.1);
 * *equed receives 'N', 'R', 'C' or 'B'. Exact for power-of-two factors. */
int remifa_laqge(remifa_handle_t h, int64_t m, int64_t n, float *d_a, int64_t lda,
                 const float *d_r, const float *d_c, float rowcnd, float colcnd, float amax,
                 char *equed);
/* The harness's conversion check (--check-conv, tests/common.hxx:107-174):
 * out[0] = max_ij |a_ij - a~_ij| / |a_ij| (|a_ij - a~_ij| where a_ij = 0),
 * out[1] = max_ij |a_ij - a~_ij|, with a~ the converted copy d_conv
 * (conv_t REMIFA_R_16F, REMIFA_R_16BF or REMIFA_R_32F) widened to FP64.
 * out is host memory (2 doubles); the call synchronises the stream. */
int remifa_conv_error(remifa_handle_t h, int conv_t, int64_t m, int64_t n, const float *d_a,
                      int64_t lda, const void *d_conv, int64_t ldc, double *out);

/* ------------------------------------------------------------ batched */
/* Factor `count` independent matrices of orders n[i] (<= max_n <= 512)
 * stored at d_ptrs[i] (fp32, leading dimension ld[i]) in place, in one
 * persistent launch. d_n, d_ld, d_ptrs are device arrays. max_n picks the
 * size bucket (64, 128, 256, 512) whose kernel runs; a caller with mixed
 * sizes gets the best schedule by sorting the batch by order and calling
 * once per bucket. No reference counterpart (NEW row of the hot path). */
int remifa_lu_nopiv_batched(remifa_handle_t h, int ctype, int64_t count, const int *d_n,
                            const int *d_ld, float *const *d_ptrs, int *d_info, int max_n);
/* Threads and persistent workgroups per CU of remifa_lu_nopiv_batched per
 * size bucket: rules[3i..3i+2] = (bucket 64 | 128 | 256 | 512, threads
 * (64 or 128 | 128 or 256 | 256 or 512 | 512), workgroups per CU 1..16);
 * count = 0: defaults. Filled by the autotuner (remifa_amd/autotune.py). */
int remifa_set_batched_tuning(remifa_handle_t h, int count, const int *rules);
/* Solve A_i x_i = b_i for every matrix of the batch with its factors
 * (fp32 right-hand sides, one per matrix, overwritten). */
int remifa_getrs_nopiv_batched(remifa_handle_t h, int64_t count, const int *d_n, const int *d_ld,
                               float *const *d_ptrs, float *const *d_b, int max_n);

#ifdef __cplusplus
}
#endif

#endif
