/*
 * hplai.h — C ABI of the MI355X-native HPL-AI mixed-precision solver
 * (libhplai_gfx950.so, built from hpl-ai_amd/csrc).
 *
 * Drop-in boundary for the hot path of RIKEN-RCCS/hpl-ai: each entry point
 * names the reference interface it replaces (file:line in that tree).
 *
 * Conventions
 *  - Matrices are column-major; leading dimensions are in elements (int64).
 *  - Pointers marked "device" are HIP device pointers owned by the caller.
 *  - `stream` is a hipStream_t passed as void* (NULL = default stream); calls
 *    are stream-ordered and asynchronous unless stated otherwise.
 *  - Every function returns HPLAI_OK (0) or a negative HPLAI_E* code; nothing
 *    aborts or throws across the ABI.  hplai_last_error() gives the message
 *    of the last failure on the calling thread.
 *  - One hplai_ctx per host thread; a ctx binds one device.
 */
#ifndef HPLAI_H
#define HPLAI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HPLAI_OK        0
#define HPLAI_EINVAL   -1   /* bad argument (shape, pointer, mode)            */
#define HPLAI_EHIP     -2   /* HIP runtime error                              */
#define HPLAI_ENOMEM   -3   /* device allocation failed                       */
#define HPLAI_ENOCONV  -4   /* IR did not reach harness < 16 within maxit     */
#define HPLAI_EPIVOT   -5   /* zero or non-finite pivot in the diagonal LU    */

/* Low-precision format of the Schur update operands */
#define HPLAI_FMT_F16      0   /* binary16, exponents all zero (reference: no -d)       */
#define HPLAI_FMT_F16_TILE 1   /* binary16 with per-b x b-tile power-of-two exponents   */
#define HPLAI_FMT_E4M3     2   /* OCP FP8 e4m3 with per-b x b-tile power-of-two exps    */

typedef struct hplai_ctx_s *hplai_ctx;

int         hplai_create(int device, hplai_ctx *out);
int         hplai_destroy(hplai_ctx ctx);
const char *hplai_last_error(void);
int         hplai_version(void);              /* MAJOR*10000 + MINOR*100 + PATCH */

/* ---------------- generator: bit-exact with hpl_rand.hpp ----------------
 * Element (i,j) of the n x n matrix is LCG state 1+i+n*j mapped to
 * [-0.5,0.5) and rounded to FP32; the diagonal is the row's FP64 sum of
 * |off-diagonals| (diagonal dominance).
 */
/* Replaces fill_one_panel_with_rand<float> (hpl_rand.hpp:59-106) as used by
 * pmatgen (matgen.hpp:7-22).  Writes the m x k block at global (i0,j0) into A.
 * d (device, length n, may be NULL): when given, global diagonal entries in
 * the block become (float)d[i]; when NULL they are left as LCG values. */
int hplai_gen_matrix_f32(hplai_ctx ctx, int64_t n, uint64_t seed, int64_t i0, int64_t j0,
                         int64_t m, int64_t k, const double *d, float *A, int64_t lda,
                         void *stream);
/* Replaces calc_diag / pdiaggen (hpl_rand.hpp:108-123, matgen.hpp:209-225):
 * d[t] = sum_{j != i} |a_ij| for i = i0+t, FP64, j ascending (bit-exact). */
int hplai_gen_diag_f64(hplai_ctx ctx, int64_t n, uint64_t seed, int64_t i0, int64_t cnt,
                       double *d, void *stream);
/* Replaces pcolvgen (matgen.hpp:193-207): b[t] = value of state 1+(i0+t)+n*n. */
int hplai_gen_rhs_f64(hplai_ctx ctx, int64_t n, uint64_t seed, int64_t i0, int64_t cnt,
                      double *b, void *stream);

/* ---------------- LU building blocks (FP32 panel) ---------------- */
/* Replaces getrf_nopiv(n, a, lda, piv, ldpiv) (getrf_nopiv.hpp:143-160):
 * unpivoted LU of the n x n block in place (unit-lower L, upper U), n <= 1024.
 * info_dev (device int, may be NULL) receives 0, or k+1 for the first zero /
 * non-finite pivot. */
int hplai_getrf_nopiv_f32(hplai_ctx ctx, int n, float *A, int64_t lda, int *info_dev,
                          void *stream);
/* Replaces update_left_panels (panel_trf.hpp:143-157, strsm "R","U","N","N"):
 * L(m x b) := L * U^-1, U the b x b upper triangle of Ukk. */
int hplai_trsm_lpanel_f32(hplai_ctx ctx, int64_t m, int b, const float *Ukk, int64_t ldu,
                          float *L, int64_t ldl, void *stream);
/* Replaces update_right_panels (panel_trf.hpp:158-171, strsm "L","L","N","U"):
 * U(b x n) := Lkk^-1 * U, Lkk unit lower. */
int hplai_trsm_upanel_f32(hplai_ctx ctx, int b, int64_t n, const float *Lkk, int64_t ldl,
                          float *U, int64_t ldu, void *stream);
/* Replaces convert_left_panels / convert_right_panels (panel_trf.hpp:173-233)
 * and pack_convert_{a,b}_opt (kernels/kernel.h:24-25).
 * side 0 (L panel): src is m x k column-major; dst is written K-contiguous:
 *   dst[i*ldd + kk] for row i, kk < k (row-major m x ldd).
 * side 1 (U panel): src is k x m column-major (k rows); dst[j*ldd + kk].
 * fmt: HPLAI_FMT_*; element size 2 (F16*) or 1 (E4M3) bytes.
 * Per-tile exponents: tiles are tb x tb (tb = tile size, e.g. the block size
 * b) indexed by (row or column)/tb relative to the panel start;
 * exps[t] receives e with q = round(x * 2^e) (0 for HPLAI_FMT_F16).
 * exps may be NULL only for HPLAI_FMT_F16.  Columns kk in [k, ldd) of dst
 * are not touched. */
int hplai_quantize(hplai_ctx ctx, int fmt, int side, int64_t m, int64_t k, int tb,
                   const float *src, int64_t lds, void *dst, int64_t ldd, int32_t *exps,
                   void *stream);
/* Replaces gemmschur / GemmControl (schur_updator.hpp:9-47, 178-269) and
 * hgemm/shgemm (fp16sim.hpp:197-233):
 *   C(m x n) -= sum_s 2^-(eL[s][i/tb] + eU[s][j/tb]) * Lq(:, slice s) Uq(slice s, :)
 * Lq is K-contiguous (Lq[i*ldl + kk]), Uq is K-contiguous (Uq[j*ldu + kk]);
 * K = nslices * kslice; FP32 accumulation, C FP32.  eL/eU are
 * [nslices][ntiles] int32 with row strides ldel/ldeu (NULL for HPLAI_FMT_F16);
 * i and j count from the first row / column of C and tb must be a multiple
 * of the kernel tile (128) when exponents are used. */
int hplai_schur_update(hplai_ctx ctx, int fmt, int64_t m, int64_t n, int nslices, int kslice,
                       const void *Lq, int64_t ldl, const int32_t *eL, int64_t ldel,
                       const void *Uq, int64_t ldu, const int32_t *eU, int64_t ldeu, int tb,
                       float *C, int64_t ldc, void *stream);
/* Replaces gemmschur(int n, int m, int k, const fp16 *a, int lda, const fp16 *b, int ldb,
 * float *c, int ldc) (schur_updator.hpp:28-31 -> shgemm fp16sim.hpp:217-233) with the
 * reference's own operand layout: C(n x m) -= A(n x k) B(k x m), all column-major,
 * A = the L panel as convert_left_panels writes it (column l at a + l*lda,
 * panel_trf.hpp:182-207), B = the U panel of convert_right_panels (column j at
 * b + j*ldb, panel_trf.hpp:208-233), binary16 bits; FP32 accumulation, C FP32
 * (device).  Any k >= 0 (the reference's C1 has k = b = 60).  The operands are
 * staged K-contiguous in a stream-ordered scratch allocation (hipMallocAsync,
 * (n + m) * roundup(k, 64) * 2 bytes) freed on the same stream. */
int hplai_gemmschur_f16(hplai_ctx ctx, int64_t n, int64_t m, int64_t k, const uint16_t *a, int64_t lda,
                        const uint16_t *b, int64_t ldb, float *c, int64_t ldc, void *stream);
/* Replaces the FP16-C overload gemmschur(fp16 a, fp16 b, fp16 c) (schur_updator.hpp:24-27,
 * hgemm with alpha = -1): C := fp16(C - cscale * 2^-(eL+eU) Lq Uq) with C an FP16
 * (binary16 bits) column-major m x n array, FP32 accumulation (the reference rounds
 * its accumulator to FP16 after every product).  m, n multiples of 128, ldc % 4 == 0,
 * C 8-byte aligned; fmt HPLAI_FMT_F16 or HPLAI_FMT_E4M3; other arguments as in
 * hplai_schur_update. */
int hplai_schur_update_f16c(hplai_ctx ctx, int fmt, int64_t m, int64_t n, int nslices, int kslice,
                            const void *Lq, int64_t ldl, const int32_t *eL, int64_t ldel,
                            const void *Uq, int64_t ldu, const int32_t *eU, int64_t ldeu, int tb,
                            uint16_t *C, int64_t ldc, float cscale, void *stream);

/* ---------------- Higham's HPL-AI matrix (--highammat) ---------------- */
/* Replaces HMGen (highammgen.hpp:5-27) with higham_mat_comp_beta
 * (higham_mat_impl.cpp:83-110): alpha = rho * beta with beta chosen so that
 * A = L U (L unit lower with alpha below the diagonal, U unit upper with beta
 * above it) has cond(A, inf) = cond.  The reference uses cond = 1000,
 * rho = 0.5 (main.cpp:258).  Host-only; no context needed. */
int hplai_higham_coeffs(int64_t n, double cond, double rho, double *alpha, double *beta);
/* Replaces pmatgen(HMGen) (matgen.hpp:24-78): the FP32 block
 * A(i0:i0+m, j0:j0+k) of the Higham matrix, column-major with lda.  alpha and
 * beta are signed: (+alpha, +beta) of hplai_higham_coeffs for the reference's
 * matrix, (-alpha, -beta) for the exact-cond one. */
int hplai_gen_higham_f32(hplai_ctx ctx, int64_t n, double alpha, double beta, int64_t i0, int64_t j0,
                         int64_t m, int64_t k, float *A, int64_t lda, void *stream);
/* Replaces otf_gemv(HMGen) (panel_gemv.hpp:86, otf_gemv.cpp:133-314):
 * y := y + scale * A x in FP64 with the Higham matrix regenerated. */
int hplai_higham_gemv_f64(hplai_ctx ctx, int64_t n, double alpha, double beta, double scale,
                          const double *x, double *y, void *stream);

/* ---------------- iterative refinement (FP64) ---------------- */
/* Replaces panel_gemv / otf_gemv / otf_gemv_kernel (panel_gemv.hpp:57-180,
 * otf_gemv.cpp:5-131): y := y + alpha * A x with A regenerated on the fly
 * from (n, seed) and diagonal d (device, FP64). */
int hplai_otf_gemv_f64(hplai_ctx ctx, int64_t n, uint64_t seed, const double *d, double alpha,
                       const double *x, double *y, void *stream);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): hplai_otf_gemv_f64 with the same kernels and results,
 * its FP64 partial sums kept in a 32 MiB device array that the library holds
 * per device until the process ends, so the call allocates nothing.
 * n * ceil(n / 2048) * 8 bytes must fit in it (n <= 92160), otherwise
 * HPLAI_EINVAL.  Calls that overlap on one device share the array and race. */
int hplai_otf_gemv_f64_noalloc(hplai_ctx ctx, int64_t n, uint64_t seed, const double *d, double alpha,
                               const double *x, double *y, void *stream);
/* Replaces panel_trsvL / tttrsvL (panel_trsv.hpp:115-124, 310-535):
 * x := L^-1 x, L the unit lower triangle of LU (FP32), x FP64.  bs is the
 * blocking (<= 1024; use the LU block size). */
int hplai_trsv_lower_f32(hplai_ctx ctx, int64_t n, int bs, const float *LU, int64_t lda,
                         double *x, void *stream);
/* Replaces panel_trsvU / tttrsvU (panel_trsv.hpp:142-153, 537-758). */
int hplai_trsv_upper_f32(hplai_ctx ctx, int64_t n, int bs, const float *LU, int64_t lda,
                         double *x, void *stream);
/* Replaces colv_infnorm (panel_norm.hpp:20-41): *out_dev = max_i |x_i|. */
int hplai_norm_inf_f64(hplai_ctx ctx, int64_t n, const double *x, double *out_dev,
                       void *stream);

/* ---------------- dense factor and multi-RHS solves ---------------- */
/* Replaces RDMAPanelLU::run (panel_trf.hpp:703-856) on a caller-supplied
 * matrix: unpivoted LU of the n x n FP32 matrix A (device, column-major,
 * lda >= n) in place, with the same kernels and schedule as hplai_solve
 * (FP32 panel, Schur update in `fmt`, `outer` panels per trailing update,
 * look-ahead on a second stream).  n % b == 0; b % 128 == 0 for the
 * tile-scaled formats.  Synchronous; HPLAI_EPIVOT on a zero pivot. */
int hplai_factor_f32(hplai_ctx ctx, int64_t n, int b, int fmt, int outer, int lookahead, float *A,
                     int64_t lda);
/* Solves with the factors of hplai_factor_f32 / hplai_solve (FP32 entries
 * widened to FP64, as panel_trsvL/U panel_trsv.hpp:310-758 do):
 * trans = 0: X := (L U)^-1 X;  trans = 1: X := (L U)^-T X = L^-T U^-T X,
 * the transposed solve for A^T lambda = g (no reference counterpart: it serves
 * the gradient of x = A^-1 b).  X (device, FP64) holds nrhs columns with
 * leading dimension ldx; bs (<= 1024) is the blocking, normally the LU's b. */
int hplai_lu_solve_f64(hplai_ctx ctx, int trans, int64_t n, int nrhs, int bs, const float *LU,
                       int64_t lda, double *X, int64_t ldx, void *stream);

/* No reference counterpart (it serves the differentiable form): FP64 GEMM on
 * the FP64 matrix cores, C := alpha op(A) op(B) + beta C, column-major, op(X)
 * = X^T when trans* != 0, C m x n, op(A) m x k; beta = 0 does not read C.
 * Used for the refinement residuals R = B - op(A) X and the gradient
 * dA = -Lambda X^T of hplai.autograd; results do not depend on scheduling
 * (a split K is summed in a fixed order). */
int hplai_gemm_f64(hplai_ctx ctx, int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                   const double *A, int64_t lda, const double *B, int64_t ldb, double beta, double *C,
                   int64_t ldc, void *stream);

/* ---------------- batched solver (no reference counterpart) ----------------
 * Many independent HPL-AI systems (order 1..1024, 1..4 FP64 right-hand
 * sides, each built by the reference generator from its own seed: element
 * (i,j) = LCG state 1+i+n*j, diagonal = FP64 row sum of |a_ij|, rhs r =
 * stream column n+r) are generated, factored (FP32 panels, FP16 Schur update
 * with FP32 accumulation) and refined in FP64 to harness < 16 by one
 * persistent launch: one workgroup per system, systems taken largest first
 * from a device counter.  The LU panel width is chosen per size bucket
 * ([1,128) [128,256) [256,512) [512,1024]) by hplai_batch_autotune. */
typedef struct hplai_batch_s *hplai_batch;
typedef struct hplai_batch_stats {
    double  seconds;        /* host wall time of the run (launch to completion)      */
    double  systems_per_s;
    double  p50_us;         /* per-system device time (wall clock inside the kernel)  */
    double  p99_us;
    double  max_us;
    int32_t failed;         /* systems with a zero pivot or a harness >= 16          */
    int32_t max_steps;      /* most residual evaluations of any right-hand side      */
    double  max_harness;
} hplai_batch_stats;
int hplai_batch_create(hplai_ctx ctx, int count, const int32_t *n, const int32_t *nrhs,
                       const uint64_t *seed, hplai_batch *out);
int hplai_batch_destroy(hplai_batch b);
/* panel width (16, 32 or 64) per size bucket */
int hplai_batch_set_nb(hplai_batch b, const int32_t nb[4]);
int hplai_batch_get_nb(hplai_batch b, int32_t nb[4]);
/* times every feasible panel width on up to 2 systems per CU of each bucket
 * and keeps the fastest; nb_out (may be NULL) receives the table */
int hplai_batch_autotune(hplai_batch b, int maxit, int32_t nb_out[4]);
/* mode 0: one persistent launch; 1: the same launch captured once in a
 * hipGraph and replayed; baselines: 2: one launch per system on one stream;
 * 3: one launch per system round-robin over four streams; 4: one launch of
 * one workgroup per system in input order.  Synchronous. */
int hplai_batch_run(hplai_batch b, int mode, int maxit, hplai_batch_stats *st);
/* per system: harness[4 s + r], steps[4 s + r] (r < 4; -1 / 0 past nrhs),
 * info[s], time_us[s]; any pointer may be NULL (host arrays) */
int hplai_batch_results(hplai_batch b, double *harness, int32_t *steps, int32_t *info,
                        double *time_us);
/* solution of system s: n_s x nrhs_s doubles, column-major, to host memory */
int hplai_batch_solution(hplai_batch b, int s, double *x_host);

/* ---------------- whole solve (main.cpp:256-452) ---------------- */
typedef struct hplai_params {
    int64_t  n;          /* matrix order, n % b == 0                               */
    int      b;          /* block size (reference `b`)                             */
    uint64_t seed;       /* 42 in the reference (main.cpp:174)                     */
    int      fmt;        /* HPLAI_FMT_*                                            */
    int      outer;      /* panels per trailing update (1 = reference schedule)    */
    int      lookahead;  /* 1: factor panel k+1 on a second stream during update k;
                            2: also block row k+1 there (the update's rows of block
                            row k+1 first, then the rows below)                      */
    int      lazy;       /* 1: -d lazy initialisation (lazy_init.hpp:92-227): the trailing
                            matrix starts at zero and accumulates -sum L U; each block gets
                            its generated a_ij added when it joins a panel                */
    int      maxit;      /* IR iterations (50 in the reference, 1 under -w)        */
    int      verbose;    /* print the reference's IR trace and !epoch lines        */
    int      keep_factors; /* leave A (factors) allocated for hplai_solution_*     */
    int      profile;    /* time every Schur-update launch with HIP events         */
    int      matrix;     /* HPLAI_MATRIX_*: the input matrix family                */
    double   cond;       /* Higham matrix: cond(A, inf); 0 = 1000 as in main.cpp:258 */
    int      trail16;    /* 1 (with lazy = 1): the trailing updates keep the partial sums of
                            blocks not yet factored in an FP16 n x n array, scaled by
                            sqrt(n sqrt(n)) (the reference's -d double-decker lower deck,
                            panel.hpp:45-83, panel_trf.hpp:343-345); b % 128 == 0.
                            2: also the compensated back buffer (--back_buffer,
                            back_buffer.hpp:81-170), a second FP16 n x n array           */
    int      split;      /* 0, or h with 0 < h < outer: each outer panel is factored in two
                            parts; after its first h inner blocks the rest of the outer
                            block and of its block row take them in one update (K = h b),
                            and the later inner blocks look left only as far as block h.
                            The trailing update keeps K = outer b.                       */
} hplai_params;

#define HPLAI_MATRIX_HPLAI        0   /* the reference's LCG matrix with diagonal row sums */
#define HPLAI_MATRIX_HIGHAM       1   /* --highammat as the reference generates it (main.cpp:85,
                                         258-276; pmatgen(HMGen) matgen.hpp:24-78): L U with
                                         +alpha below and +beta above the unit diagonals */
#define HPLAI_MATRIX_HIGHAM_KAPPA 2   /* no reference counterpart: Fasi and Higham's signs
                                         (-alpha, -beta, higham_mat_impl.cpp:113-125), so that
                                         cond(A, inf) = cond exactly */

/* Phases of the reference's Timer (timer.hpp:58-83), same order and names;
 * the communication phases are 0 on one GPU. */
#define HPLAI_PHASE_DIAG_BCAST    0
#define HPLAI_PHASE_LCOL_BCAST    1
#define HPLAI_PHASE_RROW_BCAST    2
#define HPLAI_PHASE_TEST          3
#define HPLAI_PHASE_WAIT          4
#define HPLAI_PHASE_DIAG_LU       5
#define HPLAI_PHASE_TRSM_L        6
#define HPLAI_PHASE_TRSM_R        7
#define HPLAI_PHASE_CONV_L        8
#define HPLAI_PHASE_CONV_R        9
#define HPLAI_PHASE_GEMM_UPDATE   10
#define HPLAI_PHASE_GEMM_PROGRESS 11
#define HPLAI_PHASE_LAZY_INIT     12
#define HPLAI_PHASE_WRITE_BACK    13
#define HPLAI_PHASE_IR_GEMV       14
#define HPLAI_PHASE_IR_GEMV_COMM  15
#define HPLAI_PHASE_IR_TRSV       16
#define HPLAI_PHASE_IR_TRSV_MV    17
#define HPLAI_PHASE_IR_TRSV_COMM  18
#define HPLAI_PHASE_TOTAL         19
#define HPLAI_PHASE_MISC          20
#define HPLAI_PHASES              21

typedef struct hplai_result {
    double time_gen;     /* matrix + diag + rhs generation (not in time_total)     */
    double time_lu;
    double time_ir;
    double time_total;   /* LU + IR, the reference's timed region (main.cpp:379-433) */
    double gflops;       /* (2/3 n^3 + 3/2 n^2) / time_total * 1e-9 (main.cpp:441)  */
    int    ir_steps;     /* residual evaluations                                   */
    double residual;     /* final ||b - Ax||_inf (-1 if not converged)             */
    double harness;      /* final HPL-AI harness (-1 if not converged)             */
    double trace_resid[64];
    double trace_harness[64];
    double checksum[3];  /* panel_check fingerprint of the FP32 factors             */
    /* with profile: Schur-update launches, their algorithmic flops (2 m n K each)
       and summed event-measured durations */
    int64_t schur_launches;
    double  schur_flops;
    double  schur_ms;
    /* the subset of those launches that ran the streamed 256 x 256 kernel, and
       their algorithmic HBM bytes ((m + n) K operand bytes + 8 m n for C) */
    int64_t schur_large_launches;
    double  schur_large_flops;
    double  schur_large_ms;
    double  schur_large_bytes;
    /* the updates inside the outer panel (block column and row block t take
       the t earlier inner blocks, K = t b, on the panel's stream) */
    int64_t schur_inner_launches;
    double  schur_inner_flops;
    double  schur_inner_ms;
    /* with profile: every Schur launch by class -- [0] the outer steps' trailing
       updates (K = outer b), [1] the rows below the outer block taking the
       earlier panels of the block (K = t b), [2] the updates inside the outer
       panel, [3] the row blocks of the block row (K = t b) */
    int64_t schur_class_launches[4];
    double  schur_class_flops[4];
    double  schur_class_ms[4];
    /* with profile: per-phase device time (seconds, summed kernel durations
       from HIP events; phases on the two streams overlap) and operation
       counts as the reference's Timer accumulates them; TOTAL is the LU + IR
       wall time and MISC = TOTAL - the other phases.  IR_TRSV covers the
       whole triangular solves (diagonal blocks and panel products). */
    double  phase_seconds[HPLAI_PHASES];
    double  phase_ops[HPLAI_PHASES];
} hplai_result;

/* Generate, factor and refine; synchronous.  Returns HPLAI_ENOCONV if the
 * harness stays >= 16 (result still filled). */
int hplai_solve(hplai_ctx ctx, const hplai_params *p, hplai_result *r);
/* After hplai_solve with keep_factors: copy the solution x (n doubles) and
 * the factors (n x n FP32, ld n) to host memory. */
int hplai_solution_x(hplai_ctx ctx, double *x_host);
int hplai_solution_factors(hplai_ctx ctx, float *lu_host, int64_t ldh);
/* After hplai_solve with keep_factors: copy the m x k block of the factors at
 * (i0, j0) to host memory (column-major, ldh >= m).  Lets a caller read the
 * factors of a matrix too large to copy whole (155 GB at n = 196608); no
 * reference counterpart (the reference's factors stay in its panels). */
int hplai_solution_block(hplai_ctx ctx, int64_t i0, int64_t j0, int64_t m, int64_t k, float *host,
                         int64_t ldh);
int hplai_release(hplai_ctx ctx);   /* free buffers kept by keep_factors */

#ifdef __cplusplus
}
#endif
#endif /* HPLAI_H */
