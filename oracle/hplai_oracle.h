/*
 * hplai_oracle.h — CPU restatement of HPL-AI's mixed-precision solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the oracle the HIP
 * implementation in hpl-ai_amd/ is tested against: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product path never calls it.
 *
 * Every function restates the reference algorithm (RIKEN-RCCS/hpl-ai) and
 * cites the file:line it follows.  Matrices are column-major.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   orc_lcg_*, orc_gen_*     : against oracle/_ref (the reference's own
 *                              hpl_rand.hpp compiled here), closed forms, and
 *                              the README:50 step-0 residual/harness.
 *   orc_f32_to_f16           : exhaustive over all 2^32 inputs vs F16C (x86).
 *   orc_f32_to_e4m3          : exhaustive vs a brute-force nearest-even
 *                              search over the 256 codes.
 *   orc_schur_*              : vs oracle/_ref's hgemm_opt/shgemm_opt
 *                              (kernels/avx2.cpp) bit-exactly.
 *   orc_getrf/trsm           : residual invariants ||LU - A|| and textbook
 *                              brute force at tiny sizes.
 *   orc_lu_* + orc_ir        : reference IR traces and --checksum values
 *                              (golden fixtures from the reference run).
 *   orc_otf_gemv             : vs oracle/_ref's otf_gemv_kernel and a dense
 *                              FP64 product.
 */
#ifndef HPLAI_ORACLE_H
#define HPLAI_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- LCG (hpl_rand.hpp:8-57) ---- */
void     orc_lcg_pow(uint64_t n, uint64_t *a, uint64_t *c);
uint64_t orc_lcg_state(uint64_t seed, uint64_t idx);  /* jump(idx) * initialize(seed) */
double   orc_lcg_value(uint64_t state);               /* RandStat::operator double    */
double   orc_elem(int64_t n, uint64_t seed, int64_t i, int64_t j);

/* ---- generator (hpl_rand.hpp:59-123, matgen.hpp:7-22,193-225) ---- */
/* Fill A(0:m-1,0:k-1) = global block starting at (i0,j0).  Diagonal
 * entries of the global matrix become (float)d_i when with_diag != 0. */
void orc_gen_block_f32(int64_t n, uint64_t seed, int64_t i0, int64_t j0,
                       int64_t m, int64_t k, float *A, int64_t lda, int with_diag);
void orc_gen_diag_f64(int64_t n, uint64_t seed, int64_t i0, int64_t cnt, double *d);
void orc_gen_rhs_f64(int64_t n, uint64_t seed, int64_t i0, int64_t cnt, double *b);

/* ---- low-precision rounding ---- */
uint16_t orc_f32_to_f16(float x);   /* IEEE binary16, round-to-nearest-even */
float    orc_f16_to_f32(uint16_t h);
uint8_t  orc_f32_to_e4m3(float x);  /* OCP FP8 e4m3fn, RNE, saturating at 448 */
float    orc_e4m3_to_f32(uint8_t q);
void     orc_round_f16_array(int64_t cnt, const float *x, float *y);  /* y = (float)(f16)x */
void     orc_f32_to_f16_array(int64_t cnt, const float *x, uint16_t *q);
void     orc_f32_to_e4m3_array(int64_t cnt, const float *x, uint8_t *q);
/* Per-tile exponent: largest e with amax * 2^e <= 448 (fmt 2) or 65504 (fmt 1) */
int      orc_tile_exp(int fmt, float amax);

/* ---- LU building blocks (FP32) ---- */
void orc_getrf_nopiv_f32(int n, float *a, int64_t lda);                 /* getrf_nopiv.hpp:124-141 */
void orc_trsm_runn_f32(int64_t m, int n, const float *u, int64_t ldu,
                       float *b, int64_t ldb);                           /* B := B U^-1 (trsmR)     */
void orc_trsm_llnu_f32(int m, int64_t n, const float *l, int64_t ldl,
                       float *b, int64_t ldb);                           /* B := L^-1 B (trsmL)     */
/* C(mxn) -= A(mxk) B(kxn): A, B hold FP16 values (as float), FP32 accumulate,
 * k sequential in blocks of 100 (kernels/avx2.cpp:193-302 shgemm_opt). */
void orc_schur_f32acc(int64_t m, int64_t n, int64_t k, const float *a, int64_t lda,
                      const float *b, int64_t ldb, float *c, int64_t ldc);
/* C(mxn, FP16 bits) -= A B with the accumulator rounded to FP16 after every
 * FMA (kernels/avx2.cpp:82-189 hgemm_opt). */
void orc_schur_f16acc(int64_t m, int64_t n, int64_t k, const uint16_t *a, int64_t lda,
                      const uint16_t *b, int64_t ldb, uint16_t *c, int64_t ldc);

/* ---- whole LU ---- */
enum { ORC_LU_PRECOMP = 0,   /* FP32 C, FP16 panels, FP32 accumulate (no -d)        */
       ORC_LU_LAZY    = 1,   /* -d: FP16 partial sums, FP16 accumulate, lazy init   */
       ORC_LU_FP8     = 2,   /* FP32 C, e4m3 panels with per-bxb-tile 2^e scaling   */
       ORC_LU_F16TILE = 3 }; /* FP32 C, binary16 panels with per-bxb-tile 2^e       */
/* Factor in place.  A must hold the generated matrix (with diagonal) except in
 * LAZY mode, where A is ignored on input and regenerated from (n, seed).
 * dhigh = (float)d is needed by LAZY mode (main.cpp:272). */
int orc_lu(int mode, int64_t n, int b, uint64_t seed, const float *dhigh,
           float *A, int64_t lda, int nthreads);

/* The library's 16-bit trailing storage (lazy initialisation, partial sums
 * of unfactored blocks in binary16 scaled by sqrt(n sqrt(n)), FP32-summed
 * trailing products rounded once per update; backbuf: the compensated back
 * buffer folded every 32 updates).  qmode: ORC_LU_PRECOMP (binary16 panels,
 * no exponents) or ORC_LU_FP8.  A is written (regenerated lazily). */
int orc_lu_trail16(int qmode, int backbuf, int64_t n, int b, uint64_t seed, const float *dhigh,
                   float *A, int64_t lda, int nthreads);

/* ---- IR (iterative_refinement.hpp:18-57) ---- */
/* y += alpha * A x with A regenerated (panel_gemv.hpp:16-83, otf_gemv.cpp) */
void   orc_otf_gemv(int64_t n, uint64_t seed, const double *d, double alpha,
                    const double *x, double *y, int nthreads);
void   orc_otf_gemv_rows(int64_t n, uint64_t seed, const double *d, double alpha,
                         const double *x, double *y, int64_t i0, int64_t cnt, int nthreads);
void   orc_trsv_lower_unit(int64_t n, const float *lu, int64_t lda, double *x); /* panel_trsv.hpp:115-124 */
void   orc_trsv_upper(int64_t n, const float *lu, int64_t lda, double *x);      /* panel_trsv.hpp:142-153 */
double orc_infnorm(int64_t n, const double *x);                                 /* panel_norm.hpp:8-19    */
/* Runs IR; writes per-step residual/harness into trace arrays (size maxit).
 * Returns the number of residual evaluations; *final_resid/_harness = -1 on
 * non-convergence, as the reference does. */
int orc_ir(int64_t n, uint64_t seed, const float *lu, int64_t lda, const double *d,
           const double *rhs, double *x, int maxit, double *trace_resid,
           double *trace_harness, double *final_resid, double *final_harness,
           int nthreads);

/* ---- factor fingerprint (panel_check.hpp:9-65) ---- */
void orc_checksum(int64_t n, const float *lu, int64_t lda, double out[3]);

#ifdef __cplusplus
}
#endif
#endif
