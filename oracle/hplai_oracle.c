/*
 * hplai_oracle.c — CPU restatement of HPL-AI's mixed-precision solve.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 *
 * Each function cites the reference file:line (RIKEN-RCCS/hpl-ai) whose
 * arithmetic it restates.  Built with -ffp-contract=off so every rounding
 * step is explicit: where the reference relies on a fused multiply-add the
 * code says fmaf/fma.
 */
#include "hplai_oracle.h"
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* LCG: x <- a x + c (mod 2^64), a = 6364136223846793005, c = 1        */
/* hpl_rand.hpp:8-33 (RandCoeff), 35-57 (RandStat)                     */
/* ------------------------------------------------------------------ */
#define LCG_A 6364136223846793005ULL
#define LCG_C 1ULL

typedef struct { uint64_t a, c; } coeff_t;

/* (l * r)(x) = l(r(x)): hpl_rand.hpp:16-18 */
static coeff_t coeff_mul(coeff_t l, coeff_t r)
{
    coeff_t o = { l.a * r.a, l.a * r.c + l.c };
    return o;
}

/* RandCoeff::pow, hpl_rand.hpp:20-32.  The reference recurses on n/2; all
 * factors are powers of one affine map and commute, so square-and-multiply
 * yields the identical 64-bit coefficients. */
static coeff_t coeff_pow(coeff_t base, uint64_t n)
{
    coeff_t r = { 1, 0 };
    while (n) {
        if (n & 1) r = coeff_mul(r, base);
        base = coeff_mul(base, base);
        n >>= 1;
    }
    return r;
}

static const coeff_t LCG1 = { LCG_A, LCG_C };

void orc_lcg_pow(uint64_t n, uint64_t *a, uint64_t *c)
{
    coeff_t p = coeff_pow(LCG1, n);
    *a = p.a;
    *c = p.c;
}

/* RandStat::initialize (hpl_rand.hpp:38-43) then jump(idx) (hpl_rand.hpp:157-159) */
uint64_t orc_lcg_state(uint64_t seed, uint64_t idx)
{
    uint64_t x0 = LCG_A * seed + LCG_C;
    coeff_t j = coeff_pow(LCG1, idx);
    return j.a * x0 + j.c;
}

/* RandStat::operator double, hpl_rand.hpp:50-52: (int64)x * 0x1.fffffffffffffP-65 */
double orc_lcg_value(uint64_t x)
{
    return (double)(int64_t)x * 0x1.fffffffffffffP-65;
}

double orc_elem(int64_t n, uint64_t seed, int64_t i, int64_t j)
{
    return orc_lcg_value(orc_lcg_state(seed, (uint64_t)i + (uint64_t)n * (uint64_t)j));
}

/* calc_diag, hpl_rand.hpp:108-123: d_i = sum_{j != i} |a_ij|, FP64, j ascending */
static double calc_diag(int64_t i, int64_t n, uint64_t seed)
{
    coeff_t jn = coeff_pow(LCG1, (uint64_t)n);
    uint64_t s = orc_lcg_state(seed, (uint64_t)i);
    double sum = 0.0;
    for (int64_t j = 0; j < n; ++j) {
        if (i != j) sum += fabs(orc_lcg_value(s));
        s = jn.a * s + jn.c;
    }
    return sum;
}

/* fill_one_panel_with_rand, hpl_rand.hpp:59-106 (with calc_diag=true for
 * diagonal blocks, as pmatgen does, matgen.hpp:7-22).  Element (i,j) is LCG
 * state 1+i+n*j; diagonal elements become (float)d_i. */
void orc_gen_block_f32(int64_t n, uint64_t seed, int64_t i0, int64_t j0,
                       int64_t m, int64_t k, float *A, int64_t lda, int with_diag)
{
    coeff_t jn = coeff_pow(LCG1, (uint64_t)n);
    #pragma omp parallel for schedule(static)
    for (int64_t jj = 0; jj < k; ++jj) {
        uint64_t s = orc_lcg_state(seed, (uint64_t)i0 + (uint64_t)n * (uint64_t)(j0 + jj));
        for (int64_t ii = 0; ii < m; ++ii) {
            A[jj * lda + ii] = (float)orc_lcg_value(s);
            s = LCG_A * s + LCG_C;
        }
    }
    (void)jn;
    if (with_diag) {
        int64_t lo = i0 > j0 ? i0 : j0;
        int64_t hi = (i0 + m) < (j0 + k) ? (i0 + m) : (j0 + k);
        #pragma omp parallel for schedule(dynamic, 16)
        for (int64_t g = lo; g < hi; ++g)
            A[(g - j0) * lda + (g - i0)] = (float)calc_diag(g, n, seed);
    }
}

/* pdiaggen, matgen.hpp:209-225 */
void orc_gen_diag_f64(int64_t n, uint64_t seed, int64_t i0, int64_t cnt, double *d)
{
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < cnt; ++t) d[t] = calc_diag(i0 + t, n, seed);
}

/* pcolvgen, matgen.hpp:193-207: b is column n of the stream, b_i = state 1+i+n*n */
void orc_gen_rhs_f64(int64_t n, uint64_t seed, int64_t i0, int64_t cnt, double *b)
{
    uint64_t s = orc_lcg_state(seed, (uint64_t)i0 + (uint64_t)n * (uint64_t)n);
    for (int64_t t = 0; t < cnt; ++t) {
        b[t] = orc_lcg_value(s);
        s = LCG_A * s + LCG_C;
    }
}

/* ------------------------------------------------------------------ */
/* Low-precision rounding                                              */
/* ------------------------------------------------------------------ */
static uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* Round the 24-bit significand `sig` (with implicit bit) right by `sh` bits,
 * nearest-even.  sh >= 1. */
static uint32_t rne_shift(uint32_t sig, int sh)
{
    if (sh >= 32) return 0;
    uint32_t q = sig >> sh;
    uint32_t rem = sig & ((1u << sh) - 1u);
    uint32_t half = 1u << (sh - 1);
    if (rem > half || (rem == half && (q & 1u))) ++q;
    return q;
}

/* IEEE binary16, RNE: the conversion F16C performs with imm8=0
 * (fp16sim.hpp:111, kernels/avx2.cpp:95-109). */
uint16_t orc_f32_to_f16(float x)
{
    uint32_t u = f2u(x);
    uint16_t sign = (uint16_t)((u >> 16) & 0x8000u);
    uint32_t ax = u & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint16_t)(sign | 0x7e00u | ((ax >> 13) & 0x3ffu)); /* NaN */
    if (ax == 0x7f800000u) return (uint16_t)(sign | 0x7c00u);
    int e = (int)(ax >> 23) - 127;           /* unbiased exponent */
    uint32_t sig = (ax & 0x7fffffu) | (ax >= 0x00800000u ? 0x800000u : 0u);
    if (ax < 0x00800000u) e = -126;          /* FP32 subnormal */
    if (e >= -14) {
        /* normal half: keep 10 fraction bits; carry may bump the exponent */
        uint32_t q = rne_shift(sig, 13);     /* 11-bit significand, maybe 2^11 */
        int he = e + 15;
        if (q >= 0x800u) { q >>= 1; ++he; }
        if (he >= 31) return (uint16_t)(sign | 0x7c00u);
        return (uint16_t)(sign | ((uint32_t)he << 10) | (q & 0x3ffu));
    }
    /* subnormal half: value = q * 2^-24 */
    int sh = 13 + (-14 - e);
    uint32_t q = rne_shift(sig, sh);         /* may round up to 0x400 = min normal */
    return (uint16_t)(sign | q);
}

float orc_f16_to_f32(uint16_t h)
{
    uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
    uint32_t e = (h >> 10) & 0x1fu;
    uint32_t f = h & 0x3ffu;
    if (e == 0x1f) return u2f(sign | 0x7f800000u | (f << 13));
    if (e == 0) {
        float v = (float)f * 0x1p-24f;
        return sign ? -v : v;
    }
    return u2f(sign | ((e + 112u) << 23) | (f << 13));
}

void orc_round_f16_array(int64_t cnt, const float *x, float *y)
{
    for (int64_t i = 0; i < cnt; ++i) y[i] = orc_f16_to_f32(orc_f32_to_f16(x[i]));
}

void orc_f32_to_f16_array(int64_t cnt, const float *x, uint16_t *q)
{
    for (int64_t i = 0; i < cnt; ++i) q[i] = orc_f32_to_f16(x[i]);
}

/* OCP FP8 e4m3fn: bias 7, 3 fraction bits, max 448 (0x7e), 0x7f = NaN, no
 * infinities; subnormals q * 2^-9.  Round-to-nearest-even; magnitudes that
 * round above 448 saturate to 448. */
uint8_t orc_f32_to_e4m3(float x)
{
    uint32_t u = f2u(x);
    uint8_t sign = (uint8_t)((u >> 24) & 0x80u);
    uint32_t ax = u & 0x7fffffffu;
    if (ax > 0x7f800000u) return (uint8_t)(sign | 0x7fu);
    if (ax >= 0x7f800000u) return (uint8_t)(sign | 0x7eu);
    int e = (int)(ax >> 23) - 127;
    uint32_t sig = (ax & 0x7fffffu) | (ax >= 0x00800000u ? 0x800000u : 0u);
    if (ax < 0x00800000u) e = -126;
    if (e >= -6) {
        uint32_t q = rne_shift(sig, 20);     /* 4-bit significand */
        int qe = e + 7;
        if (q >= 0x10u) { q >>= 1; ++qe; }
        if (qe > 15 || (qe == 15 && (q & 7u) == 7u)) return (uint8_t)(sign | 0x7eu);
        return (uint8_t)(sign | ((uint32_t)qe << 3) | (q & 7u));
    }
    int sh = 20 + (-6 - e);
    uint32_t q = rne_shift(sig, sh);         /* may round to 8 = min normal */
    return (uint8_t)(sign | q);
}

float orc_e4m3_to_f32(uint8_t q)
{
    uint32_t sign = (q & 0x80u) ? 1u : 0u;
    uint32_t e = (q >> 3) & 0xfu, f = q & 7u;
    float v;
    if (e == 15 && f == 7) return NAN;
    if (e == 0) v = (float)f * 0x1p-9f;
    else v = ldexpf(1.0f + (float)f * 0.125f, (int)e - 7);
    return sign ? -v : v;
}

/* Per-tile power-of-two exponent: the largest e with amax * 2^e <= fmt_max,
 * where fmt_max = mr * 2^er (448 = 1.75*2^8 for e4m3; 65504 = (2-2^-10)*2^15
 * for binary16).  Exact integer arithmetic on the FP32 bits of amax.
 * (The e4m3 form generalises the reference's scalea, panel_trf.hpp:717.) */
static int tile_exp(float amax, uint32_t mr_frac_bits, int er)
{
    if (!(amax > 0.0f)) return 0;
    uint32_t u = f2u(amax);
    int E;
    uint32_t frac;
    if (u >= 0x00800000u) {
        E = (int)(u >> 23) - 127;
        frac = u & 0x7fffffu;
    } else {                        /* subnormal amax: normalise */
        int lz = __builtin_clz(u) - 8;
        E = -126 - lz;
        frac = (u << lz) & 0x7fffffu;
    }
    return er - E - (frac > mr_frac_bits ? 1 : 0);
}
/* e4m3 exponents become E8M0 block scales 127 - e of the scaled MFMA, so the
 * library keeps them in [-127, 127]: a tile whose amax is below about 2^-119
 * takes e = 127 and quantizes to (near) zero. */
static int tile_exp_e4m3(float amax)
{
    int e = tile_exp(amax, 0x600000u, 8);
    return e > 127 ? 127 : (e < -127 ? -127 : e);
}
#define E4M3_MR_FRAC 0x600000u   /* 1.75  */
#define E4M3_ER 8
#define F16_MR_FRAC  0x7fe000u   /* 1.9990234375 */
#define F16_ER 15

int orc_tile_exp(int fmt, float amax)
{
    return fmt == ORC_LU_FP8 ? tile_exp_e4m3(amax) : tile_exp(amax, F16_MR_FRAC, F16_ER);
}

void orc_f32_to_e4m3_array(int64_t cnt, const float *x, uint8_t *q)
{
    for (int64_t i = 0; i < cnt; ++i) q[i] = orc_f32_to_e4m3(x[i]);
}

/* ------------------------------------------------------------------ */
/* LU building blocks (FP32)                                           */
/* ------------------------------------------------------------------ */

/* getrf_nopiv_small, getrf_nopiv.hpp:30-46 */
static void getrf_small(int n, float *a, int64_t lda)
{
    for (int k = 0; k < n; ++k) {
        float inv = 1.0f / a[lda * k + k];
        for (int i = k + 1; i < n; ++i) a[k * lda + i] *= inv;
        for (int j = k + 1; j < n; ++j) {
            float akj = a[j * lda + k];
            for (int i = k + 1; i < n; ++i)
                a[j * lda + i] = a[j * lda + i] - a[k * lda + i] * akj;
        }
    }
}

/* strsm_("R","U","N","N") (getrf_nopiv.hpp:14-18); netlib column order with
 * reciprocal of the pivot. */
void orc_trsm_runn_f32(int64_t m, int n, const float *u, int64_t ldu, float *b, int64_t ldb)
{
    #pragma omp parallel for schedule(static) if (m > 4096)
    for (int64_t i0 = 0; i0 < m; i0 += 256) {
        int64_t i1 = i0 + 256 < m ? i0 + 256 : m;
        for (int j = 0; j < n; ++j) {
            for (int k = 0; k < j; ++k) {
                float ukj = u[(int64_t)j * ldu + k];
                if (ukj != 0.0f)
                    for (int64_t i = i0; i < i1; ++i)
                        b[(int64_t)j * ldb + i] = b[(int64_t)j * ldb + i] - ukj * b[(int64_t)k * ldb + i];
            }
            float t = 1.0f / u[(int64_t)j * ldu + j];
            for (int64_t i = i0; i < i1; ++i) b[(int64_t)j * ldb + i] *= t;
        }
    }
}

/* strsm_("L","L","N","U") (getrf_nopiv.hpp:24-28); netlib column order. */
void orc_trsm_llnu_f32(int m, int64_t n, const float *l, int64_t ldl, float *b, int64_t ldb)
{
    #pragma omp parallel for schedule(static) if (n > 256)
    for (int64_t j = 0; j < n; ++j) {
        float *bj = b + j * ldb;
        for (int k = 0; k < m; ++k) {
            float bk = bj[k];
            if (bk != 0.0f)
                for (int i = k + 1; i < m; ++i) bj[i] = bj[i] - bk * l[(int64_t)k * ldl + i];
        }
    }
}

/* FP32 C -= A B with A, B FP32 (getrf's inner sgemm_, schur_updator.hpp:18-23). */
static void sgemm_minus(int64_t m, int64_t n, int64_t k, const float *a, int64_t lda,
                        const float *b, int64_t ldb, float *c, int64_t ldc)
{
    for (int64_t j = 0; j < n; ++j)
        for (int64_t l = 0; l < k; ++l) {
            float blj = b[j * ldb + l];
            for (int64_t i = 0; i < m; ++i)
                c[j * ldc + i] = c[j * ldc + i] - a[l * lda + i] * blj;
        }
}

/* getrf_nopiv<float>, getrf_nopiv.hpp:124-141 (NSMALL = 16) */
void orc_getrf_nopiv_f32(int n, float *a, int64_t lda)
{
    const int ns = 16;
    for (int k = 0; k < n; k += ns) {
        int nn = n - k < ns ? n - k : ns;
        getrf_small(nn, a + lda * k + k, lda);
        if (n - k > nn) {
            orc_trsm_llnu_f32(nn, n - k - nn, a + lda * k + k, lda, a + lda * (k + nn) + k, lda);
            orc_trsm_runn_f32(n - k - nn, nn, a + lda * k + k, lda, a + lda * k + k + nn, lda);
            sgemm_minus(n - k - nn, n - k - nn, nn, a + lda * k + k + nn, lda,
                        a + lda * (k + nn) + k, lda, a + lda * (k + nn) + (k + nn), lda);
        }
    }
}

/* shgemm_opt (kernels/avx2.cpp:249-302) with alpha = -1, beta = 1:
 * for each k-block of 100, acc = sum_k a_ik b_kj in FP32 (k ascending), then
 * c = c - acc.  Products of FP16 values are exact in FP32, so the reference's
 * fused multiply-add and this mul+add round identically. */
void orc_schur_f32acc(int64_t m, int64_t n, int64_t k, const float *a, int64_t lda,
                      const float *b, int64_t ldb, float *c, int64_t ldc)
{
    const int64_t BK = 100, RI = 64;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t j0 = 0; j0 < n; j0 += 4) {
        int64_t nj = n - j0 < 4 ? n - j0 : 4;
        float acc[4][64];
        for (int64_t i0 = 0; i0 < m; i0 += RI) {
            int64_t ni = m - i0 < RI ? m - i0 : RI;
            for (int64_t l0 = 0; l0 < k; l0 += BK) {
                int64_t l1 = l0 + BK < k ? l0 + BK : k;
                for (int jj = 0; jj < 4; ++jj)
                    for (int ii = 0; ii < 64; ++ii) acc[jj][ii] = 0.0f;
                for (int64_t l = l0; l < l1; ++l) {
                    const float *al = a + l * lda + i0;
                    for (int jj = 0; jj < nj; ++jj) {
                        float blj = b[(j0 + jj) * ldb + l];
                        for (int64_t ii = 0; ii < ni; ++ii) acc[jj][ii] = acc[jj][ii] + al[ii] * blj;
                    }
                }
                for (int jj = 0; jj < nj; ++jj)
                    for (int64_t ii = 0; ii < ni; ++ii)
                        c[(j0 + jj) * ldc + i0 + ii] = c[(j0 + jj) * ldc + i0 + ii] - acc[jj][ii];
            }
        }
    }
}

/* hgemm_opt (kernels/avx2.cpp:82-189), alpha = -1: the accumulator is rounded
 * to FP16 after every FMA (_mm256_cvtps_ph(_mm256_fmadd_ps(..),0)); the
 * epilogue is c = fp16(alpha*acc + c). */
void orc_schur_f16acc(int64_t m, int64_t n, int64_t k, const uint16_t *a, int64_t lda,
                      const uint16_t *b, int64_t ldb, uint16_t *c, int64_t ldc)
{
    const int64_t BK = 100;
    #pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < n; ++j)
        for (int64_t l0 = 0; l0 < k; l0 += BK) {
            int64_t l1 = l0 + BK < k ? l0 + BK : k;
            for (int64_t i = 0; i < m; ++i) {
                uint16_t acc = 0;
                for (int64_t l = l0; l < l1; ++l) {
                    float p = fmaf(orc_f16_to_f32(a[l * lda + i]), orc_f16_to_f32(b[j * ldb + l]),
                                   orc_f16_to_f32(acc));
                    acc = orc_f32_to_f16(p);
                }
                float cn = fmaf(-1.0f, orc_f16_to_f32(acc), orc_f16_to_f32(c[j * ldc + i]));
                c[j * ldc + i] = orc_f32_to_f16(cn);
            }
        }
}

/* ------------------------------------------------------------------ */
/* Whole LU: right-looking, no pivoting, block b                       */
/* RDMAPanelLU::run, panel_trf.hpp:703-856 (arithmetic; the grid and   */
/* the look-ahead only reorder which trailing blocks update when)      */
/* ------------------------------------------------------------------ */

/* Quantize a panel (rows r0..r1 of block-column or row) with per-b x b tile
 * exponents; writes dequantized FP32 values q * 2^-e into out. */
static void quant_tiles(int mode, int64_t m, int64_t k, int b, const float *src, int64_t lds,
                        float *out, int64_t ldo)
{
    for (int64_t tj = 0; tj < k; tj += b)
        for (int64_t ti = 0; ti < m; ti += b) {
            int64_t mi = m - ti < b ? m - ti : b, kj = k - tj < b ? k - tj : b;
            float amax = 0.0f;
            for (int64_t j = tj; j < tj + kj; ++j)
                for (int64_t i = ti; i < ti + mi; ++i) {
                    float v = fabsf(src[j * lds + i]);
                    if (v > amax) amax = v;
                }
            int e = mode == ORC_LU_FP8 ? tile_exp_e4m3(amax) : tile_exp(amax, F16_MR_FRAC, F16_ER);
            for (int64_t j = tj; j < tj + kj; ++j)
                for (int64_t i = ti; i < ti + mi; ++i) {
                    float s = ldexpf(src[j * lds + i], e);
                    float q = mode == ORC_LU_FP8 ? orc_e4m3_to_f32(orc_f32_to_e4m3(s))
                                                 : orc_f16_to_f32(orc_f32_to_f16(s));
                    out[j * ldo + i] = ldexpf(q, -e);
                }
        }
}

/* With HPLAI_ORACLE_PROGRESS set, the long oracle runs (the LU and the
 * refinement at full size take minutes) report each step on stderr. */
static int orc_progress(void)
{
    return getenv("HPLAI_ORACLE_PROGRESS") != NULL;
}

int orc_lu(int mode, int64_t n, int b, uint64_t seed, const float *dhigh,
           float *A, int64_t lda, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if (n % b) return -1;
    int64_t nb = n / b;
    float *Lq = (float *)malloc(sizeof(float) * (size_t)n * b);
    float *Uq = (float *)malloc(sizeof(float) * (size_t)n * b);
    uint16_t *P = NULL, *L16 = NULL, *U16 = NULL;
    float scalea = 1.0f, downscale = 1.0f;
    coeff_t jn = coeff_pow(LCG1, (uint64_t)n);
    if (mode == ORC_LU_LAZY) {
        /* pmatgen0 (matgen.hpp:80-105): partial sums start at zero */
        P = (uint16_t *)calloc((size_t)n * n, sizeof(uint16_t));
        L16 = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)n * b);
        U16 = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)n * b);
        /* Matgen: scalea = sqrt(n sqrt(n)), hpl_rand.hpp:154; panel_trf.hpp:717-719 */
        double sa = sqrt((double)n * sqrt((double)n));
        scalea = (float)sa;
        downscale = (float)(1. / (double)(scalea * 1.0f));
    }
    if (!Lq || !Uq || (mode == ORC_LU_LAZY && (!P || !L16 || !U16))) {
        free(Lq); free(Uq); free(P); free(L16); free(U16);
        return -2;
    }
    for (int64_t k = 0; k < nb; ++k) {
        int64_t kk = k * b, m = n - kk - b;
        float *Akk = A + kk * lda + kk;
        if (mode == ORC_LU_LAZY) {
            /* LazyInitializer::update_diag / update, lazy_init.hpp:92-227:
             * A = downscale * (float)P + a_ij over the block column and row. */
            #pragma omp parallel for schedule(static)
            for (int64_t j = kk; j < n; ++j) {
                int64_t ilo = j < kk + b ? kk : kk, ihi = j < kk + b ? n : kk + b;
                uint64_t s = orc_lcg_state(seed, (uint64_t)ilo + (uint64_t)n * (uint64_t)j);
                for (int64_t i = ilo; i < ihi; ++i) {
                    float aij = (i == j) ? dhigh[i] : (float)orc_lcg_value(s);
                    A[j * lda + i] = fmaf(downscale, orc_f16_to_f32(P[j * n + i]), aij);
                    s = LCG_A * s + LCG_C;
                }
            }
            (void)jn;
        }
        if (orc_progress() && k % 8 == 0) {
            fprintf(stderr, "orc_lu: block %lld of %lld\n", (long long)k, (long long)nb);
            fflush(stderr);
        }
        /* getrf_nopiv on the diagonal block (panel_trf.hpp:771) */
        orc_getrf_nopiv_f32(b, Akk, lda);
        if (m == 0) break;
        float *L = Akk + b, *U = Akk + (int64_t)b * lda;
        /* update_left_panels / update_right_panels, panel_trf.hpp:143-171 */
        orc_trsm_runn_f32(m, b, Akk, lda, L, lda);
        orc_trsm_llnu_f32(b, m, Akk, lda, U, lda);
        float *C = Akk + (int64_t)b * lda + b;
        if (mode == ORC_LU_PRECOMP) {
            /* convert_*_panels with scale 1 (panel_trf.hpp:173-233), shgemm */
            for (int64_t j = 0; j < b; ++j)
                orc_round_f16_array(m, L + j * lda, Lq + j * m);
            for (int64_t j = 0; j < m; ++j)
                orc_round_f16_array(b, U + j * lda, Uq + j * b);
            orc_schur_f32acc(m, m, b, Lq, m, Uq, b, C, lda);
        } else if (mode == ORC_LU_LAZY) {
            /* L panel scaled by scalea before the FP16 cast (panel_trf.hpp:790) */
            for (int64_t j = 0; j < b; ++j)
                for (int64_t i = 0; i < m; ++i)
                    L16[j * m + i] = orc_f32_to_f16(scalea * L[j * lda + i]);
            for (int64_t j = 0; j < m; ++j)
                for (int64_t i = 0; i < b; ++i)
                    U16[j * b + i] = orc_f32_to_f16(U[j * lda + i]);
            orc_schur_f16acc(m, m, b, L16, m, U16, b, P + (kk + b) * n + kk + b, n);
        } else {
            /* FP8 e4m3 / per-tile power-of-two scales (generalises scalea, panel_trf.hpp:717-719) */
            quant_tiles(mode, m, b, b, L, lda, Lq, m);
            quant_tiles(mode, b, m, b, U, lda, Uq, b);
            orc_schur_f32acc(m, m, b, Lq, m, Uq, b, C, lda);
        }
    }
    free(Lq); free(Uq); free(P); free(L16); free(U16);
    return 0;
}

/* 16-bit trailing storage: the reference's -d keeps the partial sums of the
 * blocks that have not joined a panel in binary16 (the lower deck of the
 * double-decker layout, panel.hpp:45-83), scaled by scalea = sqrt(n sqrt(n))
 * (hpl_rand.hpp:154, panel_trf.hpp:717-719), and merges them when a block
 * joins a panel, A = downscale * P + a_ij (LazyInitializer::update_diag /
 * update, lazy_init.hpp:92-227).  This restates the library's form of it
 * (its trail16 option): the products are summed in FP32 (the reference's
 * hgemm rounds its accumulator to binary16 after every product,
 * kernels/avx2.cpp:82-189), and each trailing update rounds once,
 * P := fp16(P - scalea * sum), with the L panel quantized unscaled (binary16,
 * qmode ORC_LU_PRECOMP) or with per-tile e4m3 exponents (ORC_LU_FP8).
 * backbuf != 0 adds the compensated back buffer Q (back_buffer.hpp:81-170,
 * bb_writeback_b_impl_fp16 back_buffer.cpp:32-55 in binary16 arithmetic):
 * after every 32nd trailing update the blocks that do not join the next
 * panel fold P into Q, Q := fl(Q + P/32), P := 32 (P/32 - (Q_new - Q)),
 * and the merge uses x = P + 32 Q (lazy_init.hpp:229-329).  Binary16
 * additions are computed in FP32 and rounded once, which equals binary16
 * arithmetic (24 >= 2 * 11 + 2 significand bits). */
static float h2f(uint16_t h) { return orc_f16_to_f32(h); }

int orc_lu_trail16(int qmode, int backbuf, int64_t n, int b, uint64_t seed, const float *dhigh,
                   float *A, int64_t lda, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    if (n % b || (qmode != ORC_LU_PRECOMP && qmode != ORC_LU_FP8)) return -1;
    const int64_t nb = n / b;
    float *Lq = (float *)malloc(sizeof(float) * (size_t)n * b);
    float *Uq = (float *)malloc(sizeof(float) * (size_t)n * b);
    float *T = (float *)malloc(sizeof(float) * (size_t)n * n);
    uint16_t *P = (uint16_t *)calloc((size_t)n * n, sizeof(uint16_t));
    uint16_t *Q = backbuf ? (uint16_t *)calloc((size_t)n * n, sizeof(uint16_t)) : NULL;
    if (!Lq || !Uq || !T || !P || (backbuf && !Q)) {
        free(Lq); free(Uq); free(T); free(P); free(Q);
        return -2;
    }
    const float pscale = (float)sqrt((double)n * sqrt((double)n));
    const float pdown = (float)(1.0 / (double)pscale);
    for (int64_t k = 0; k < nb; ++k) {
        const int64_t kk = k * b, e = kk + b, m = n - e;
        /* merge: block column k (rows kk..n) and block row k (columns e..n) */
        #pragma omp parallel for schedule(static)
        for (int64_t j = kk; j < n; ++j) {
            const int64_t ihi = j < e ? n : e;
            uint64_t s = orc_lcg_state(seed, (uint64_t)kk + (uint64_t)n * (uint64_t)j);
            for (int64_t i = kk; i < ihi; ++i) {
                const float aij = (i == j) ? dhigh[i] : (float)orc_lcg_value(s);
                float x = h2f(P[j * n + i]);
                if (Q) x = fmaf(32.0f, h2f(Q[j * n + i]), x);
                A[j * lda + i] = fmaf(pdown, x, aij);
                s = LCG_A * s + LCG_C;
            }
        }
        float *Akk = A + kk * lda + kk;
        orc_getrf_nopiv_f32(b, Akk, lda);
        if (m == 0) break;
        float *L = Akk + b, *U = Akk + (int64_t)b * lda;
        orc_trsm_runn_f32(m, b, Akk, lda, L, lda);
        orc_trsm_llnu_f32(b, m, Akk, lda, U, lda);
        if (qmode == ORC_LU_PRECOMP) {
            for (int64_t j = 0; j < b; ++j) orc_round_f16_array(m, L + j * lda, Lq + j * m);
            for (int64_t j = 0; j < m; ++j) orc_round_f16_array(b, U + j * lda, Uq + j * b);
        } else {
            quant_tiles(ORC_LU_FP8, m, b, b, L, lda, Lq, m);
            quant_tiles(ORC_LU_FP8, b, m, b, U, lda, Uq, b);
        }
        /* T = -sum_l Lq Uq in FP32, then one binary16 rounding per element */
        #pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < m; ++j) memset(T + j * m, 0, sizeof(float) * (size_t)m);
        orc_schur_f32acc(m, m, b, Lq, m, Uq, b, T, m);
        #pragma omp parallel for schedule(static)
        for (int64_t j = 0; j < m; ++j)
            for (int64_t i = 0; i < m; ++i) {
                uint16_t *p = P + (e + j) * n + e + i;
                *p = orc_f32_to_f16(fmaf(pscale, T[j * m + i], h2f(*p)));
            }
        if (Q && (k + 1) % 32 == 0 && e + b < n) {
            const int64_t r1 = e + b;
            #pragma omp parallel for schedule(static)
            for (int64_t j = r1; j < n; ++j)
                for (int64_t i = r1; i < n; ++i) {
                    const float q0 = h2f(Q[j * n + i]);
                    const float t1 = h2f(orc_f32_to_f16(h2f(P[j * n + i]) * h2f(orc_f32_to_f16(1.0f / 32.0f))));
                    const float t2 = h2f(orc_f32_to_f16(q0 + t1));
                    const float t3 = h2f(orc_f32_to_f16(t2 - q0));
                    const float t4 = h2f(orc_f32_to_f16(t1 - t3));
                    Q[j * n + i] = orc_f32_to_f16(t2);
                    P[j * n + i] = orc_f32_to_f16(t4 * 32.0f);
                }
        }
    }
    free(Lq); free(Uq); free(T); free(P); free(Q);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Iterative refinement                                                */
/* ------------------------------------------------------------------ */

/* y += alpha * A x with A regenerated on the fly: off-diagonal a_ij from the
 * LCG, diagonal d_i (otf_gemv_small, panel_gemv.hpp:16-31; otf_gemv_kernel,
 * otf_gemv.cpp:5-131).  FP64, j ascending per row. */
void orc_otf_gemv(int64_t n, uint64_t seed, const double *d, double alpha,
                  const double *x, double *y, int nthreads)
{
    orc_otf_gemv_rows(n, seed, d, alpha, x, y, 0, n, nthreads);
}

/* The same for rows i0 .. i0+cnt-1 only (y indexed by global row). */
void orc_otf_gemv_rows(int64_t n, uint64_t seed, const double *d, double alpha,
                       const double *x, double *y, int64_t i0, int64_t cnt, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    coeff_t jn = coeff_pow(LCG1, (uint64_t)n);
    #pragma omp parallel for schedule(static, 16)
    for (int64_t i = i0; i < i0 + cnt; ++i) {
        uint64_t s = orc_lcg_state(seed, (uint64_t)i);
        double t0 = 0.0, t1 = 0.0;
        int64_t j = 0;
        for (; j + 1 < n; j += 2) {
            double a0 = (j == i) ? d[i] : orc_lcg_value(s);
            s = jn.a * s + jn.c;
            double a1 = (j + 1 == i) ? d[i] : orc_lcg_value(s);
            s = jn.a * s + jn.c;
            t0 = t0 + a0 * x[j];
            t1 = t1 + a1 * x[j + 1];
        }
        if (j < n) {
            double a0 = (j == i) ? d[i] : orc_lcg_value(s);
            t0 = t0 + a0 * x[j];
        }
        y[i] = y[i] + alpha * (t0 + t1);
    }
}

/* tttrsvL, panel_trsv.hpp:115-124 (FAcc=double, FPanel=float) */
void orc_trsv_lower_unit(int64_t n, const float *lu, int64_t lda, double *x)
{
    for (int64_t j = 0; j + 1 < n; ++j) {
        double alpha = x[j];
        for (int64_t i = j + 1; i < n; ++i) x[i] = x[i] - alpha * (double)lu[i + j * lda];
    }
}

/* tttrsvU, panel_trsv.hpp:142-153 */
void orc_trsv_upper(int64_t n, const float *lu, int64_t lda, double *x)
{
    for (int64_t jj = n - 1; jj >= 0; --jj) {
        x[jj] /= (double)lu[jj * lda + jj];
        double alpha = x[jj];
        for (int64_t ii = 0; ii < jj; ++ii) x[ii] = x[ii] - alpha * (double)lu[jj * lda + ii];
    }
}

/* calc_infnorm, panel_norm.hpp:8-19 */
double orc_infnorm(int64_t n, const double *x)
{
    double nrm = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double t = x[i] >= 0.0 ? x[i] : -x[i];
        nrm = nrm >= t ? nrm : t;
    }
    return nrm;
}

/* iterative_refinement, iterative_refinement.hpp:18-57; norms as in
 * main.cpp:273-275 (normb = ||b||, norma = hpl_infnorm = 2 max d). */
int orc_ir(int64_t n, uint64_t seed, const float *lu, int64_t lda, const double *d,
           const double *rhs, double *x, int maxit, double *trace_resid,
           double *trace_harness, double *final_resid, double *final_harness, int nthreads)
{
    double *r = (double *)malloc(sizeof(double) * (size_t)n);
    if (!r) return -2;
    double normb = orc_infnorm(n, rhs);
    double norma = 2.0 * orc_infnorm(n, d);
    for (int64_t i = 0; i < n; ++i) x[i] = rhs[i] / d[i];
    *final_resid = -1.0;
    *final_harness = -1.0;
    int iter;
    for (iter = 0; iter < maxit; ++iter) {
        memcpy(r, rhs, sizeof(double) * (size_t)n);
        double normx = orc_infnorm(n, x);
        orc_otf_gemv(n, seed, d, -1.0, x, r, nthreads);
        double normr = orc_infnorm(n, r);
        double h = normr / (norma * normx + normb) * 1. / ((double)n * DBL_EPSILON / 2);
        if (trace_resid) trace_resid[iter] = normr;
        if (trace_harness) trace_harness[iter] = h;
        if (orc_progress()) {
            fprintf(stderr, "orc_ir: step %d harness %g\n", iter, h);
            fflush(stderr);
        }
        if (h < 16.) {
            *final_resid = normr;
            *final_harness = h;
            free(r);
            return iter + 1;
        }
        orc_trsv_lower_unit(n, lu, lda, r);
        orc_trsv_upper(n, lu, lda, r);
        for (int64_t i = 0; i < n; ++i) x[i] += r[i];
    }
    free(r);
    return iter;
}

/* panel_check, panel_check.hpp:9-65 */
void orc_checksum(int64_t n, const float *lu, int64_t lda, double out[3])
{
    double sd = 0, su = 0, sl = 0;
    for (int64_t j = 0; j < n; ++j) {
        for (int64_t i = 0; i < j; ++i) su += fabs((double)lu[j * lda + i]);
        sd += fabs((double)lu[j * lda + j]);
        for (int64_t i = j + 1; i < n; ++i) sl += fabs((double)lu[j * lda + i]);
    }
    out[0] = sd;
    out[1] = su;
    out[2] = sl;
}
