// internal.hpp — host-side plumbing shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "../../include/hplai.h"

struct hplai_ctx_s {
    int device = 0;
    int num_cu = 0;
    hipStream_t side = nullptr;    // second stream for look-ahead panel work
    hipStream_t bulk = nullptr;    // third stream: the panel's rows below the outer block
    hipEvent_t ev[8] = {};
    // buffers kept by hplai_solve(keep_factors)
    float *A = nullptr;
    int64_t n = 0, lda = 0;
    double *x = nullptr;
    // device address of this device's scratch array (hplai_otf_gemv_f64_noalloc)
    void *scratch = nullptr;
};

namespace hplai {

void set_error(const char *fmt, ...);
int hip_fail(hipError_t e, const char *what);

#define HPLAI_HIP(call)                                              \
    do {                                                             \
        hipError_t _e = (call);                                      \
        if (_e != hipSuccess) return ::hplai::hip_fail(_e, #call);   \
    } while (0)

#define HPLAI_CHECK(cond, ...)                                       \
    do {                                                             \
        if (!(cond)) {                                               \
            ::hplai::set_error(__VA_ARGS__);                         \
            return HPLAI_EINVAL;                                     \
        }                                                            \
    } while (0)

inline hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }
__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- kernel launchers (no argument validation; callers validate) ----
// accumulate != 0: A += generated block (lazy initialisation); P (FP16,
// leading dimension ldp, may be NULL): A += pscale * P as well (the FP16
// trailing storage's partial sums of the block)
int launch_gen_matrix(int64_t n, uint64_t seed, int64_t i0, int64_t j0, int64_t m, int64_t k,
                      const double *d, float *A, int64_t lda, int accumulate, hipStream_t s,
                      const uint16_t *P = nullptr, int64_t ldp = 0, float pscale = 0.0f,
                      const uint16_t *Q = nullptr);
// back-buffer write-back of an m x k block of the FP16 trailing storage:
// Q := fl16(Q + P / 32), P := the rounding error times 32
int launch_backbuffer(int64_t m, int64_t k, uint16_t *P, uint16_t *Q, int64_t ldp, hipStream_t s);
int launch_gen_diag(int64_t n, uint64_t seed, int64_t i0, int64_t cnt, double *d, hipStream_t s);
// Higham's HPL-AI matrix (higham.hip)
int higham_coeffs(int64_t n, double cond, double rho, double *alpha, double *beta);
int launch_gen_higham(double alpha, double beta, int64_t i0, int64_t j0, int64_t m, int64_t k, float *A,
                      int64_t lda, int accumulate, hipStream_t s);
int launch_higham_diag(double alpha, double beta, int64_t i0, int64_t cnt, double *d, hipStream_t s);
int64_t higham_gemv_work_doubles(int64_t n);
int launch_higham_gemv(int64_t n, double alpha, double beta, double scale, const double *x, double *y,
                       double *work, hipStream_t s);
int launch_higham_norm(int64_t n, double alpha, double beta, double *ones, double *rows, double *work,
                       double *out, hipStream_t s);
int launch_gen_rhs(int64_t n, uint64_t seed, int64_t i0, int64_t cnt, double *b, hipStream_t s);

int launch_getrf(int n, float *A, int64_t lda, int *info, hipStream_t s);
// Diagonal-block sub-inverses used by the strip TRSMs: for each 32 x 32
// diagonal sub-block of a factored b x b block, Linv then Uinv (row-major).
int64_t diag_inv_floats(int b);
int launch_diag_inv(int b, const float *LU, int64_t lda, float *inv, hipStream_t s);
// inv: the factored diagonal block's sub-inverses, or NULL (computed on the fly)
int launch_trsm_l(int64_t m, int b, const float *U, int64_t ldu, float *L, int64_t ldl,
                  float *amax, int tb, const float *inv, hipStream_t s);
int launch_trsm_u(int b, int64_t n, const float *Lkk, int64_t ldl, float *U, int64_t ldu,
                  float *amax, int tb, const float *inv, hipStream_t s);
// amax: per-tile maxima of |src| (nullable for FMT_F16 → exponents 0).
// packed: dst in the panel-packed layout (packed_off), rows a multiple of 256 from its base.
int launch_quantize(int fmt, int side, int64_t m, int64_t k, int tb, const float *src, int64_t lds,
                    void *dst, int64_t ldd, const float *amax, int32_t *exps, hipStream_t s,
                    bool packed = false);
int launch_tile_amax(int side, int64_t m, int64_t k, int tb, const float *src, int64_t lds,
                     float *amax, hipStream_t s);
int launch_fill_f32(float *p, int64_t cnt, float v, hipStream_t s);
// packed: Lq and Uq in the panel-packed layout (packed_off) from their bases.
int launch_schur(int fmt, int64_t m, int64_t n, int nslices, int kslice, const void *Lq,
                 int64_t ldl, const int32_t *eL, int64_t ldel, const void *Uq, int64_t ldu,
                 const int32_t *eU, int64_t ldeu, int tb, float *C, int64_t ldc, int num_cu,
                 hipStream_t s, bool packed = false);
// The same update of the FP16 trailing storage: C := fp16(C - cscale * L U),
// C holding cscale times the partial sums (m, n multiples of 128).
int launch_schur_c16(int fmt, int64_t m, int64_t n, int nslices, int kslice, const void *Lq,
                     int64_t ldl, const int32_t *eL, int64_t ldel, const void *Uq, int64_t ldu,
                     const int32_t *eU, int64_t ldeu, int tb, uint16_t *C, int64_t ldc, float cscale,
                     int num_cu, hipStream_t s, bool packed = false);
// true when launch_schur runs the streamed 256 x 256 kernel for this shape
bool schur_large_tiles(int fmt, int64_t m, int64_t n, int tb, int num_cu);

// the reference's column-major gemmschur seam (gemmschur.hip): both operands
// staged K-contiguous with K padded to 64 in `scratch`
// (gemmschur_scratch_bytes), then C(n x m) -= A B with FP32 accumulation
int64_t gemmschur_scratch_bytes(int64_t n, int64_t m, int64_t k);
int launch_gemmschur_f16(int64_t n, int64_t m, int64_t k, const uint16_t *a, int64_t lda, const uint16_t *b,
                         int64_t ldb, float *c, int64_t ldc, uint16_t *scratch, int num_cu, hipStream_t s);

int launch_otf_gemv(int64_t n, uint64_t seed, const double *d, double alpha, const double *x,
                    double *y, double *work, hipStream_t s);
int64_t otf_gemv_work_doubles(int64_t n);
// address of the no-allocation scratch array on `device` (current), looked up once per device
int noalloc_setup(int device, void **out);
// triangular solves with the FP32 factors of LU, FP64 vectors; nrhs columns
// of x with stride ldx
bool tri_inverse_ok(int b);
int64_t tri_inverse_offset(int b);
int launch_tri_trsv(bool upper, int b, const float *tri, const double *xin, double *xout, hipStream_t s);
int launch_trsv_lower_inv(int64_t n, int bs, const float *LU, int64_t lda, double *x, const float *inv,
                          int64_t inv_stride, double *xsc, hipStream_t s);
int launch_trsv_upper_inv(int64_t n, int bs, const float *LU, int64_t lda, double *x, const float *inv,
                          int64_t inv_stride, double *xsc, hipStream_t s);
int launch_trsv_lower(int64_t n, int bs, const float *LU, int64_t lda, double *x, hipStream_t s,
                      int nrhs = 1, int64_t ldx = 0);   // x := L^-1 x
int launch_trsv_upper(int64_t n, int bs, const float *LU, int64_t lda, double *x, hipStream_t s,
                      int nrhs = 1, int64_t ldx = 0);   // x := U^-1 x
int launch_trsv_upper_t(int64_t n, int bs, const float *LU, int64_t lda, double *x, hipStream_t s,
                        int nrhs = 1, int64_t ldx = 0);   // x := U^-T x
int launch_trsv_lower_t(int64_t n, int bs, const float *LU, int64_t lda, double *x, hipStream_t s,
                        int nrhs = 1, int64_t ldx = 0);   // x := L^-T x
int launch_norm_inf(int64_t n, const double *x, double *out, hipStream_t s);
int launch_vec_div(int64_t n, const double *num, const double *den, double *out, hipStream_t s);
int launch_vec_add(int64_t n, const double *x, double *y, hipStream_t s);
int launch_checksum(int64_t n, const float *LU, int64_t lda, double *out3, hipStream_t s);

}  // namespace hplai
