// capi.hip — the exported C ABI (include/hplai.h): context, error reporting
// and argument-checked wrappers over the kernel launchers.
#include <cstdarg>
#include <cstdio>
#include <new>

#include "internal.hpp"

namespace hplai {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_fail(hipError_t e, const char *what)
{
    set_error("%s: %s", what, hipGetErrorString(e));
    // the runtime keeps the last error until it is read: clear it, so that a
    // failed allocation is not reported again by the next call's launch check
    (void)hipGetLastError();
    return e == hipErrorOutOfMemory ? HPLAI_ENOMEM : HPLAI_EHIP;
}

}  // namespace hplai

using namespace hplai;

#define CTX_CHECK(ctx)                                                  \
    do {                                                                \
        if ((ctx) == nullptr) {                                         \
            set_error("null hplai_ctx");                                \
            return HPLAI_EINVAL;                                        \
        }                                                               \
        HPLAI_HIP(hipSetDevice((ctx)->device));                         \
    } while (0)

extern "C" {

int hplai_version(void) { return 100; }

const char *hplai_last_error(void) { return g_err; }

int hplai_create(int device, hplai_ctx *out)
{
    HPLAI_CHECK(out != nullptr, "hplai_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    HPLAI_HIP(hipGetDeviceCount(&ndev));
    HPLAI_CHECK(device >= 0 && device < ndev, "hplai_create: device %d of %d", device, ndev);
    HPLAI_HIP(hipSetDevice(device));
    hplai_ctx c = new (std::nothrow) hplai_ctx_s();
    if (c == nullptr) {
        set_error("hplai_create: host allocation failed");
        return HPLAI_ENOMEM;
    }
    c->device = device;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        delete c;
        return hip_fail(e, "hipGetDeviceProperties");
    }
    c->num_cu = prop.multiProcessorCount;
    e = hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->bulk, hipStreamNonBlocking);
    for (int i = 0; e == hipSuccess && i < 8; ++i)
        e = hipEventCreateWithFlags(&c->ev[i], hipEventDisableTiming);
    if (e != hipSuccess) {
        hplai_destroy(c);
        return hip_fail(e, "hplai_create streams/events");
    }
    // the device address of this device's no-allocation scratch array
    const int rc = noalloc_setup(device, &c->scratch);
    if (rc != HPLAI_OK) {
        hplai_destroy(c);
        return rc;
    }
    *out = c;
    return HPLAI_OK;
}

int hplai_destroy(hplai_ctx ctx)
{
    if (ctx == nullptr) return HPLAI_OK;
    (void)hipSetDevice(ctx->device);
    hplai_release(ctx);
    for (int i = 0; i < 8; ++i)
        if (ctx->ev[i]) (void)hipEventDestroy(ctx->ev[i]);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
    if (ctx->bulk) (void)hipStreamDestroy(ctx->bulk);
    delete ctx;
    return HPLAI_OK;
}

int hplai_gen_matrix_f32(hplai_ctx ctx, int64_t n, uint64_t seed, int64_t i0, int64_t j0, int64_t m,
                         int64_t k, const double *d, float *A, int64_t lda, void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(n > 0 && i0 >= 0 && j0 >= 0 && m >= 0 && k >= 0 && i0 + m <= n && j0 + k <= n,
                "gen_matrix: block (%lld,%lld)+(%lld,%lld) outside n=%lld", (long long)i0,
                (long long)j0, (long long)m, (long long)k, (long long)n);
    HPLAI_CHECK(A != nullptr && lda >= m, "gen_matrix: A NULL or lda < m");
    return launch_gen_matrix(n, seed, i0, j0, m, k, d, A, lda, 0, as_stream(stream));
}

int hplai_gen_diag_f64(hplai_ctx ctx, int64_t n, uint64_t seed, int64_t i0, int64_t cnt, double *d,
                       void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(n > 0 && i0 >= 0 && cnt >= 0 && i0 + cnt <= n && d != nullptr,
                "gen_diag: bad range or NULL output");
    return launch_gen_diag(n, seed, i0, cnt, d, as_stream(stream));
}

int hplai_gen_rhs_f64(hplai_ctx ctx, int64_t n, uint64_t seed, int64_t i0, int64_t cnt, double *b,
                      void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(n > 0 && i0 >= 0 && cnt >= 0 && i0 + cnt <= n && b != nullptr,
                "gen_rhs: bad range or NULL output");
    return launch_gen_rhs(n, seed, i0, cnt, b, as_stream(stream));
}

int hplai_getrf_nopiv_f32(hplai_ctx ctx, int n, float *A, int64_t lda, int *info_dev, void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(n >= 0 && n <= 1024, "getrf_nopiv: n=%d outside [0,1024]", n);
    HPLAI_CHECK(A != nullptr && lda >= n, "getrf_nopiv: A NULL or lda < n");
    return launch_getrf(n, A, lda, info_dev, as_stream(stream));
}

int hplai_trsm_lpanel_f32(hplai_ctx ctx, int64_t m, int b, const float *Ukk, int64_t ldu, float *L,
                          int64_t ldl, void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(m >= 0 && b >= 0 && b <= 1024, "trsm_lpanel: bad shape");
    HPLAI_CHECK(Ukk != nullptr && L != nullptr && ldu >= b && ldl >= m, "trsm_lpanel: bad pointers/ld");
    return launch_trsm_l(m, b, Ukk, ldu, L, ldl, nullptr, 1, nullptr, as_stream(stream));
}

int hplai_trsm_upanel_f32(hplai_ctx ctx, int b, int64_t n, const float *Lkk, int64_t ldl, float *U,
                          int64_t ldu, void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(n >= 0 && b >= 0 && b <= 1024, "trsm_upanel: bad shape");
    HPLAI_CHECK(Lkk != nullptr && U != nullptr && ldl >= b && ldu >= b, "trsm_upanel: bad pointers/ld");
    return launch_trsm_u(b, n, Lkk, ldl, U, ldu, nullptr, 1, nullptr, as_stream(stream));
}

int hplai_quantize(hplai_ctx ctx, int fmt, int side, int64_t m, int64_t k, int tb, const float *src,
                   int64_t lds, void *dst, int64_t ldd, int32_t *exps, void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(fmt == HPLAI_FMT_F16 || fmt == HPLAI_FMT_F16_TILE || fmt == HPLAI_FMT_E4M3,
                "quantize: unknown format %d", fmt);
    HPLAI_CHECK(side == 0 || side == 1, "quantize: side must be 0 (L) or 1 (U)");
    HPLAI_CHECK(m >= 0 && k >= 0 && ldd >= k && src != nullptr && dst != nullptr,
                "quantize: bad shape/pointers");
    HPLAI_CHECK(lds >= (side == 0 ? m : k), "quantize: lds too small");
    if (m == 0 || k == 0) return HPLAI_OK;
    float *amax = nullptr;
    if (fmt != HPLAI_FMT_F16) {
        HPLAI_CHECK(exps != nullptr && tb > 0, "quantize: exps and tb required for tile formats");
        const int64_t ntiles = ceil_div(m, tb);
        hipStream_t s = as_stream(stream);
        HPLAI_HIP(hipMallocAsync(reinterpret_cast<void **>(&amax), sizeof(float) * ntiles, s));
        int rc = launch_fill_f32(amax, ntiles, 0.0f, s);
        if (rc == HPLAI_OK) rc = launch_tile_amax(side, m, k, tb, src, lds, amax, s);
        if (rc == HPLAI_OK)
            rc = launch_quantize(fmt, side, m, k, tb, src, lds, dst, ldd, amax, exps, s);
        HPLAI_HIP(hipFreeAsync(amax, s));
        return rc;
    }
    return launch_quantize(fmt, side, m, k, tb > 0 ? tb : 1, src, lds, dst, ldd, nullptr, nullptr,
                           as_stream(stream));
}

int hplai_schur_update(hplai_ctx ctx, int fmt, int64_t m, int64_t n, int nslices, int kslice,
                       const void *Lq, int64_t ldl, const int32_t *eL, int64_t ldel, const void *Uq,
                       int64_t ldu, const int32_t *eU, int64_t ldeu, int tb, float *C, int64_t ldc,
                       void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(m >= 0 && n >= 0 && nslices >= 0 && Lq != nullptr && Uq != nullptr && C != nullptr,
                "schur_update: bad shape/pointers");
    return launch_schur(fmt, m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu, eU, ldeu, tb, C, ldc,
                        ctx->num_cu, as_stream(stream));
}

int hplai_gemmschur_f16(hplai_ctx ctx, int64_t n, int64_t m, int64_t k, const uint16_t *a, int64_t lda,
                        const uint16_t *b, int64_t ldb, float *c, int64_t ldc, void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(n >= 0 && m >= 0 && k >= 0, "gemmschur: negative size (n=%lld m=%lld k=%lld)", (long long)n,
                (long long)m, (long long)k);
    if (n == 0 || m == 0 || k == 0) return HPLAI_OK;
    HPLAI_CHECK(a != nullptr && b != nullptr && c != nullptr, "gemmschur: NULL operand");
    HPLAI_CHECK(lda >= n && ldb >= k && ldc >= n, "gemmschur: need lda >= n, ldb >= k, ldc >= n");
    HPLAI_CHECK(k <= (int64_t(1) << 24), "gemmschur: k=%lld too large", (long long)k);
    hipStream_t s = as_stream(stream);
    void *scratch = nullptr;
    HPLAI_HIP(hipMallocAsync(&scratch, static_cast<size_t>(gemmschur_scratch_bytes(n, m, k)), s));
    const int rc = launch_gemmschur_f16(n, m, k, a, lda, b, ldb, c, ldc, static_cast<uint16_t *>(scratch),
                                        ctx->num_cu, s);
    HPLAI_HIP(hipFreeAsync(scratch, s));
    return rc;
}

int hplai_schur_update_f16c(hplai_ctx ctx, int fmt, int64_t m, int64_t n, int nslices, int kslice,
                            const void *Lq, int64_t ldl, const int32_t *eL, int64_t ldel, const void *Uq,
                            int64_t ldu, const int32_t *eU, int64_t ldeu, int tb, uint16_t *C, int64_t ldc,
                            float cscale, void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(m >= 0 && n >= 0 && nslices >= 0 && Lq != nullptr && Uq != nullptr && C != nullptr,
                "schur_update_f16c: bad shape/pointers");
    return launch_schur_c16(fmt, m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu, eU, ldeu, tb, C, ldc, cscale,
                            ctx->num_cu, as_stream(stream));
}

int hplai_otf_gemv_f64(hplai_ctx ctx, int64_t n, uint64_t seed, const double *d, double alpha,
                       const double *x, double *y, void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(n > 0 && d != nullptr && x != nullptr && y != nullptr, "otf_gemv: bad args");
    hipStream_t s = as_stream(stream);
    double *work = nullptr;
    HPLAI_HIP(hipMallocAsync(reinterpret_cast<void **>(&work), sizeof(double) * otf_gemv_work_doubles(n), s));
    int rc = launch_otf_gemv(n, seed, d, alpha, x, y, work, s);
    HPLAI_HIP(hipFreeAsync(work, s));
    return rc;
}

int hplai_trsv_lower_f32(hplai_ctx ctx, int64_t n, int bs, const float *LU, int64_t lda, double *x,
                         void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(n >= 0 && bs > 0 && bs <= 1024 && LU != nullptr && x != nullptr && lda >= n,
                "trsv_lower: bad args");
    return launch_trsv_lower(n, bs, LU, lda, x, as_stream(stream));
}

int hplai_trsv_upper_f32(hplai_ctx ctx, int64_t n, int bs, const float *LU, int64_t lda, double *x,
                         void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(n >= 0 && bs > 0 && bs <= 1024 && LU != nullptr && x != nullptr && lda >= n,
                "trsv_upper: bad args");
    return launch_trsv_upper(n, bs, LU, lda, x, as_stream(stream));
}

int hplai_norm_inf_f64(hplai_ctx ctx, int64_t n, const double *x, double *out_dev, void *stream)
{
    CTX_CHECK(ctx);
    HPLAI_CHECK(n >= 0 && x != nullptr && out_dev != nullptr, "norm_inf: bad args");
    return launch_norm_inf(n, x, out_dev, as_stream(stream));
}

}  // extern "C"
