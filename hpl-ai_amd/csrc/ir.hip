// ir.hip — FP64 iterative-refinement kernels on gfx950: the matrix-free
// residual GEMV (A regenerated from the LCG), the blocked triangular solves
// with FP32 factors and FP64 vectors, norms, vector updates and the factor
// checksum.
#include "device_common.hpp"
#include "internal.hpp"
#include "trsv_block.hpp"

namespace hplai {

// ---------------------------------------------------------------------------
// otf_gemv (panel_gemv.hpp:57-83, otf_gemv.cpp:5-131).
// Pass 1: partial[c][i] = sum_{j in chunk c, j != i} (double)(int64)state_ij * x_j
// (raw integers, like otf_gemv_kernel; the 2^-64-style scale is applied once).
// Pass 2: y_i += alpha * (scale * sum_c partial[c][i] + d_i x_i), chunks in order.
// ---------------------------------------------------------------------------
constexpr int kGemvRows = 256;
constexpr int kGemvChunk = 2048;

__global__ __launch_bounds__(kGemvRows) void otf_gemv_partial_kernel(
    int64_t n, uint64_t x1, int64_t row_blocks, LcgCoeff jn, LcgCoeff jn2,
    const double *__restrict__ x, double *__restrict__ partial)
{
    __shared__ double xs[kGemvChunk];
    const int64_t rb = blockIdx.x % row_blocks;
    const int64_t cb = blockIdx.x / row_blocks;
    const int64_t j0 = cb * kGemvChunk;
    const int64_t j1 = j0 + kGemvChunk < n ? j0 + kGemvChunk : n;
    {
        double v[kGemvChunk / kGemvRows];   // all loads in flight together
#pragma unroll
        for (int u = 0; u < kGemvChunk / kGemvRows; ++u) v[u] = x[min(j0 + u * kGemvRows + threadIdx.x, j1 - 1)];
#pragma unroll
        for (int u = 0; u < kGemvChunk / kGemvRows; ++u) xs[u * kGemvRows + threadIdx.x] = v[u];
    }
    __syncthreads();
    const int64_t i = rb * kGemvRows + threadIdx.x;
    if (i >= n) return;
    uint64_t s0 = lcg_jump(x1, static_cast<uint64_t>(i) + static_cast<uint64_t>(n) * static_cast<uint64_t>(j0));
    uint64_t s1 = jn.a * s0 + jn.c;
    double acc0 = 0.0, acc1 = 0.0;
    const int64_t len = j1 - j0;
    const int64_t idiag = i - j0;   // local column of the diagonal, if in this chunk
    int64_t t = 0;
    // pairs (t, t + 1) with no diagonal element up to te; the one pair that
    // holds the diagonal (when it is in this chunk) is masked on its own
    auto pairs = [&](int64_t te) {
        for (; t + 1 < te; t += 2) {
            const double a0 = static_cast<double>(static_cast<int64_t>(s0));
            const double a1 = static_cast<double>(static_cast<int64_t>(s1));
            acc0 = fma(a0, xs[t], acc0);
            acc1 = fma(a1, xs[t + 1], acc1);
            s0 = jn2.a * s0 + jn2.c;
            s1 = jn2.a * s1 + jn2.c;
        }
    };
    const int64_t pd = (idiag >= 0 && idiag < len) ? (idiag & ~static_cast<int64_t>(1)) : len;
    pairs(pd);
    if (t == pd && t + 1 < len) {
        double a0 = static_cast<double>(static_cast<int64_t>(s0));
        double a1 = static_cast<double>(static_cast<int64_t>(s1));
        a0 = (t == idiag) ? 0.0 : a0;
        a1 = (t + 1 == idiag) ? 0.0 : a1;
        acc0 = fma(a0, xs[t], acc0);
        acc1 = fma(a1, xs[t + 1], acc1);
        s0 = jn2.a * s0 + jn2.c;
        s1 = jn2.a * s1 + jn2.c;
        t += 2;
    }
    pairs(len);
    if (t < len) {
        double a0 = static_cast<double>(static_cast<int64_t>(s0));
        a0 = (t == idiag) ? 0.0 : a0;
        acc0 = fma(a0, xs[t], acc0);
    }
    partial[cb * n + i] = acc0 + acc1;
}

__global__ void otf_gemv_reduce_kernel(int64_t n, int64_t nchunks, double alpha,
                                       const double *__restrict__ d, const double *__restrict__ x,
                                       const double *__restrict__ partial, double *__restrict__ y)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int64_t c = 0; c < nchunks; ++c) s += partial[c * n + i];
    const double ax = fma(s, 0x1.fffffffffffffP-65, d[i] * x[i]);
    y[i] = fma(alpha, ax, y[i]);
}

int64_t otf_gemv_work_doubles(int64_t n) { return ceil_div(n, kGemvChunk) * n; }

int launch_otf_gemv(int64_t n, uint64_t seed, const double *d, double alpha, const double *x,
                    double *y, double *work, hipStream_t s)
{
    if (n <= 0) return HPLAI_OK;
    const int64_t row_blocks = ceil_div(n, kGemvRows);
    const int64_t nchunks = ceil_div(n, kGemvChunk);
    const LcgCoeff jn = lcg_pow_host(static_cast<uint64_t>(n));
    const LcgCoeff jn2 = lcg_mul(jn, jn);
    hipLaunchKernelGGL(otf_gemv_partial_kernel, dim3(static_cast<unsigned>(row_blocks * nchunks)),
                       dim3(kGemvRows), 0, s, n, lcg_init(seed), row_blocks, jn, jn2, x, work);
    HPLAI_HIP(hipGetLastError());
    hipLaunchKernelGGL(otf_gemv_reduce_kernel, dim3(static_cast<unsigned>(ceil_div(n, 256))),
                       dim3(256), 0, s, n, nchunks, alpha, d, x, work, y);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

// ---------------------------------------------------------------------------
// Triangular solves (panel_trsvL / panel_trsvU, panel_trsv.hpp:310-758, with
// tttrsvL/U, 115-153, and ttgemv, 10-45): FP32 factor entries widened to FP64.
// Per block of bs rows: (1) one workgroup solves the diagonal block
// (trsv_block, trsv_block.hpp), (2) a GEMV updates the rest of the vector
// with the block's column panel.
// ---------------------------------------------------------------------------
constexpr int kTrsvMax = 1024;

// Right-hand side blockIdx.y at x + y * ldx.
template <int RPT, bool UPPER, bool UNIT, bool TRANS>
__global__ __launch_bounds__(256) void trsv_diag_kernel(int bs, const float *__restrict__ T, int64_t lda,
                                                        double *__restrict__ x, int64_t ldx)
{
    __shared__ double xs[32];
    trsv_block<RPT, UPPER, UNIT, TRANS>(bs, T, lda, x + static_cast<int64_t>(blockIdx.y) * ldx, xs);
}

template <bool UPPER, bool UNIT, bool TRANS>
static void launch_trsv_diag(int w, const float *T, int64_t lda, double *x, int64_t ldx, int nrhs,
                             hipStream_t s)
{
    const dim3 grid(1, static_cast<unsigned>(nrhs));
    if (w <= 256)
        hipLaunchKernelGGL((trsv_diag_kernel<1, UPPER, UNIT, TRANS>), grid, dim3(256), 0, s, w, T, lda, x, ldx);
    else if (w <= 512)
        hipLaunchKernelGGL((trsv_diag_kernel<2, UPPER, UNIT, TRANS>), grid, dim3(256), 0, s, w, T, lda, x, ldx);
    else
        hipLaunchKernelGGL((trsv_diag_kernel<4, UPPER, UNIT, TRANS>), grid, dim3(256), 0, s, w, T, lda, x, ldx);
}

// y(0:m) -= P(m x w) * xb(0:w), P column-major (ttgemv, panel_trsv.hpp:10-45).
// Workgroup = 64 rows x 4 column groups; the 4 partial sums of a row are added
// in a fixed order through LDS.  Right-hand side blockIdx.y at offset y * ldx.
constexpr int kGemvRowsT = 64;

// With xcopy != nullptr (one right-hand side), workgroup 0 also stores xb
// into xcopy(0:w): the solved segment goes from its scratch copy to x while
// the update runs.
__global__ __launch_bounds__(256) void trsv_gemv_kernel(int64_t m, int w, const float *__restrict__ P,
                                                        int64_t lda, const double *__restrict__ xb,
                                                        double *__restrict__ y, int64_t ldx,
                                                        double *__restrict__ xcopy)
{
    __shared__ double xs[kTrsvMax];
    __shared__ double part[4][kGemvRowsT];
    xb += static_cast<int64_t>(blockIdx.y) * ldx;
    y += static_cast<int64_t>(blockIdx.y) * ldx;
    for (int c = threadIdx.x; c < w; c += 256) {
        xs[c] = xb[c];
        if (xcopy != nullptr && blockIdx.x == 0) xcopy[c] = xb[c];
    }
    __syncthreads();
    const int lr = threadIdx.x % kGemvRowsT, g = threadIdx.x / kGemvRowsT;
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kGemvRowsT + lr;
    const int cw = (w + 3) / 4, c0 = g * cw, c1 = c0 + cw < w ? c0 + cw : w;
    double acc0 = 0.0, acc1 = 0.0;
    if (r < m) {
        const float *pr = P + r;
        int c = c0;
        for (; c + 7 < c1; c += 8) {
            float v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = pr[static_cast<int64_t>(c + q) * lda];
#pragma unroll
            for (int q = 0; q < 8; q += 2) {
                acc0 = fma(static_cast<double>(v[q]), xs[c + q], acc0);
                acc1 = fma(static_cast<double>(v[q + 1]), xs[c + q + 1], acc1);
            }
        }
        for (; c < c1; ++c) acc0 = fma(static_cast<double>(pr[static_cast<int64_t>(c) * lda]), xs[c], acc0);
    }
    part[g][lr] = acc0 + acc1;
    __syncthreads();
    if (g == 0 && r < m) y[r] -= ((part[0][lr] + part[1][lr]) + part[2][lr]) + part[3][lr];
}

// Transposed panel: y(r) -= sum_c P[r*lda + c] * xb(c), r < m, c < w — each
// row is contiguous; one wave per row at a time, lanes split the columns and
// the partial sums are combined by a fixed shuffle tree.
constexpr int kGemvTRows = 16;

__global__ __launch_bounds__(256) void trsv_gemv_t_kernel(int64_t m, int w, const float *__restrict__ P,
                                                          int64_t lda, const double *__restrict__ xb,
                                                          double *__restrict__ y, int64_t ldx)
{
    __shared__ double xs[kTrsvMax];
    xb += static_cast<int64_t>(blockIdx.y) * ldx;
    y += static_cast<int64_t>(blockIdx.y) * ldx;
    for (int c = threadIdx.x; c < w; c += 256) xs[c] = xb[c];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int rr = wave; rr < kGemvTRows; rr += 4) {
        const int64_t r = static_cast<int64_t>(blockIdx.x) * kGemvTRows + rr;
        if (r >= m) break;
        const float *pr = P + r * lda;
        double acc = 0.0;
        // columns lane, lane + 64, ... in that order; 8 loads issued together
        int c = lane;
        for (; c + 64 * 7 < w; c += 64 * 8) {
            float v[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) v[q] = pr[c + 64 * q];
#pragma unroll
            for (int q = 0; q < 8; ++q) acc = fma(static_cast<double>(v[q]), xs[c + 64 * q], acc);
        }
        for (; c < w; c += 64) acc = fma(static_cast<double>(pr[c]), xs[c], acc);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
        if (lane == 0) y[r] -= acc;
    }
}

static void launch_gemv(bool trans, int64_t m, int w, const float *P, int64_t lda, const double *xb, double *y,
                        int64_t ldx, int nrhs, hipStream_t s, double *xcopy = nullptr)
{
    if (trans)
        hipLaunchKernelGGL(trsv_gemv_t_kernel, dim3(static_cast<unsigned>(ceil_div(m, kGemvTRows)), nrhs),
                           dim3(256), 0, s, m, w, P, lda, xb, y, ldx);
    else
        hipLaunchKernelGGL(trsv_gemv_kernel, dim3(static_cast<unsigned>(ceil_div(m, kGemvRowsT)), nrhs),
                           dim3(256), 0, s, m, w, P, lda, xb, y, ldx, xcopy);
}

// The same block solves with the diagonal blocks applied through the inverses
// the LU stored (tri_inv_kernel layout, one set per block at inv + q
// inv_stride + tri_inverse_offset(bs)): x_q := Tinv_q r_q into the scratch
// segment xsc (bs doubles), then the update of the other rows from xsc, whose
// workgroup 0 also moves x_q into x.  One right-hand side; n % bs == 0 and
// tri_inverse_ok(bs).
int launch_trsv_lower_inv(int64_t n, int bs, const float *LU, int64_t lda, double *x, const float *inv,
                          int64_t inv_stride, double *xsc, hipStream_t s)
{
    HPLAI_CHECK(tri_inverse_ok(bs) && n % bs == 0, "trsv: no stored inverses for bs=%d", bs);
    const int64_t off = tri_inverse_offset(bs);
    for (int64_t kb = 0; kb < n; kb += bs) {
        int rc = launch_tri_trsv(false, bs, inv + (kb / bs) * inv_stride + off, x + kb, xsc, s);
        if (rc) return rc;
        const int64_t m = n - kb - bs;
        if (m > 0) {
            launch_gemv(false, m, bs, LU + kb * lda + kb + bs, lda, xsc, x + kb + bs, 0, 1, s, x + kb);
            HPLAI_HIP(hipGetLastError());
        } else {
            HPLAI_HIP(hipMemcpyAsync(x + kb, xsc, sizeof(double) * bs, hipMemcpyDeviceToDevice, s));
        }
    }
    return HPLAI_OK;
}

int launch_trsv_upper_inv(int64_t n, int bs, const float *LU, int64_t lda, double *x, const float *inv,
                          int64_t inv_stride, double *xsc, hipStream_t s)
{
    HPLAI_CHECK(tri_inverse_ok(bs) && n % bs == 0, "trsv: no stored inverses for bs=%d", bs);
    const int64_t off = tri_inverse_offset(bs);
    for (int64_t kb = n - bs; kb >= 0; kb -= bs) {
        int rc = launch_tri_trsv(true, bs, inv + (kb / bs) * inv_stride + off, x + kb, xsc, s);
        if (rc) return rc;
        if (kb > 0) {
            launch_gemv(false, kb, bs, LU + kb * lda, lda, xsc, x, 0, 1, s, x + kb);
            HPLAI_HIP(hipGetLastError());
        } else {
            HPLAI_HIP(hipMemcpyAsync(x + kb, xsc, sizeof(double) * bs, hipMemcpyDeviceToDevice, s));
        }
    }
    return HPLAI_OK;
}

// x := L^-1 x (unit lower L of LU), nrhs columns of x with stride ldx
int launch_trsv_lower(int64_t n, int bs, const float *LU, int64_t lda, double *x, hipStream_t s, int nrhs,
                      int64_t ldx)
{
    for (int64_t kb = 0; kb < n; kb += bs) {
        const int w = n - kb < bs ? static_cast<int>(n - kb) : bs;
        const float *Lkk = LU + kb * lda + kb;
        launch_trsv_diag<false, true, false>(w, Lkk, lda, x + kb, ldx, nrhs, s);
        HPLAI_HIP(hipGetLastError());
        const int64_t m = n - kb - w;
        if (m > 0) {
            launch_gemv(false, m, w, Lkk + w, lda, x + kb, x + kb + w, ldx, nrhs, s);
            HPLAI_HIP(hipGetLastError());
        }
    }
    return HPLAI_OK;
}

// x := U^-1 x (upper U of LU)
int launch_trsv_upper(int64_t n, int bs, const float *LU, int64_t lda, double *x, hipStream_t s, int nrhs,
                      int64_t ldx)
{
    const int64_t nblk = ceil_div(n, bs);
    for (int64_t q = nblk - 1; q >= 0; --q) {
        const int64_t kb = q * bs;
        const int w = n - kb < bs ? static_cast<int>(n - kb) : bs;
        const float *Ukk = LU + kb * lda + kb;
        launch_trsv_diag<true, false, false>(w, Ukk, lda, x + kb, ldx, nrhs, s);
        HPLAI_HIP(hipGetLastError());
        if (kb > 0) {
            launch_gemv(false, kb, w, LU + kb * lda, lda, x + kb, x, ldx, nrhs, s);
            HPLAI_HIP(hipGetLastError());
        }
    }
    return HPLAI_OK;
}

// x := U^-T x: U^T is lower with U's diagonal; element (r, c) = LU[r*lda + c]
int launch_trsv_upper_t(int64_t n, int bs, const float *LU, int64_t lda, double *x, hipStream_t s, int nrhs,
                        int64_t ldx)
{
    for (int64_t kb = 0; kb < n; kb += bs) {
        const int w = n - kb < bs ? static_cast<int>(n - kb) : bs;
        const float *Tkk = LU + kb * lda + kb;
        launch_trsv_diag<false, false, true>(w, Tkk, lda, x + kb, ldx, nrhs, s);
        HPLAI_HIP(hipGetLastError());
        const int64_t m = n - kb - w;
        if (m > 0) {
            // rows kb + w + r', columns kb + c: LU[(kb + w + r') * lda + kb + c]
            launch_gemv(true, m, w, LU + (kb + w) * lda + kb, lda, x + kb, x + kb + w, ldx, nrhs, s);
            HPLAI_HIP(hipGetLastError());
        }
    }
    return HPLAI_OK;
}

// x := L^-T x: L^T is unit upper; element (r, c) = LU[r*lda + c]
int launch_trsv_lower_t(int64_t n, int bs, const float *LU, int64_t lda, double *x, hipStream_t s, int nrhs,
                        int64_t ldx)
{
    const int64_t nblk = ceil_div(n, bs);
    for (int64_t q = nblk - 1; q >= 0; --q) {
        const int64_t kb = q * bs;
        const int w = n - kb < bs ? static_cast<int>(n - kb) : bs;
        const float *Tkk = LU + kb * lda + kb;
        launch_trsv_diag<true, true, true>(w, Tkk, lda, x + kb, ldx, nrhs, s);
        HPLAI_HIP(hipGetLastError());
        if (kb > 0) {
            // rows r < kb, columns kb + c: LU[r * lda + kb + c]
            launch_gemv(true, kb, w, LU + kb, lda, x + kb, x, ldx, nrhs, s);
            HPLAI_HIP(hipGetLastError());
        }
    }
    return HPLAI_OK;
}

// ---------------------------------------------------------------------------
// norms and vector ops (panel_norm.hpp:8-41, panel.hpp:277-330)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void norm_inf_kernel(int64_t n, const double *__restrict__ x,
                                                        double *__restrict__ out)
{
    __shared__ double red[1024];
    double m = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) m = fmax(m, fabs(x[i]));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int off = 512; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + off]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

__global__ void vec_div_kernel(int64_t n, const double *num, const double *den, double *out)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = num[i] / den[i];
}

__global__ void vec_add_kernel(int64_t n, const double *x, double *y)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) y[i] += x[i];
}

int launch_norm_inf(int64_t n, const double *x, double *out, hipStream_t s)
{
    hipLaunchKernelGGL(norm_inf_kernel, dim3(1), dim3(1024), 0, s, n, x, out);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

int launch_vec_div(int64_t n, const double *num, const double *den, double *out, hipStream_t s)
{
    if (n <= 0) return HPLAI_OK;
    hipLaunchKernelGGL(vec_div_kernel, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0, s,
                       n, num, den, out);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

int launch_vec_add(int64_t n, const double *x, double *y, hipStream_t s)
{
    if (n <= 0) return HPLAI_OK;
    hipLaunchKernelGGL(vec_add_kernel, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0, s,
                       n, x, y);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

// ---------------------------------------------------------------------------
// panel_check fingerprint (panel_check.hpp:9-65): sum |diag|, |strict U|,
// |strict L| in FP64.  Column sums then a fixed-order final pass, so the
// result is deterministic.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void checksum_col_kernel(int64_t n, const float *__restrict__ LU,
                                                           int64_t lda, double *__restrict__ cols)
{
    __shared__ double red[3][256];
    const int64_t j = blockIdx.x;
    const float *c = LU + j * lda;
    double su = 0.0, sl = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double v = fabs(static_cast<double>(c[i]));
        if (i < j) su += v;
        else if (i > j) sl += v;
    }
    red[0][threadIdx.x] = su;
    red[1][threadIdx.x] = sl;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (threadIdx.x < off) {
            red[0][threadIdx.x] += red[0][threadIdx.x + off];
            red[1][threadIdx.x] += red[1][threadIdx.x + off];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        cols[j] = fabs(static_cast<double>(c[j]));
        cols[n + j] = red[0][0];
        cols[2 * n + j] = red[1][0];
    }
}

__global__ __launch_bounds__(1024) void checksum_final_kernel(int64_t n, const double *__restrict__ cols,
                                                              double *__restrict__ out)
{
    __shared__ double red[1024];
    for (int q = 0; q < 3; ++q) {
        double s = 0.0;
        for (int64_t j = threadIdx.x; j < n; j += 1024) s += cols[q * n + j];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int off = 512; off > 0; off >>= 1) {
            if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = red[0];
        __syncthreads();
    }
}

int launch_checksum(int64_t n, const float *LU, int64_t lda, double *out3, hipStream_t s)
{
    double *cols = nullptr;
    HPLAI_HIP(hipMallocAsync(reinterpret_cast<void **>(&cols), sizeof(double) * 3 * n, s));
    hipLaunchKernelGGL(checksum_col_kernel, dim3(static_cast<unsigned>(n)), dim3(256), 0, s, n, LU, lda,
                       cols);
    hipLaunchKernelGGL(checksum_final_kernel, dim3(1), dim3(1024), 0, s, n, cols, out3);
    HPLAI_HIP(hipGetLastError());
    HPLAI_HIP(hipFreeAsync(cols, s));
    return HPLAI_OK;
}

}  // namespace hplai
