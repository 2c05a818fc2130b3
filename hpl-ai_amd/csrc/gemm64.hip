// gemm64.hip — FP64 GEMM on CDNA4's FP64 matrix cores for the differentiable
// form (no reference counterpart): the residuals R = B - op(A) X of the
// forward and transposed refinements and the gradient dA = -Lambda X^T.
//
//   C := alpha op(A) op(B) + beta C   (column-major; beta = 0: C is not read)
//
// 64 x 64 tiles of C per workgroup (4 waves, 2 x 2, each 32 x 32 as 2 x 2
// blocks of v_mfma_f64_16x16x4_f64), K in chunks of 16 staged through LDS
// with the next chunk's global loads in flight during the current chunk's
// MFMAs.  When the tiles cannot fill the device (the skinny residual shapes,
// n x 64 by n x n), K is split over several workgroups whose FP64 partial sums
// go to a workspace and are added in split order by a second kernel, so the
// result does not depend on scheduling.
#include "internal.hpp"

namespace hplai {

typedef double doublex4_t __attribute__((ext_vector_type(4)));

constexpr int kG64T = 64;        // tile rows and columns
constexpr int kG64K = 16;        // K per chunk
constexpr int kG64Pad = 2;       // LDS row padding (doubles)

template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gemm64_kernel(int64_t m, int64_t n, int64_t k, int64_t kchunk, double alpha,
                                                     const double *__restrict__ A, int64_t lda,
                                                     const double *__restrict__ B, int64_t ldb, double beta,
                                                     double *__restrict__ C, int64_t ldc, double *__restrict__ W,
                                                     int tiles_m)
{
    __shared__ double As[kG64K][kG64T + kG64Pad];   // As[l][i] = op(A)(i0 + i, l0 + l)
    __shared__ double Bs[kG64K][kG64T + kG64Pad];   // Bs[l][j] = op(B)(l0 + l, j0 + j)
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wm = wave & 1, wn = wave >> 1;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x % tiles_m) * kG64T;
    const int64_t j0 = static_cast<int64_t>(blockIdx.x / tiles_m) * kG64T;
    const int64_t kb = static_cast<int64_t>(blockIdx.y) * kchunk;
    const int64_t ke = kb + kchunk < k ? kb + kchunk : k;

    // this thread's 4 elements of each operand chunk: (tile index, chunk index)
    // pairs chosen so that consecutive lanes read consecutive addresses
    int ai[4], al[4], bl[4], bj[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        if (TA) { al[r] = tid % kG64K; ai[r] = tid / kG64K + 16 * r; }
        else    { ai[r] = tid % kG64T; al[r] = tid / kG64T + 4 * r; }
        if (TB) { bj[r] = tid % kG64T; bl[r] = tid / kG64T + 4 * r; }
        else    { bl[r] = tid % kG64K; bj[r] = tid / kG64K + 16 * r; }
    }
    auto load = [&](int64_t l0, double (&av)[4], double (&bv)[4]) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int64_t i = i0 + ai[r], l = l0 + al[r];
            av[r] = (i < m && l < ke) ? (TA ? A[l + i * lda] : A[i + l * lda]) : 0.0;
            const int64_t j = j0 + bj[r], lb = l0 + bl[r];
            bv[r] = (j < n && lb < ke) ? (TB ? B[j + lb * ldb] : B[lb + j * ldb]) : 0.0;
        }
    };

    doublex4_t acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = doublex4_t{0.0, 0.0, 0.0, 0.0};

    double av[4], bv[4];
    load(kb, av, bv);
    for (int64_t l0 = kb; l0 < ke; l0 += kG64K) {
        __syncthreads();   // the previous chunk's MFMAs are done with the LDS
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            As[al[r]][ai[r]] = av[r];
            Bs[bl[r]][bj[r]] = bv[r];
        }
        __syncthreads();
        if (l0 + kG64K < ke) load(l0 + kG64K, av, bv);   // next chunk in flight during the MFMAs
        // v_mfma_f64_16x16x4_f64: lane l holds A(l % 16, l / 16) and B(l / 16, l % 16)
#pragma unroll
        for (int kk = 0; kk < kG64K; kk += 4) {
            double a[2], b[2];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                a[t] = As[kk + lane / 16][wm * 32 + t * 16 + lane % 16];
                b[t] = Bs[kk + lane / 16][wn * 32 + t * 16 + lane % 16];
            }
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x], b[y], acc[x][y], 0, 0, 0);
        }
    }

    // acc[x][y][q] is C(row l / 16 + 4 q, column l % 16) of block (x, y)
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) {
            const int64_t col = j0 + wn * 32 + y * 16 + lane % 16;
            if (col >= n) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t row = i0 + wm * 32 + x * 16 + lane / 16 + 4 * q;
                if (row >= m) continue;
                if (W != nullptr) {
                    W[static_cast<int64_t>(blockIdx.y) * m * n + col * m + row] = acc[x][y][q];
                } else {
                    double c = alpha * acc[x][y][q];
                    if (beta != 0.0) c = fma(beta, C[row + col * ldc], c);
                    C[row + col * ldc] = c;
                }
            }
        }
}

// C := alpha * (sum over the splits of W, in split order) + beta C
__global__ void gemm64_reduce_kernel(int64_t m, int64_t n, int splits, double alpha, const double *__restrict__ W,
                                     double beta, double *__restrict__ C, int64_t ldc)
{
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const int64_t row = e % m, col = e / m;
    double s = 0.0;
    for (int t = 0; t < splits; ++t) s += W[static_cast<int64_t>(t) * m * n + e];
    double c = alpha * s;
    if (beta != 0.0) c = fma(beta, C[row + col * ldc], c);
    C[row + col * ldc] = c;
}

int gemm64_splits(int64_t m, int64_t n, int64_t k, int num_cu)
{
    const int64_t tiles = ceil_div(m, kG64T) * ceil_div(n, kG64T);
    if (num_cu <= 0) num_cu = 256;
    if (tiles >= 2 * num_cu || k < 8 * kG64K) return 1;
    int64_t s = ceil_div(2 * static_cast<int64_t>(num_cu), tiles);
    const int64_t smax = k / (8 * kG64K);   // at least 8 chunks per split
    if (s > smax) s = smax;
    if (s > 64) s = 64;
    return static_cast<int>(s < 1 ? 1 : s);
}

int launch_gemm64(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha, const double *A, int64_t lda,
                  const double *B, int64_t ldb, double beta, double *C, int64_t ldc, double *W, int splits,
                  hipStream_t s)
{
    if (m <= 0 || n <= 0) return HPLAI_OK;
    const int64_t tiles_m = ceil_div(m, kG64T), tiles_n = ceil_div(n, kG64T);
    if (k <= 0) splits = 1;
    const int64_t kchunk = splits > 1 ? ceil_div(ceil_div(k, splits), kG64K) * kG64K : (k > 0 ? k : 1);
    const dim3 grid(static_cast<unsigned>(tiles_m * tiles_n), static_cast<unsigned>(splits));
    double *Wk = splits > 1 ? W : nullptr;
    if (ta && tb)
        hipLaunchKernelGGL((gemm64_kernel<true, true>), grid, dim3(256), 0, s, m, n, k, kchunk, alpha, A, lda, B,
                           ldb, beta, C, ldc, Wk, static_cast<int>(tiles_m));
    else if (ta)
        hipLaunchKernelGGL((gemm64_kernel<true, false>), grid, dim3(256), 0, s, m, n, k, kchunk, alpha, A, lda, B,
                           ldb, beta, C, ldc, Wk, static_cast<int>(tiles_m));
    else if (tb)
        hipLaunchKernelGGL((gemm64_kernel<false, true>), grid, dim3(256), 0, s, m, n, k, kchunk, alpha, A, lda, B,
                           ldb, beta, C, ldc, Wk, static_cast<int>(tiles_m));
    else
        hipLaunchKernelGGL((gemm64_kernel<false, false>), grid, dim3(256), 0, s, m, n, k, kchunk, alpha, A, lda, B,
                           ldb, beta, C, ldc, Wk, static_cast<int>(tiles_m));
    HPLAI_HIP(hipGetLastError());
    if (splits > 1) {
        hipLaunchKernelGGL(gemm64_reduce_kernel, dim3(static_cast<unsigned>(ceil_div(m * n, 256))), dim3(256), 0, s,
                           m, n, splits, alpha, W, beta, C, ldc);
        HPLAI_HIP(hipGetLastError());
    }
    return HPLAI_OK;
}

}  // namespace hplai

using namespace hplai;

extern "C" int hplai_gemm_f64(hplai_ctx ctx, int transa, int transb, int64_t m, int64_t n, int64_t k, double alpha,
                              const double *A, int64_t lda, const double *B, int64_t ldb, double beta, double *C,
                              int64_t ldc, void *stream)
{
    HPLAI_CHECK(ctx != nullptr, "null hplai_ctx");
    HPLAI_HIP(hipSetDevice(ctx->device));
    HPLAI_CHECK(m >= 0 && n >= 0 && k >= 0, "gemm_f64: negative size");
    if (m == 0 || n == 0) return HPLAI_OK;
    HPLAI_CHECK(C != nullptr && ldc >= m, "gemm_f64: C NULL or ldc < m");
    if (k > 0) {
        HPLAI_CHECK(A != nullptr && B != nullptr, "gemm_f64: A or B NULL");
        HPLAI_CHECK(lda >= (transa ? k : m) && ldb >= (transb ? n : k), "gemm_f64: lda or ldb too small");
    }
    hipStream_t s = as_stream(stream);
    const int splits = gemm64_splits(m, n, k, ctx->num_cu);
    double *W = nullptr;
    if (splits > 1)
        HPLAI_HIP(hipMallocAsync(reinterpret_cast<void **>(&W), sizeof(double) * splits * m * n, s));
    const int rc = launch_gemm64(transa != 0, transb != 0, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, W, splits, s);
    if (W != nullptr) {
        const hipError_t e = hipFreeAsync(W, s);
        if (rc == HPLAI_OK && e != hipSuccess) return hip_fail(e, "hipFreeAsync");
    }
    return rc;
}
