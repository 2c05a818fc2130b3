// gemmschur.hip — the reference's column-major Schur seam.
//
// gemmschur(int n, int m, int k, const fp16 *a, int lda, const fp16 *b, int ldb,
//           float *c, int ldc)  (schur_updator.hpp:28-31 -> shgemm,
// fp16sim.hpp:217-233) computes C(n x m) -= A(n x k) B(k x m) with every
// operand column-major as the reference lays it out: A is the L panel that
// convert_left_panels writes (column l at a + l*lda, panel_trf.hpp:182-207),
// B the U panel of convert_right_panels (column j at b + j*ldb,
// panel_trf.hpp:208-233), both binary16, C FP32.
//
// The MFMA Schur kernels read K-contiguous operand rows whose K is a whole
// number of 128-byte stages.  The seam therefore stages both operands in a
// scratch buffer: A is transposed through LDS (64 rows x 64 K per
// workgroup, coalesced on both sides), B's columns are copied, and K is
// padded with zeros to a multiple of 64; then the update runs with FP32
// accumulation.  Any k >= 0 is accepted (the reference's C1 uses k = 60).
#include "internal.hpp"

namespace hplai {

constexpr int kSeamTile = 64;

// dst[i * kpad + kk] = a[kk * lda + i] (0 for kk >= k): one 64 x 64 tile
__global__ __launch_bounds__(256) void seam_pack_a_kernel(int64_t n, int64_t k, int64_t kpad,
                                                          const uint16_t *__restrict__ a, int64_t lda,
                                                          uint16_t *__restrict__ dst)
{
    __shared__ uint16_t t[kSeamTile][kSeamTile + 2];
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kSeamTile;
    const int64_t k0 = static_cast<int64_t>(blockIdx.y) * kSeamTile;
    const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
    for (int kk = ty; kk < kSeamTile; kk += 4) {
        const int64_t gk = k0 + kk, gi = i0 + tx;
        t[kk][tx] = (gk < k && gi < n) ? a[gk * lda + gi] : uint16_t(0);
    }
    __syncthreads();
    for (int r = ty; r < kSeamTile; r += 4) {
        const int64_t gi = i0 + r, gk = k0 + tx;
        if (gi < n && gk < kpad) dst[gi * kpad + gk] = t[tx][r];
    }
}

// dst[j * kpad + kk] = b[j * ldb + kk] (0 for kk >= k)
__global__ __launch_bounds__(256) void seam_pack_b_kernel(int64_t m, int64_t k, int64_t kpad,
                                                          const uint16_t *__restrict__ b, int64_t ldb,
                                                          uint16_t *__restrict__ dst)
{
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= m * kpad) return;
    const int64_t j = idx / kpad, kk = idx - j * kpad;
    dst[idx] = kk < k ? b[j * ldb + kk] : uint16_t(0);
}

int64_t gemmschur_scratch_bytes(int64_t n, int64_t m, int64_t k)
{
    const int64_t kpad = ceil_div(k, kSeamTile) * kSeamTile;
    return (n + m) * kpad * static_cast<int64_t>(sizeof(uint16_t));
}

int launch_gemmschur_f16(int64_t n, int64_t m, int64_t k, const uint16_t *a, int64_t lda, const uint16_t *b,
                         int64_t ldb, float *c, int64_t ldc, uint16_t *scratch, int num_cu, hipStream_t s)
{
    if (n <= 0 || m <= 0 || k <= 0) return HPLAI_OK;
    const int64_t kpad = ceil_div(k, kSeamTile) * kSeamTile;
    uint16_t *ap = scratch, *bp = scratch + n * kpad;
    HPLAI_CHECK(ceil_div(n, kSeamTile) < (int64_t(1) << 31) && ceil_div(kpad, kSeamTile) <= 65535,
                "gemmschur: n=%lld k=%lld too large", (long long)n, (long long)k);
    hipLaunchKernelGGL(seam_pack_a_kernel,
                       dim3(static_cast<unsigned>(ceil_div(n, kSeamTile)), static_cast<unsigned>(kpad / kSeamTile)),
                       dim3(256), 0, s, n, k, kpad, a, lda, ap);
    HPLAI_HIP(hipGetLastError());
    const int64_t nb = m * kpad;
    HPLAI_CHECK(ceil_div(nb, 256) < (int64_t(1) << 31), "gemmschur: m=%lld k=%lld too large", (long long)m,
                (long long)k);
    hipLaunchKernelGGL(seam_pack_b_kernel, dim3(static_cast<unsigned>(ceil_div(nb, 256))), dim3(256), 0, s, m, k,
                       kpad, b, ldb, bp);
    HPLAI_HIP(hipGetLastError());
    // kslice must be a multiple of 64 binary16 values; one slice covers K
    return launch_schur(HPLAI_FMT_F16, n, m, 1, static_cast<int>(kpad), ap, kpad, nullptr, 0, bp, kpad, nullptr,
                        0, 0, c, ldc, num_cu, s, false);
}

}  // namespace hplai
