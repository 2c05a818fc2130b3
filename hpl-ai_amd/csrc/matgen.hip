// matgen.hip — the HPL-AI matrix generator on gfx950, bit-exact with
// hpl_rand.hpp (RIKEN-RCCS/hpl-ai).
//
// Element (i,j) of the n x n matrix is the LCG state reached after 1+i+n*j
// steps from the seed, mapped to [-0.5, 0.5).  Rows are consecutive states,
// columns jump by (a,c)^n.  The diagonal entry is sum_{j != i} |a_ij|,
// accumulated in FP64 with j ascending, exactly as calc_diag does.
#include "device_common.hpp"
#include "internal.hpp"

namespace hplai {

constexpr int kGenRows = 256;   // rows per block (one per thread)
constexpr int kGenCols = 64;    // columns per thread: amortises the jump-ahead

// fill_one_panel_with_rand (hpl_rand.hpp:59-106) for a block at (i0, j0).
__global__ __launch_bounds__(kGenRows) void gen_matrix_kernel(
    int64_t n, uint64_t x1, int64_t i0, int64_t j0, int64_t m, int64_t k, int64_t row_blocks,
    LcgCoeff jn, const double *__restrict__ d, float *__restrict__ A, int64_t lda, int accumulate,
    const uint16_t *__restrict__ P, const uint16_t *__restrict__ Q, int64_t ldp, float pscale)
{
    const int64_t rb = blockIdx.x % row_blocks;
    const int64_t cg = blockIdx.x / row_blocks;
    const int64_t il = rb * kGenRows + threadIdx.x;
    if (il >= m) return;
    const int64_t jl0 = cg * kGenCols;
    const int64_t jl1 = jl0 + kGenCols < k ? jl0 + kGenCols : k;
    const int64_t gi = i0 + il;
    uint64_t s = lcg_jump(x1, static_cast<uint64_t>(gi) +
                                  static_cast<uint64_t>(n) * static_cast<uint64_t>(j0 + jl0));
    float *col = A + jl0 * lda + il;
    const uint16_t *pcol = P != nullptr ? P + jl0 * ldp + il : nullptr;
    const uint16_t *qcol = Q != nullptr ? Q + jl0 * ldp + il : nullptr;
    for (int64_t jl = jl0; jl < jl1; ++jl, col += lda) {
        float v = static_cast<float>(lcg_value(s));
        if (d != nullptr && gi == j0 + jl) v = static_cast<float>(d[gi]);
        // lazy initialisation merges the partial sum: to = scale * from + a_ij
        // (LazyInitializer::update, lazy_init.hpp:147-227); the FP16 trailing
        // storage holds the partial sums of earlier outer steps times 1 / scale
        if (pcol != nullptr) {
            // with the back buffer Q (lazy_init.hpp:229-329): x = P + 32 Q
            float x = f16_bits_to_f32(*pcol);
            if (qcol != nullptr) {
                x = fmaf(32.0f, f16_bits_to_f32(*qcol), x);
                qcol += ldp;
            }
            v = fmaf(pscale, x, v);
            pcol += ldp;
        }
        *col = accumulate ? *col + v : v;
        s = jn.a * s + jn.c;
    }
}

// calc_diag (hpl_rand.hpp:108-123): one thread per row; the FP64 sum runs
// j = 0..n-1 in order.  Four interleaved LCG chains (stride 4 columns) break
// the 64-bit multiply dependency; the additions stay strictly sequential.
__global__ __launch_bounds__(256) void gen_diag_kernel(int64_t n, uint64_t x1, int64_t i0,
                                                       int64_t cnt, LcgCoeff jn, LcgCoeff jn4,
                                                       double *__restrict__ d)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    const int64_t i = i0 + t;
    uint64_t s0 = lcg_jump(x1, static_cast<uint64_t>(i));
    uint64_t s1 = jn.a * s0 + jn.c;
    uint64_t s2 = jn.a * s1 + jn.c;
    uint64_t s3 = jn.a * s2 + jn.c;
    double sum = 0.0;
    int64_t j = 0;
    for (; j + 3 < n; j += 4) {
        double v0 = fabs(lcg_value(s0));
        double v1 = fabs(lcg_value(s1));
        double v2 = fabs(lcg_value(s2));
        double v3 = fabs(lcg_value(s3));
        // adding +0.0 leaves a non-negative sum unchanged, so the skipped
        // diagonal term costs no branch
        sum = __dadd_rn(sum, j + 0 == i ? 0.0 : v0);
        sum = __dadd_rn(sum, j + 1 == i ? 0.0 : v1);
        sum = __dadd_rn(sum, j + 2 == i ? 0.0 : v2);
        sum = __dadd_rn(sum, j + 3 == i ? 0.0 : v3);
        s0 = jn4.a * s0 + jn4.c;
        s1 = jn4.a * s1 + jn4.c;
        s2 = jn4.a * s2 + jn4.c;
        s3 = jn4.a * s3 + jn4.c;
    }
    for (; j < n; ++j) {
        double v0 = fabs(lcg_value(s0));
        sum = __dadd_rn(sum, j == i ? 0.0 : v0);
        s0 = jn.a * s0 + jn.c;
    }
    d[t] = sum;
}

// pcolvgen (matgen.hpp:193-207): b_i = value of state 1 + i + n*n.
__global__ void gen_rhs_kernel(int64_t n, uint64_t x1, int64_t i0, int64_t cnt,
                               double *__restrict__ b)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= cnt) return;
    uint64_t idx = static_cast<uint64_t>(i0 + t) + static_cast<uint64_t>(n) * static_cast<uint64_t>(n);
    b[t] = lcg_value(lcg_jump(x1, idx));
}

int launch_gen_matrix(int64_t n, uint64_t seed, int64_t i0, int64_t j0, int64_t m, int64_t k,
                      const double *d, float *A, int64_t lda, int accumulate, hipStream_t s,
                      const uint16_t *P, int64_t ldp, float pscale, const uint16_t *Q)
{
    if (m <= 0 || k <= 0) return HPLAI_OK;
    const int64_t row_blocks = ceil_div(m, kGenRows);
    const int64_t col_groups = ceil_div(k, kGenCols);
    const LcgCoeff jn = lcg_pow_host(static_cast<uint64_t>(n));
    // split launches so the grid stays below 2^31 blocks
    const int64_t max_groups = (int64_t(1) << 30) / row_blocks;
    for (int64_t g0 = 0; g0 < col_groups; g0 += max_groups) {
        int64_t g = col_groups - g0 < max_groups ? col_groups - g0 : max_groups;
        int64_t jj = g0 * kGenCols;
        int64_t kk = (k - jj) < g * kGenCols ? (k - jj) : g * kGenCols;
        hipLaunchKernelGGL(gen_matrix_kernel, dim3(static_cast<unsigned>(row_blocks * g)),
                           dim3(kGenRows), 0, s, n, lcg_init(seed), i0, j0 + jj, m, kk, row_blocks,
                           jn, d, A + jj * lda, lda, accumulate, P != nullptr ? P + jj * ldp : nullptr,
                           Q != nullptr ? Q + jj * ldp : nullptr, ldp, pscale);
        HPLAI_HIP(hipGetLastError());
    }
    return HPLAI_OK;
}

// Back-buffer write-back (bb_writeback_b_impl_fp16, back_buffer.cpp:32-55 in
// its non-SVE form): the front P (the trailing updates' target) folds into
// the back Q with a compensated sum in FP16 arithmetic, Q := fl(Q + P/32),
// P := 32 (P/32 - (Q_new - Q)), so P keeps the rounding error.
__global__ __launch_bounds__(256) void backbuffer_kernel(int64_t m, int64_t k, uint16_t *__restrict__ P,
                                                         uint16_t *__restrict__ Q, int64_t ldp)
{
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (i >= m) return;
    const int64_t j = blockIdx.y;
    uint16_t *pp = P + j * ldp + i, *qq = Q + j * ldp + i;
    const _Float16 t0 = __builtin_bit_cast(_Float16, *qq);
    const _Float16 t1 = __builtin_bit_cast(_Float16, *pp) * static_cast<_Float16>(1.0f / 32.0f);
    const _Float16 t2 = t0 + t1;
    const _Float16 t3 = t2 - t0;
    const _Float16 t4 = t1 - t3;
    *qq = __builtin_bit_cast(uint16_t, t2);
    *pp = __builtin_bit_cast(uint16_t, static_cast<_Float16>(t4 * static_cast<_Float16>(32.0f)));
    (void)k;
}

int launch_backbuffer(int64_t m, int64_t k, uint16_t *P, uint16_t *Q, int64_t ldp, hipStream_t s)
{
    if (m <= 0 || k <= 0) return HPLAI_OK;
    for (int64_t j0 = 0; j0 < k; j0 += 65535) {
        const int64_t kk = k - j0 < 65535 ? k - j0 : 65535;
        hipLaunchKernelGGL(backbuffer_kernel, dim3(static_cast<unsigned>(ceil_div(m, 256)), static_cast<unsigned>(kk)),
                           dim3(256), 0, s, m, kk, P + j0 * ldp, Q + j0 * ldp, ldp);
        HPLAI_HIP(hipGetLastError());
    }
    return HPLAI_OK;
}

int launch_gen_diag(int64_t n, uint64_t seed, int64_t i0, int64_t cnt, double *d, hipStream_t s)
{
    if (cnt <= 0) return HPLAI_OK;
    const LcgCoeff jn = lcg_pow_host(static_cast<uint64_t>(n));
    const LcgCoeff jn4 = lcg_mul(lcg_mul(jn, jn), lcg_mul(jn, jn));
    hipLaunchKernelGGL(gen_diag_kernel, dim3(static_cast<unsigned>(ceil_div(cnt, 256))), dim3(256),
                       0, s, n, lcg_init(seed), i0, cnt, jn, jn4, d);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

int launch_gen_rhs(int64_t n, uint64_t seed, int64_t i0, int64_t cnt, double *b, hipStream_t s)
{
    if (cnt <= 0) return HPLAI_OK;
    hipLaunchKernelGGL(gen_rhs_kernel, dim3(static_cast<unsigned>(ceil_div(cnt, 256))), dim3(256), 0,
                       s, n, lcg_init(seed), i0, cnt, b);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

}  // namespace hplai
