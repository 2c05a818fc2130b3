// trsm.hip — panel TRSMs of the blocked LU with the whole panel height in LDS.
//
// One workgroup owns a strip of 32 columns (U panel) or 32 rows (L panel) of
// the panel and all b <= 512 entries of its other dimension, in LDS.  The
// strip is solved 32 at a time: the 32 x 32 diagonal sub-block is applied
// through its precomputed inverse (a dense 32 x 32 product, parallel over the
// workgroup), then a rank-32 update removes the solved part from the rest of
// the strip.  The inverses of the diagonal sub-blocks of L and U come from
// diag_inv32_kernel, run once per factored diagonal block.
// Panels with b > 512 use the register-blocked kernels in panel.hip.
#include "device_common.hpp"
#include "internal.hpp"

namespace hplai {

constexpr int kStrip = 32;          // strip width
constexpr int kPanelMax = 512;      // panel height held in LDS
constexpr int kXS = kStrip + 1;     // LDS stride (bank spread)
constexpr int kInvBlock = 2 * 32 * 32;   // floats per sub-block: Linv then Uinv

// inv[q] for q = 0..ceil(b/32)-1: Linv (row-major, unit lower) of the
// sub-block at (32q, 32q), then Uinv (row-major, upper) of the same.
__global__ __launch_bounds__(256) void diag_inv32_kernel(int b, const float *__restrict__ LU, int64_t lda,
                                                         float *__restrict__ inv)
{
    __shared__ float ML[32 * 33], MU[32 * 33], T[32 * 33];
    const int tid = threadIdx.x;
    const int sb = blockIdx.x * 32;
    const int w = b - sb < 32 ? b - sb : 32;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const int e = u * 256 + tid;
        const int r = e / 32, c = e % 32;
        const bool in = r < w && c < w;
        const float lv = LU[static_cast<int64_t>(sb + min(c, w - 1)) * lda + sb + min(r, w - 1)];
        const float v = in ? lv : 0.0f;   // LU(r, c)
        ML[r * 33 + c] = r > c ? v : (r == c ? 1.0f : 0.0f);                // unit lower L
        MU[c * 33 + r] = r < c ? v : (r == c ? (in ? v : 1.0f) : 0.0f);     // U^T (lower)
    }
    __syncthreads();
    lower_inv32(ML, T, tid, 256);
    lower_inv32(MU, T, tid, 256);
    float *o = inv + static_cast<int64_t>(blockIdx.x) * kInvBlock;
    for (int e = tid; e < 32 * 32; e += 256) {
        const int i = e / 32, j = e % 32;
        o[e] = ML[i * 33 + j];              // Linv(i, j)
        o[1024 + e] = MU[j * 33 + i];       // Uinv(i, j) = ((U^T)^-1)(j, i)
    }
}

// U(b x n) := Lkk^-1 U, Lkk unit lower (update_right_panels, panel_trf.hpp:158-171).
__global__ __launch_bounds__(256) void trsm_u_strip_kernel(int b, int64_t n, const float *__restrict__ Lk,
                                                           int64_t ldl, const float *__restrict__ inv,
                                                           float *__restrict__ U, int64_t ldu,
                                                           float *__restrict__ amax, int tb)
{
    __shared__ float X[kPanelMax * kXS];   // X[r * kXS + c]: row r of column j0 + c
    __shared__ float Li[kStrip * 33];      // Li[i * 33 + r] = Linv(i, r) of the current sub-block
    const int tid = threadIdx.x;
    const int64_t j0 = static_cast<int64_t>(blockIdx.x) * kStrip;
    const int nc = n - j0 < kStrip ? static_cast<int>(n - j0) : kStrip;
    // load the strip through registers, 16 elements per thread per batch, with
    // clamped (always valid) addresses so the loads issue together
    const int tot = b * kStrip;
    for (int base = 0; base < tot; base += 256 * 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = min(base + u * 256 + tid, tot - 1);
            const int c = idx / b, r = idx % b;
            v[u] = U[(j0 + min(c, nc - 1)) * ldu + r];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = base + u * 256 + tid;
            const int c = idx / b, r = idx % b;
            if (idx < tot) X[r * kXS + c] = c < nc ? v[u] : 0.0f;
        }
    }
    const int cc = tid & 31, g = tid >> 5;   // sub-block product: column cc, rows 4g..4g+3
    for (int kb = 0; kb < b; kb += kStrip) {
        const int w = b - kb < kStrip ? b - kb : kStrip;
        const float *src = inv + static_cast<int64_t>(kb / kStrip) * kInvBlock;
        {
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = src[u * 256 + tid];
#pragma unroll
            for (int u = 0; u < 4; ++u) Li[((u * 256 + tid) / 32) * 33 + (u * 256 + tid) % 32] = v[u];
        }
        __syncthreads();
        float out[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int r = 0; r < w; ++r) {
            const float x = X[(kb + r) * kXS + cc];
#pragma unroll
            for (int q = 0; q < 4; ++q) out[q] = fmaf(Li[(4 * g + q) * 33 + r], x, out[q]);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (4 * g + q < w) X[(kb + 4 * g + q) * kXS + cc] = out[q];
        __syncthreads();
        const int r0 = kb + w;
        if (r0 < b) {
            // rows r0.. : thread = 8 columns x up to 8 rows spaced by 64
            const int cs = (tid & 3) * 8, rs = tid >> 2;
            float acc[8][8] = {};
            for (int r = 0; r < w; ++r) {
                float xv[8];
#pragma unroll
                for (int c = 0; c < 8; ++c) xv[c] = X[(kb + r) * kXS + cs + c];
                const float *lcol = Lk + static_cast<int64_t>(kb + r) * ldl;
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int i = r0 + rs + 64 * q;
                    const float lv = lcol[min(i, b - 1)];
                    const float l = i < b ? lv : 0.0f;
#pragma unroll
                    for (int c = 0; c < 8; ++c) acc[q][c] = fmaf(l, xv[c], acc[q][c]);
                }
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int i = r0 + rs + 64 * q;
                if (i < b)
#pragma unroll
                    for (int c = 0; c < 8; ++c) X[i * kXS + cs + c] -= acc[q][c];
            }
        }
        __syncthreads();
    }
    float vmax = 0.0f;
    for (int idx = tid; idx < b * kStrip; idx += 256) {
        const int c = idx / b, r = idx % b;
        if (c < nc) {
            const float v = X[r * kXS + c];
            U[(j0 + c) * ldu + r] = v;
            vmax = fmaxf(vmax, fabsf(v));
        }
    }
    // all 32 columns of the strip lie in one tile when tb % 32 == 0
    tile_max_reduce(amax, j0 / tb, vmax, true);
}

// L(m x b) := L Ukk^-1, Ukk upper non-unit (update_left_panels, panel_trf.hpp:143-157).
__global__ __launch_bounds__(256) void trsm_l_strip_kernel(int64_t m, int b, const float *__restrict__ Uk,
                                                           int64_t ldu, const float *__restrict__ inv,
                                                           float *__restrict__ L, int64_t ldl,
                                                           float *__restrict__ amax, int tb)
{
    __shared__ float X[kPanelMax * kXS];   // X[c * kXS + r]: column c, row i0 + r
    __shared__ float Ui[kStrip * 33];      // Ui[l * 33 + c] = Uinv(l, c) of the current sub-block
    const int tid = threadIdx.x;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kStrip;
    const int nr = m - i0 < kStrip ? static_cast<int>(m - i0) : kStrip;
    const int tot = b * kStrip;
    for (int base = 0; base < tot; base += 256 * 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = min(base + u * 256 + tid, tot - 1);
            const int c = idx / kStrip, r = idx % kStrip;
            v[u] = L[static_cast<int64_t>(c) * ldl + i0 + min(r, nr - 1)];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = base + u * 256 + tid;
            const int c = idx / kStrip, r = idx % kStrip;
            if (idx < tot) X[c * kXS + r] = r < nr ? v[u] : 0.0f;
        }
    }
    const int rr = tid & 31, g = tid >> 5;   // sub-block product: row rr, columns 4g..4g+3
    for (int jb = 0; jb < b; jb += kStrip) {
        const int w = b - jb < kStrip ? b - jb : kStrip;
        const float *src = inv + static_cast<int64_t>(jb / kStrip) * kInvBlock + 1024;
        {
            float v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = src[u * 256 + tid];
#pragma unroll
            for (int u = 0; u < 4; ++u) Ui[((u * 256 + tid) / 32) * 33 + (u * 256 + tid) % 32] = v[u];
        }
        __syncthreads();
        float out[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int l = 0; l < w; ++l) {
            const float x = X[(jb + l) * kXS + rr];
#pragma unroll
            for (int q = 0; q < 4; ++q) out[q] = fmaf(x, Ui[l * 33 + 4 * g + q], out[q]);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (4 * g + q < w) X[(jb + 4 * g + q) * kXS + rr] = out[q];
        __syncthreads();
        // columns jb+w..b: thread = one column, all 32 rows; U(jb..jb+w, c) in registers
        for (int c = jb + w + tid; c < b; c += 256) {
            float u[kStrip];
            const float *ucol = Uk + static_cast<int64_t>(c) * ldu + jb;
#pragma unroll
            for (int l = 0; l < kStrip; ++l) {
                const float v = ucol[min(l, w - 1)];
                u[l] = l < w ? v : 0.0f;
            }
            for (int r = 0; r < kStrip; ++r) {
                float acc = 0.0f;
#pragma unroll
                for (int l = 0; l < kStrip; ++l) acc = fmaf(X[(jb + l) * kXS + r], u[l], acc);
                X[c * kXS + r] -= acc;
            }
        }
        __syncthreads();
    }
    float vmax = 0.0f;
    for (int idx = tid; idx < b * kStrip; idx += 256) {
        const int c = idx / kStrip, r = idx % kStrip;
        if (r < nr) {
            const float v = X[c * kXS + r];
            L[static_cast<int64_t>(c) * ldl + i0 + r] = v;
            vmax = fmaxf(vmax, fabsf(v));
        }
    }
    tile_max_reduce(amax, i0 / tb, vmax, true);
}

// ---------------------------------------------------------------------------
// MFMA forms (b % 32 == 0): the same strip algorithm with every 32 x 32 x 32
// product on v_mfma_f32_32x32x2_f32 (FP32 inputs, FP32 accumulation), twice
// the FP32 rate of the vector ALU.  Operands: the strip from LDS, the panel
// block from global memory (L2-resident), each lane loading its 128-byte
// column segment once and picking the even or odd K entries it feeds.
// Output layout of the MFMA: acc[4 q + r] is row 8 q + 4 (lane / 32) + r,
// column lane % 32 of the 32 x 32 block.
// ---------------------------------------------------------------------------
typedef float floatx16t_t __attribute__((ext_vector_type(16)));

// v[0..15] = seg[2 kk + h] for kk = 0..15 from a 32-float contiguous segment
__device__ __forceinline__ void load_k_pairs(const float *__restrict__ seg, int h, float (&v)[16])
{
    float t[32];
#pragma unroll
    for (int e = 0; e < 8; ++e) {
        const float4 q = reinterpret_cast<const float4 *>(seg)[e];
        t[4 * e] = q.x;
        t[4 * e + 1] = q.y;
        t[4 * e + 2] = q.z;
        t[4 * e + 3] = q.w;
    }
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) v[kk] = h ? t[2 * kk + 1] : t[2 * kk];
}

// L(m x b) := L Ukk^-1 (update_left_panels, panel_trf.hpp:143-157), b % 32 == 0.
__global__ __launch_bounds__(256) void trsm_l_mfma_kernel(int64_t m, int b, const float *__restrict__ Uk,
                                                          int64_t ldu, const float *__restrict__ inv,
                                                          float *__restrict__ L, int64_t ldl,
                                                          float *__restrict__ amax, int tb)
{
    __shared__ float X[kPanelMax * kXS];   // X[c * kXS + r]: column c, row i0 + r
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, lh = lane >> 5;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * kStrip;
    const int nr = m - i0 < kStrip ? static_cast<int>(m - i0) : kStrip;
    const int tot = b * kStrip;
    for (int base = 0; base < tot; base += 256 * 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = min(base + u * 256 + tid, tot - 1);
            const int c = idx / kStrip, r = idx % kStrip;
            v[u] = L[static_cast<int64_t>(c) * ldl + i0 + min(r, nr - 1)];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = base + u * 256 + tid;
            const int c = idx / kStrip, r = idx % kStrip;
            if (idx < tot) X[c * kXS + r] = r < nr ? v[u] : 0.0f;
        }
    }
    __syncthreads();
    for (int jb = 0; jb < b; jb += kStrip) {
        // X(:, jb:jb+32) := X(:, jb:jb+32) Uinv  (wave 0)
        if (wave == 0) {
            const float *Ui = inv + static_cast<int64_t>(jb / kStrip) * kInvBlock + 1024;   // row-major
            floatx16t_t acc;
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) {
                const int k = 2 * kk + lh;
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(X[(jb + k) * kXS + l32], Ui[k * 32 + l32], acc, 0, 0, 0);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) X[(jb + l32) * kXS + 8 * q + 4 * lh + r] = acc[4 * q + r];
        }
        __syncthreads();
        // X(:, c0:c0+32) -= X(:, jb:jb+32) U(jb:jb+32, c0:c0+32) for the blocks right of jb
        const int nblk = (b - jb - kStrip) / kStrip;
        for (int blk = wave; blk < nblk; blk += 4) {
            const int c0 = jb + kStrip + blk * kStrip;
            float bv[16];
            load_k_pairs(Uk + jb + static_cast<int64_t>(c0 + l32) * ldu, lh, bv);
            floatx16t_t acc;
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(X[(jb + 2 * kk + lh) * kXS + l32], bv[kk], acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) X[(c0 + l32) * kXS + 8 * q + 4 * lh + r] -= acc[4 * q + r];
        }
        __syncthreads();
    }
    float vmax = 0.0f;
    for (int idx = tid; idx < b * kStrip; idx += 256) {
        const int c = idx / kStrip, r = idx % kStrip;
        if (r < nr) {
            const float v = X[c * kXS + r];
            L[static_cast<int64_t>(c) * ldl + i0 + r] = v;
            vmax = fmaxf(vmax, fabsf(v));
        }
    }
    tile_max_reduce(amax, i0 / tb, vmax, true);
}

// U(b x n) := Lkk^-1 U (update_right_panels, panel_trf.hpp:158-171), b % 32 == 0.
__global__ __launch_bounds__(256) void trsm_u_mfma_kernel(int b, int64_t n, const float *__restrict__ Lk,
                                                          int64_t ldl, const float *__restrict__ inv,
                                                          float *__restrict__ U, int64_t ldu,
                                                          float *__restrict__ amax, int tb)
{
    __shared__ float X[kPanelMax * kXS];   // X[r * kXS + c]: row r of column j0 + c
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, lh = lane >> 5;
    const int64_t j0 = static_cast<int64_t>(blockIdx.x) * kStrip;
    const int nc = n - j0 < kStrip ? static_cast<int>(n - j0) : kStrip;
    const int tot = b * kStrip;
    for (int base = 0; base < tot; base += 256 * 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = min(base + u * 256 + tid, tot - 1);
            const int c = idx / b, r = idx % b;
            v[u] = U[(j0 + min(c, nc - 1)) * ldu + r];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = base + u * 256 + tid;
            const int c = idx / b, r = idx % b;
            if (idx < tot) X[r * kXS + c] = c < nc ? v[u] : 0.0f;
        }
    }
    __syncthreads();
    for (int kb = 0; kb < b; kb += kStrip) {
        // X(kb:kb+32, :) := Linv X(kb:kb+32, :)  (wave 0)
        if (wave == 0) {
            const float *Li = inv + static_cast<int64_t>(kb / kStrip) * kInvBlock;   // row-major Linv
            float av[16];
            load_k_pairs(Li + l32 * 32, lh, av);
            floatx16t_t acc;
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[kk], X[(kb + 2 * kk + lh) * kXS + l32], acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) X[(kb + 8 * q + 4 * lh + r) * kXS + l32] = acc[4 * q + r];
        }
        __syncthreads();
        // X(r0:r0+32, :) -= L(r0:r0+32, kb:kb+32) X(kb:kb+32, :) for the blocks below kb
        const int nblk = (b - kb - kStrip) / kStrip;
        for (int blk = wave; blk < nblk; blk += 4) {
            const int r0 = kb + kStrip + blk * kStrip;
            float av[16];
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) av[kk] = Lk[(r0 + l32) + static_cast<int64_t>(kb + 2 * kk + lh) * ldl];
            floatx16t_t acc;
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[kk], X[(kb + 2 * kk + lh) * kXS + l32], acc, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) X[(r0 + 8 * q + 4 * lh + r) * kXS + l32] -= acc[4 * q + r];
        }
        __syncthreads();
    }
    float vmax = 0.0f;
    for (int idx = tid; idx < b * kStrip; idx += 256) {
        const int c = idx / b, r = idx % b;
        if (c < nc) {
            const float v = X[r * kXS + c];
            U[(j0 + c) * ldu + r] = v;
            vmax = fmaxf(vmax, fabsf(v));
        }
    }
    tile_max_reduce(amax, j0 / tb, vmax, true);
}

// The strip kernels need b <= 512 and, when tile maxima are collected, every
// strip inside one tile (tb % 32 == 0).
bool trsm_strip_ok(int b, const float *amax, int tb)
{
    return b <= kPanelMax && (amax == nullptr || (tb % kStrip) == 0);
}

int64_t diag_inv_floats(int b) { return ceil_div(b, kStrip) * kInvBlock; }

int launch_diag_inv(int b, const float *LU, int64_t lda, float *inv, hipStream_t s)
{
    hipLaunchKernelGGL(diag_inv32_kernel, dim3(static_cast<unsigned>(ceil_div(b, kStrip))), dim3(256), 0,
                       s, b, LU, lda, inv);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

int launch_trsm_l_strip(int64_t m, int b, const float *U, int64_t ldu, const float *inv, float *L,
                        int64_t ldl, float *amax, int tb, hipStream_t s)
{
    if (b % kStrip == 0 && (ldu % 4) == 0 && (reinterpret_cast<uintptr_t>(U) % 16) == 0)
        hipLaunchKernelGGL(trsm_l_mfma_kernel, dim3(static_cast<unsigned>(ceil_div(m, kStrip))), dim3(256),
                           0, s, m, b, U, ldu, inv, L, ldl, amax, tb);
    else
        hipLaunchKernelGGL(trsm_l_strip_kernel, dim3(static_cast<unsigned>(ceil_div(m, kStrip))), dim3(256),
                           0, s, m, b, U, ldu, inv, L, ldl, amax, tb);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

int launch_trsm_u_strip(int b, int64_t n, const float *Lkk, int64_t ldl, const float *inv, float *U,
                        int64_t ldu, float *amax, int tb, hipStream_t s)
{
    if (b % kStrip == 0)
        hipLaunchKernelGGL(trsm_u_mfma_kernel, dim3(static_cast<unsigned>(ceil_div(n, kStrip))), dim3(256),
                           0, s, b, n, Lkk, ldl, inv, U, ldu, amax, tb);
    else
        hipLaunchKernelGGL(trsm_u_strip_kernel, dim3(static_cast<unsigned>(ceil_div(n, kStrip))), dim3(256),
                           0, s, b, n, Lkk, ldl, inv, U, ldu, amax, tb);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

}  // namespace hplai
