// panel.hip — FP32 panel kernels of the blocked no-pivot LU on gfx950:
// diagonal-block LU, the two panel TRSMs, per-tile |max| and the
// FP32 -> FP16 / FP8-e4m3 quantizer that feeds the MFMA Schur update.
#include "device_common.hpp"
#include "internal.hpp"

namespace hplai {

// ---------------------------------------------------------------------------
// Diagonal-block LU (getrf_nopiv, getrf_nopiv.hpp:124-141 with NSMALL panels)
//
// One workgroup of 512 threads factors an n x n block (n <= NMAX) in place.
// Right-looking over column panels of width W held in LDS (column-major,
// rows contiguous), the U strip L11^-1 A12 in LDS, and a register-blocked
// 4x4 trailing update A22 -= L21 U12 read from / written to global (L2).
// ---------------------------------------------------------------------------
constexpr int kGetrfThreads = 512;

template <int NMAX, int W>
__global__ __launch_bounds__(kGetrfThreads) void getrf_kernel(int n, float *__restrict__ A,
                                                              int64_t lda, int *__restrict__ info)
{
    constexpr int NT = kGetrfThreads;
    constexpr int RPT = NMAX / NT;   // panel rows per thread
    // +8: the 8 x 8 tiles of the trailing update may read past the last row /
    // column (values unused)
    __shared__ __attribute__((aligned(16))) float Lc[W * NMAX + 8];   // panel, Lc[c * NMAX + r]
    __shared__ __attribute__((aligned(16))) float Us[W * NMAX + 8];   // U strip, Us[r * NMAX + c]
    __shared__ float prow[2][W];     // pivot row, double-buffered
    __shared__ float Mi[32 * 33], Ti[32 * 33];   // L11 inverse (W == 32)
    const int tid = threadIdx.x;
    int bad = 0;
    for (int k = 0; k < n; k += W) {
        const int w = n - k < W ? n - k : W;
        const int h = n - k;
        float *Ak = A + static_cast<int64_t>(k) * lda + k;
        // each thread holds its panel rows in registers
        float p[RPT][W];
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + q * NT;
#pragma unroll
            for (int c = 0; c < W; ++c) {
                const float v = Ak[static_cast<int64_t>(min(c, w - 1)) * lda + min(r, h - 1)];
                p[q][c] = (r < h && c < w) ? v : 0.0f;
            }
        }
        // unblocked right-looking LU of the h x w panel (getrf_nopiv_small,
        // getrf_nopiv.hpp:30-46, continued down the rows = trsmR); the owner of
        // row kk publishes it, one barrier per pivot
#pragma unroll
        for (int kk = 0; kk < W; ++kk) {
            if (kk < w) {
                if (tid == kk)
#pragma unroll
                    for (int c = 0; c < W; ++c) prow[kk & 1][c] = p[0][c];
                __syncthreads();
                const float piv = prow[kk & 1][kk];
                const float inv = 1.0f / piv;
                if (tid == 0 && bad == 0 && !(piv != 0.0f && isfinite(piv))) bad = k + kk + 1;
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int r = tid + q * NT;
                    if (r > kk && r < h) {
                        const float l = p[q][kk] * inv;
                        p[q][kk] = l;
#pragma unroll
                        for (int c = kk + 1; c < W; ++c) p[q][c] = p[q][c] - l * prow[kk & 1][c];
                    }
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int q = 0; q < RPT; ++q) {
            const int r = tid + q * NT;
            if (r < h)
#pragma unroll
                for (int c = 0; c < W; ++c)
                    if (c < w) {
                        Ak[static_cast<int64_t>(c) * lda + r] = p[q][c];
                        Lc[c * NMAX + r] = p[q][c];
                    }
        }
        __syncthreads();
        const int wr = n - k - w;   // columns right of the panel
        if constexpr (W == 32) {
            // U strip through the inverse of the unit-lower L11: x := L11^-1 x
            for (int e = tid; e < 32 * 32; e += NT) {
                const int r = e / 32, c = e % 32;
                Mi[r * 33 + c] = (r > c && r < w) ? Lc[c * NMAX + r] : (r == c ? 1.0f : 0.0f);
            }
            __syncthreads();
            lower_inv32(Mi, Ti, tid, NT);
            for (int c = tid; c < wr; c += NT) {
                float x[W];
                float *col = Ak + static_cast<int64_t>(w + c) * lda;
#pragma unroll
                for (int r = 0; r < W; ++r) {
                    const float v = col[min(r, w - 1)];
                    x[r] = r < w ? v : 0.0f;
                }
                for (int i = 0; i < w; ++i) {
                    float acc = 0.0f;
#pragma unroll
                    for (int r = 0; r < W; ++r) acc = fmaf(Mi[i * 33 + r], x[r], acc);
                    col[i] = acc;
                    Us[i * NMAX + c] = acc;
                }
            }
        } else {
            // U strip: x := L11^-1 x per column (strsm "L","L","N","U" order)
            for (int c = tid; c < wr; c += NT) {
                float x[W];
                float *col = Ak + static_cast<int64_t>(w + c) * lda;
#pragma unroll
                for (int r = 0; r < W; ++r) {
                    const float v = col[min(r, w - 1)];
                    x[r] = r < w ? v : 0.0f;
                }
#pragma unroll
                for (int r = 0; r < W; ++r)
#pragma unroll
                    for (int i = r + 1; i < W; ++i) x[i] = x[i] - x[r] * Lc[r * NMAX + i];
#pragma unroll
                for (int r = 0; r < W; ++r)
                    if (r < w) {
                        col[r] = x[r];
                        Us[r * NMAX + c] = x[r];
                    }
            }
        }
        __syncthreads();
        // trailing update A22 -= L21 U12, 8 x 8 register tiles (rows fastest
        // across threads); L21 and U12 come from LDS as 16-byte reads.  Rows /
        // columns past the edge read padding and are never stored.
        const int hr = h - w;
        if (hr > 0 && wr > 0) {
            const int RT = (hr + 7) / 8, CT = (wr + 7) / 8;
            for (int t = tid; t < RT * CT; t += NT) {
                const int r0 = w + 8 * (t % RT), c0 = 8 * (t / RT);
                float acc[8][8];   // [column][row], starts from A22
#pragma unroll
                for (int bq = 0; bq < 8; ++bq)
#pragma unroll
                    for (int a = 0; a < 8; ++a)
                        acc[bq][a] = Ak[static_cast<int64_t>(w + min(c0 + bq, wr - 1)) * lda + min(r0 + a, h - 1)];
                for (int kk = 0; kk < w; ++kk) {
                    const float4 l0 = *reinterpret_cast<const float4 *>(&Lc[kk * NMAX + r0]);
                    const float4 l1 = *reinterpret_cast<const float4 *>(&Lc[kk * NMAX + r0 + 4]);
                    const float4 u0 = *reinterpret_cast<const float4 *>(&Us[kk * NMAX + c0]);
                    const float4 u1 = *reinterpret_cast<const float4 *>(&Us[kk * NMAX + c0 + 4]);
                    const float l[8] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w};
                    const float u[8] = {u0.x, u0.y, u0.z, u0.w, u1.x, u1.y, u1.z, u1.w};
#pragma unroll
                    for (int bq = 0; bq < 8; ++bq)
#pragma unroll
                        for (int a = 0; a < 8; ++a) acc[bq][a] = fmaf(-l[a], u[bq], acc[bq][a]);
                }
#pragma unroll
                for (int bq = 0; bq < 8; ++bq)
#pragma unroll
                    for (int a = 0; a < 8; ++a)
                        if (r0 + a < h && c0 + bq < wr)
                            Ak[static_cast<int64_t>(w + c0 + bq) * lda + r0 + a] = acc[bq][a];
            }
        }
        __syncthreads();
    }
    if (tid == 0 && info != nullptr) *info = bad;
}

// ---------------------------------------------------------------------------
// The same blocked LU for n <= 512 with n % 32 == 0 and 16-byte aligned
// columns, built around the FP32 MFMA (v_mfma_f32_32x32x2_f32, FP32 inputs
// and accumulation).  Per 32-column step k:
//  1. the h x 32 panel and the 32 x (n - k - 32) U strip are staged in LDS;
//  2. wave 0 factors the 32 x 32 diagonal block in registers (lane i holds
//     row i; pivot rows are broadcast with v_readlane, no barriers);
//  3. wave 0 inverts the unit-lower L11, wave 1 the upper U11 (Gauss-Jordan
//     on the rows, again broadcast by readlane);
//  4. every wave takes 32 x 32 tiles of L21 := A21 U11^-1 and
//     U12 := L11^-1 A12 on the MFMA;
//  5. every wave updates 32 x 32 tiles of A22 -= L21 U12 (C in global memory,
//     four tiles with their sixteen float4 loads in flight together).
// L21 and U12 come from LDS: Lc column-major (stride kLcS), Us column-major
// (stride 33, one column of the strip per 33 floats).
// ---------------------------------------------------------------------------
typedef float floatx16g_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float readlane_f(float v, int l)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}

constexpr int kLcS = 512 + 4;   // Lc column stride (float4 column writes spread over the banks)

__global__ __launch_bounds__(kGetrfThreads) void getrf_mfma_kernel(int n, float *__restrict__ A, int64_t lda,
                                                                   int *__restrict__ info)
{
    constexpr int NT = kGetrfThreads, NMAX = 512, W = 32;
    __shared__ __attribute__((aligned(16))) float Lc[W * kLcS];          // panel, Lc[c * kLcS + r]
    __shared__ __attribute__((aligned(16))) float Us[(NMAX - W) * 33];   // U strip, Us[c * 33 + r]
    __shared__ float Li[32 * 33], Ui[32 * 33];                           // L11^-1, U11^-1 row-major
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, lh = lane >> 5;
    int bad = 0;
    for (int k = 0; k < n; k += W) {
        const int h = n - k, wr = n - k - W;
        float *Ak = A + static_cast<int64_t>(k) * lda + k;
        // 1. stage the panel (h x 32) and the U strip (32 x wr): every load of
        //    the step is issued before the first LDS write (one L2 latency per
        //    step instead of one per loop trip).  Panel: thread takes row chunk
        //    tid % 128 of columns tid / 128 + 4 u; U strip: row chunk tid % 8 of
        //    columns tid / 8 + 64 u.
        {
            const int pf = tid & 127, pg = tid >> 7;
            const int ui = tid & 7, uc = tid >> 3;
            const bool pin = 4 * pf < h;
            const float *pcol = Ak + static_cast<int64_t>(pg) * lda + 4 * pf;
            const float *ucol = Ak + static_cast<int64_t>(W + uc) * lda + 4 * ui;
            float4 vp[8], vu[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                vp[u] = pin ? *reinterpret_cast<const float4 *>(pcol + static_cast<int64_t>(4 * u) * lda)
                            : float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int u = 0; u < 8; ++u)
                vu[u] = uc + 64 * u < wr ? *reinterpret_cast<const float4 *>(ucol + static_cast<int64_t>(64 * u) * lda)
                                         : float4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (pin) *reinterpret_cast<float4 *>(&Lc[(pg + 4 * u) * kLcS + 4 * pf]) = vp[u];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                if (uc + 64 * u < wr) {
                    float *d = &Us[(uc + 64 * u) * 33 + 4 * ui];
                    d[0] = vu[u].x;
                    d[1] = vu[u].y;
                    d[2] = vu[u].z;
                    d[3] = vu[u].w;
                }
            }
        }
        __syncthreads();
        // 2. LU of the 32 x 32 diagonal block by wave 0
        if (wave == 0) {
            const int i = l32;
            float a[W];
#pragma unroll
            for (int j = 0; j < W; ++j) a[j] = Lc[j * kLcS + i];
#pragma unroll
            for (int kk = 0; kk < W; ++kk) {
                const float piv = readlane_f(a[kk], kk);
                if (bad == 0 && !(piv != 0.0f && isfinite(piv))) bad = k + kk + 1;
                const float inv = 1.0f / piv;
                const bool below = i > kk;
                const float l = a[kk] * inv;
                a[kk] = below ? l : a[kk];
#pragma unroll
                for (int j = kk + 1; j < W; ++j) {
                    const float ukj = readlane_f(a[j], kk);
                    a[j] = below ? fmaf(-l, ukj, a[j]) : a[j];
                }
            }
            if (lh == 0) {
#pragma unroll
                for (int j = 0; j < W; ++j) {
                    Lc[j * kLcS + i] = a[j];
                    Ak[static_cast<int64_t>(j) * lda + i] = a[j];
                }
            }
        }
        __syncthreads();
        if (wr <= 0) break;
        // 3. L11^-1 (waves 0 and 2) and U11^-1 (waves 1 and 3), Gauss-Jordan on
        //    rows held one per lane; the columns of an inverse are independent,
        //    so each of the two waves takes every second one (j = 2 jj + par)
        if (wave < 4) {
            const int i = l32, par = wave >> 1;
            float a[W], x[W / 2];
#pragma unroll
            for (int j = 0; j < W; ++j) a[j] = Lc[j * kLcS + i];
#pragma unroll
            for (int jj = 0; jj < W / 2; ++jj) x[jj] = 2 * jj + par == i ? 1.0f : 0.0f;
            if ((wave & 1) == 0) {
                // rows i > c: X(i, :) -= L(i, c) X(c, :), X(c, :) final at step c
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    const bool below = i > c;
#pragma unroll
                    for (int jj = 0; jj < W / 2; ++jj)
                        if (2 * jj + par <= c) {
                            const float xcj = readlane_f(x[jj], c);
                            x[jj] = below ? fmaf(-a[c], xcj, x[jj]) : x[jj];
                        }
                }
                if (lh == 0)
#pragma unroll
                    for (int jj = 0; jj < W / 2; ++jj) Li[i * 33 + 2 * jj + par] = x[jj];
            } else {
                // from the bottom: Y(c, :) /= U(c, c), then rows i < c: Y(i, :) -= U(i, c) Y(c, :)
#pragma unroll
                for (int c = W - 1; c >= 0; --c) {
                    const float rd = 1.0f / readlane_f(a[c], c);
                    if (i == c)
#pragma unroll
                        for (int jj = 0; jj < W / 2; ++jj)
                            if (2 * jj + par >= c) x[jj] *= rd;
                    const bool above = i < c;
#pragma unroll
                    for (int jj = 0; jj < W / 2; ++jj)
                        if (2 * jj + par >= c) {
                            const float ycj = readlane_f(x[jj], c);
                            x[jj] = above ? fmaf(-a[c], ycj, x[jj]) : x[jj];
                        }
                }
                if (lh == 0)
#pragma unroll
                    for (int jj = 0; jj < W / 2; ++jj) Ui[i * 33 + 2 * jj + par] = x[jj];
            }
        }
        __syncthreads();
        // 4. L21 := A21 U11^-1 (TR tiles) and U12 := L11^-1 A12 (TC tiles)
        const int TR = (h - W) / 32, TC = wr / 32;
        for (int t = wave; t < TR + TC; t += NT / 64) {
            floatx16g_t acc;
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
            if (t < TR) {
                const int r0 = W + 32 * t;
                float av[16], bv[16];
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) {
                    av[kk] = Lc[(2 * kk + lh) * kLcS + r0 + l32];
                    bv[kk] = Ui[(2 * kk + lh) * 33 + l32];
                }
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[kk], bv[kk], acc, 0, 0, 0);
                float *g = Ak + static_cast<int64_t>(l32) * lda + r0 + 4 * lh;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 o = {acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]};
                    *reinterpret_cast<float4 *>(&Lc[l32 * kLcS + r0 + 8 * q + 4 * lh]) = o;
                    *reinterpret_cast<float4 *>(g + 8 * q) = o;
                }
            } else {
                const int c0 = 32 * (t - TR);
                float av[16], bv[16];
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) {
                    av[kk] = Li[l32 * 33 + 2 * kk + lh];
                    bv[kk] = Us[(c0 + l32) * 33 + 2 * kk + lh];
                }
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av[kk], bv[kk], acc, 0, 0, 0);
                float *g = Ak + static_cast<int64_t>(W + c0 + l32) * lda + 4 * lh;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 o = {acc[4 * q], acc[4 * q + 1], acc[4 * q + 2], acc[4 * q + 3]};
#pragma unroll
                    for (int r = 0; r < 4; ++r) Us[(c0 + l32) * 33 + 8 * q + 4 * lh + r] = acc[4 * q + r];
                    *reinterpret_cast<float4 *>(g + 8 * q) = o;
                }
            }
        }
        __syncthreads();
        // 5. trailing update A22 -= L21 U12 (hr x wr, both multiples of 32):
        //    one tile per wave at a time; the C of the wave's next tile loads
        //    while this one is computed (two at a time: 5% slower, four: spills)
        const int ntile = TR * TC;
        constexpr int TPW = 1, STEP = TPW * (NT / 64);
        auto load_c = [&](int t0, float4 (&v)[TPW][4]) {
#pragma unroll
            for (int u = 0; u < TPW; ++u) {
                const int t = t0 + u < ntile ? t0 + u : ntile - 1;
                const int ti = t % TR, tj = t / TR;
                const float *ccol = Ak + static_cast<int64_t>(W + tj * 32 + l32) * lda + W + ti * 32 + 4 * lh;
#pragma unroll
                for (int q = 0; q < 4; ++q) v[u][q] = *reinterpret_cast<const float4 *>(ccol + 8 * q);
            }
        };
        float4 cv[TPW][4];
        if (wave * TPW < ntile) load_c(wave * TPW, cv);
        for (int t0 = wave * TPW; t0 < ntile; t0 += STEP) {
            float4 cn[TPW][4];
            if (t0 + STEP < ntile) load_c(t0 + STEP, cn);
#pragma unroll
            for (int u = 0; u < TPW; ++u) {
                const int t = t0 + u;
                if (t >= ntile) break;
                const int ti = t % TR, tj = t / TR;
                floatx16g_t acc;
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) {
                    const int kr = 2 * kk + lh;
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(Lc[kr * kLcS + W + ti * 32 + l32],
                                                               Us[(tj * 32 + l32) * 33 + kr], acc, 0, 0, 0);
                }
                float *ccol = Ak + static_cast<int64_t>(W + tj * 32 + l32) * lda + W + ti * 32 + 4 * lh;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    float4 c4 = cv[u][q];
                    c4.x -= acc[4 * q + 0];
                    c4.y -= acc[4 * q + 1];
                    c4.z -= acc[4 * q + 2];
                    c4.w -= acc[4 * q + 3];
                    *reinterpret_cast<float4 *>(ccol + 8 * q) = c4;
                }
            }
#pragma unroll
            for (int u = 0; u < TPW; ++u)
#pragma unroll
                for (int q = 0; q < 4; ++q) cv[u][q] = cn[u][q];
        }
        __syncthreads();
    }
    if (tid == 0 && info != nullptr) *info = bad;
}

// ---------------------------------------------------------------------------
// L panel TRSM: L(m x b) := L U^-1, U upper non-unit (update_left_panels,
// panel_trf.hpp:143-157 -> strsm "R","U","N","N").  One thread per row; the
// row is solved in column blocks of CB held in registers; each block's
// coefficients U(0:jb+CB, jb:jb+CB) are staged in LDS.  Per element the
// subtraction order is l = 0..j-1 followed by the reciprocal multiply, the
// netlib order.
// ---------------------------------------------------------------------------
constexpr int kTrsmCB = 32;
constexpr int kTrsmMaxB = 1024;

__global__ __launch_bounds__(256) void trsm_l_kernel(int64_t m, int b, const float *__restrict__ U,
                                                     int64_t ldu, float *__restrict__ L,
                                                     int64_t ldl, float *__restrict__ amax, int tb)
{
    __shared__ float Ut[kTrsmMaxB / 2 * kTrsmCB];   // Ut[l * CB + c], l < 512 per pass
    __shared__ float Dinv[kTrsmCB];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const bool active = i < m;
    float *row = L + (active ? i : 0);
    float vmax = 0.0f;
    for (int jb = 0; jb < b; jb += kTrsmCB) {
        const int w = b - jb < kTrsmCB ? b - jb : kTrsmCB;
        float x[kTrsmCB];
#pragma unroll
        for (int c = 0; c < kTrsmCB; ++c)
            x[c] = (active && c < w) ? row[static_cast<int64_t>(jb + c) * ldl] : 0.0f;
        // contributions of finished columns, staged 512 rows of U at a time
        for (int l0 = 0; l0 < jb + w; l0 += kTrsmMaxB / 2) {
            const int l1 = (jb + w) - l0 < kTrsmMaxB / 2 ? (jb + w) : l0 + kTrsmMaxB / 2;
            __syncthreads();
            for (int idx = threadIdx.x; idx < (l1 - l0) * kTrsmCB; idx += 256) {
                int l = idx / kTrsmCB, c = idx % kTrsmCB;
                Ut[idx] = c < w ? U[static_cast<int64_t>(jb + c) * ldu + l0 + l] : 0.0f;
            }
            if (l0 == 0 && threadIdx.x < kTrsmCB)
                Dinv[threadIdx.x] = threadIdx.x < w
                    ? 1.0f / U[static_cast<int64_t>(jb + threadIdx.x) * ldu + jb + threadIdx.x]
                    : 0.0f;
            __syncthreads();
            const int lend = l1 < jb ? l1 : jb;
            for (int l = l0; l < lend; ++l) {
                const float xl = active ? row[static_cast<int64_t>(l) * ldl] : 0.0f;
                const float *ur = Ut + (l - l0) * kTrsmCB;
#pragma unroll
                for (int c = 0; c < kTrsmCB; ++c) x[c] = x[c] - xl * ur[c];
            }
            if (l1 == jb + w) {
                // diagonal block: rows jb..jb+w of U are at Ut[(jb - l0 + r) * CB + c]
                const float *ud = Ut + (jb - l0) * kTrsmCB;
#pragma unroll
                for (int c = 0; c < kTrsmCB; ++c) {
#pragma unroll
                    for (int l = 0; l < c; ++l) x[c] = x[c] - x[l] * ud[l * kTrsmCB + c];
                    x[c] = x[c] * Dinv[c];
                }
            }
        }
#pragma unroll
        for (int c = 0; c < kTrsmCB; ++c)
            if (active && c < w) {
                row[static_cast<int64_t>(jb + c) * ldl] = x[c];
                vmax = fmaxf(vmax, fabsf(x[c]));
            }
    }
    tile_max_reduce(amax, active ? i / tb : 0, vmax, active);
}

// ---------------------------------------------------------------------------
// U panel TRSM: U(b x n) := Lkk^-1 U, Lkk unit lower (update_right_panels,
// panel_trf.hpp:158-171 -> strsm "L","L","N","U").  One thread per column,
// rows solved in blocks of CB in registers; Lkk coefficients are wave-uniform
// loads.  Element i subtracts rows k = 0..i-1 in order (netlib order).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void trsm_u_kernel(int b, int64_t n, const float *__restrict__ Lk,
                                                     int64_t ldl, float *__restrict__ U,
                                                     int64_t ldu, float *__restrict__ amax, int tb)
{
    const int64_t j = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const bool active = j < n;
    float *col = U + (active ? j : 0) * ldu;
    float vmax = 0.0f;
    for (int ib = 0; ib < b; ib += kTrsmCB) {
        const int w = b - ib < kTrsmCB ? b - ib : kTrsmCB;
        float x[kTrsmCB];
#pragma unroll
        for (int r = 0; r < kTrsmCB; ++r) x[r] = (active && r < w) ? col[ib + r] : 0.0f;
        for (int l = 0; l < ib; ++l) {
            const float xl = active ? col[l] : 0.0f;
            const float *lc = Lk + static_cast<int64_t>(l) * ldl + ib;
#pragma unroll
            for (int r = 0; r < kTrsmCB; ++r)
                if (r < w) x[r] = x[r] - xl * lc[r];
        }
#pragma unroll
        for (int r = 0; r < kTrsmCB; ++r) {
            const float *lc = Lk + static_cast<int64_t>(ib + r) * ldl + ib;
#pragma unroll
            for (int q = r + 1; q < kTrsmCB; ++q)
                if (q < w) x[q] = x[q] - x[r] * lc[q];
        }
#pragma unroll
        for (int r = 0; r < kTrsmCB; ++r)
            if (active && r < w) {
                col[ib + r] = x[r];
                vmax = fmaxf(vmax, fabsf(x[r]));
            }
    }
    tile_max_reduce(amax, active ? j / tb : 0, vmax, active);
}

// ---------------------------------------------------------------------------
// per-tile max|x| of a panel (side 0: m x k L panel, tiles by row / tb;
// side 1: k x m U panel, tiles by column / tb)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void tile_amax_kernel(int side, int64_t m, int64_t k, int tb,
                                                        const float *__restrict__ src, int64_t lds,
                                                        float *__restrict__ amax)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const bool active = t < m;
    float v = 0.0f;
    if (active) {
        if (side == 0)
            for (int64_t c = 0; c < k; ++c) v = fmaxf(v, fabsf(src[c * lds + t]));
        else
            for (int64_t r = 0; r < k; ++r) v = fmaxf(v, fabsf(src[t * lds + r]));
    }
    tile_max_reduce(amax, active ? t / tb : 0, v, active);
}

__global__ void fill_f32_kernel(float *p, int64_t cnt, float v)
{
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < cnt) p[t] = v;
}

// ---------------------------------------------------------------------------
// Quantizer (convert_left_panels / convert_right_panels, panel_trf.hpp:173-233)
// q = round(x * 2^e) with e from the tile's max (0 for plain FP16).
// side 0: L panel m x k col-major -> dst[i*ldd + kk] (transposed through LDS)
// side 1: U panel k x m col-major -> dst[j*ldd + kk]
// ---------------------------------------------------------------------------
template <int FMT>
__device__ __forceinline__ int fmt_exponent(const float *amax, int64_t tile)
{
    if (FMT == HPLAI_FMT_F16) return 0;
    const float a = amax[tile];
    return FMT == HPLAI_FMT_E4M3 ? e4m3_exponent(a) : tile_exponent(a, kF16MrFrac, kF16Er);
}

template <int FMT, bool PACKED>
__global__ __launch_bounds__(256) void quant_l_kernel(int64_t m, int64_t k, int tb,
                                                      const float *__restrict__ src, int64_t lds,
                                                      void *__restrict__ dst, int64_t ldd,
                                                      const float *__restrict__ amax,
                                                      int32_t *__restrict__ exps)
{
    __shared__ float T[64][65];
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * 64;
    const int64_t c0 = static_cast<int64_t>(blockIdx.y) * 64;
    const int tx = threadIdx.x % 64, ty = threadIdx.x / 64;
    for (int c = ty; c < 64; c += 4) {
        const int64_t gr = r0 + tx, gc = c0 + c;
        T[c][tx] = (gr < m && gc < k) ? src[gc * lds + gr] : 0.0f;
    }
    __syncthreads();
    // each thread writes 4 consecutive k values of one row (16 rows x 16 chunks per pass)
    for (int p = threadIdx.x; p < 64 * 16; p += 256) {
        const int r = p / 16, kc = (p % 16) * 4;
        const int64_t gr = r0 + r;
        if (gr >= m || c0 + kc >= k) continue;
        const int e = fmt_exponent<FMT>(amax, gr / tb);
        if (exps != nullptr && (gr % tb) == 0 && c0 == 0 && kc == 0) exps[gr / tb] = e;
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = ldexp_fast(T[kc + q][r], e);
        const int nk = k - (c0 + kc) < 4 ? static_cast<int>(k - (c0 + kc)) : 4;
        if (FMT == HPLAI_FMT_E4M3) {
            uint8_t *d8 = static_cast<uint8_t *>(dst) + (PACKED ? packed_off(gr, c0 + kc, ldd) : gr * ldd + c0 + kc);
            uint32_t w = f32x4_to_e4m3x4(v[0], v[1], v[2], v[3]);
            if (nk == 4 && (reinterpret_cast<uintptr_t>(d8) & 3) == 0)
                *reinterpret_cast<uint32_t *>(d8) = w;   // one 4-byte store
            else
                for (int q = 0; q < nk; ++q) d8[q] = static_cast<uint8_t>(w >> (8 * q));
        } else {
            uint16_t *d16 = PACKED ? reinterpret_cast<uint16_t *>(static_cast<uint8_t *>(dst) +
                                                                packed_off(gr, 2 * (c0 + kc), 2 * ldd))
                                   : static_cast<uint16_t *>(dst) + gr * ldd + c0 + kc;
            if (nk == 4 && (reinterpret_cast<uintptr_t>(d16) & 7) == 0) {
                uint2 w;
                w.x = f32_to_f16_bits(v[0]) | (static_cast<uint32_t>(f32_to_f16_bits(v[1])) << 16);
                w.y = f32_to_f16_bits(v[2]) | (static_cast<uint32_t>(f32_to_f16_bits(v[3])) << 16);
                *reinterpret_cast<uint2 *>(d16) = w;     // one 8-byte store
            } else {
                for (int q = 0; q < nk; ++q) d16[q] = f32_to_f16_bits(v[q]);
            }
        }
    }
}

template <int FMT, bool PACKED>
__global__ __launch_bounds__(256) void quant_u_kernel(int64_t m, int64_t k, int tb,
                                                      const float *__restrict__ src, int64_t lds,
                                                      void *__restrict__ dst, int64_t ldd,
                                                      const float *__restrict__ amax,
                                                      int32_t *__restrict__ exps)
{
    // thread handles 4 consecutive rows (k index) of one column j
    const int64_t kq = ceil_div(k, 4);
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int64_t j = t / kq, r = (t % kq) * 4;
    if (j >= m) return;
    const int e = fmt_exponent<FMT>(amax, j / tb);
    if (exps != nullptr && r == 0 && (j % tb) == 0) exps[j / tb] = e;
    const int nk = k - r < 4 ? static_cast<int>(k - r) : 4;
    float v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = q < nk ? ldexp_fast(src[j * lds + r + q], e) : 0.0f;
    if (FMT == HPLAI_FMT_E4M3) {
        uint8_t *d8 = static_cast<uint8_t *>(dst) + (PACKED ? packed_off(j, r, ldd) : j * ldd + r);
        uint32_t w = f32x4_to_e4m3x4(v[0], v[1], v[2], v[3]);
        if (nk == 4 && (reinterpret_cast<uintptr_t>(d8) & 3) == 0)
            *reinterpret_cast<uint32_t *>(d8) = w;
        else
            for (int q = 0; q < nk; ++q) d8[q] = static_cast<uint8_t>(w >> (8 * q));
    } else {
        uint16_t *d16 = PACKED ? reinterpret_cast<uint16_t *>(static_cast<uint8_t *>(dst) + packed_off(j, 2 * r, 2 * ldd))
                               : static_cast<uint16_t *>(dst) + j * ldd + r;
        if (nk == 4 && (reinterpret_cast<uintptr_t>(d16) & 7) == 0) {
            uint2 w;
            w.x = f32_to_f16_bits(v[0]) | (static_cast<uint32_t>(f32_to_f16_bits(v[1])) << 16);
            w.y = f32_to_f16_bits(v[2]) | (static_cast<uint32_t>(f32_to_f16_bits(v[3])) << 16);
            *reinterpret_cast<uint2 *>(d16) = w;
        } else {
            for (int q = 0; q < nk; ++q) d16[q] = f32_to_f16_bits(v[q]);
        }
    }
}

// 16 consecutive K values of one row scaled by 2^e and stored at d (16 bytes
// of e4m3 or 32 bytes of FP16; d 16-byte aligned).
template <int FMT>
__device__ __forceinline__ void store_q16(uint8_t *d, const float (&v)[16], int e)
{
    float x[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) x[q] = ldexp_fast(v[q], e);
    if constexpr (FMT == HPLAI_FMT_E4M3) {
        uint4 w;
        w.x = f32x4_to_e4m3x4(x[0], x[1], x[2], x[3]);
        w.y = f32x4_to_e4m3x4(x[4], x[5], x[6], x[7]);
        w.z = f32x4_to_e4m3x4(x[8], x[9], x[10], x[11]);
        w.w = f32x4_to_e4m3x4(x[12], x[13], x[14], x[15]);
        *reinterpret_cast<uint4 *>(d) = w;
    } else {
        uint32_t h[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            h[q] = f32_to_f16_bits(x[2 * q]) | (static_cast<uint32_t>(f32_to_f16_bits(x[2 * q + 1])) << 16);
        *reinterpret_cast<uint4 *>(d) = uint4{h[0], h[1], h[2], h[3]};
        *reinterpret_cast<uint4 *>(d + 16) = uint4{h[4], h[5], h[6], h[7]};
    }
}

// The same quantizers with 16 values per thread for aligned shapes (k % 16 == 0,
// 16-byte aligned source columns and destination rows): 64-byte loads and one
// 16-byte (e4m3) or 32-byte (FP16) store per thread.
// L panel: a 64 x 64 block through LDS; thread p takes row p % 64 of the block
// and K values 16 (p / 64) .. + 15, so a half-wave reads 32 distinct banks.
template <int FMT, bool PACKED>
__global__ __launch_bounds__(256) void quant_l16_kernel(int64_t m, int64_t k, int tb,
                                                        const float *__restrict__ src, int64_t lds,
                                                        void *__restrict__ dst, int64_t ldd,
                                                        const float *__restrict__ amax,
                                                        int32_t *__restrict__ exps)
{
    __shared__ float T[64][65];
    constexpr int es = FMT == HPLAI_FMT_E4M3 ? 1 : 2;
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * 64;
    const int64_t c0 = static_cast<int64_t>(blockIdx.y) * 64;
    const int tid = threadIdx.x;
    // loads: 64 columns x 16 float4; thread tid takes float4 i = tid % 16 of
    // columns tid / 16 + 16 u
    {
        const int i = tid & 15;
        const int64_t gr = r0 + 4 * i;
        float4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = (tid >> 4) + 16 * u;
            const int64_t gc = c0 + c < k ? c0 + c : k - 1;
            const float *col = src + gc * lds;
            if (gr + 3 < m) {
                v[u] = *reinterpret_cast<const float4 *>(col + gr);
            } else {
                v[u].x = gr < m ? col[gr] : 0.0f;
                v[u].y = gr + 1 < m ? col[gr + 1] : 0.0f;
                v[u].z = gr + 2 < m ? col[gr + 2] : 0.0f;
                v[u].w = 0.0f;
            }
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = (tid >> 4) + 16 * u;
            T[c][4 * i] = v[u].x;
            T[c][4 * i + 1] = v[u].y;
            T[c][4 * i + 2] = v[u].z;
            T[c][4 * i + 3] = v[u].w;
        }
    }
    __syncthreads();
    const int r = tid & 63, kc = (tid >> 6) * 16;
    const int64_t gr = r0 + r;
    if (gr >= m || c0 + kc >= k) return;
    const int e = fmt_exponent<FMT>(amax, gr / tb);
    if (exps != nullptr && (gr % tb) == 0 && c0 == 0 && kc == 0) exps[gr / tb] = e;
    float v[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) v[q] = T[kc + q][r];
    uint8_t *d = static_cast<uint8_t *>(dst) +
                 (PACKED ? packed_off(gr, es * (c0 + kc), es * ldd) : (gr * ldd + c0 + kc) * es);
    store_q16<FMT>(d, v, e);
}

// U panel: thread t takes K values 16 (t % (k / 16)) .. + 15 of column t / (k / 16).
template <int FMT, bool PACKED>
__global__ __launch_bounds__(256) void quant_u16_kernel(int64_t m, int64_t k, int tb,
                                                        const float *__restrict__ src, int64_t lds,
                                                        void *__restrict__ dst, int64_t ldd,
                                                        const float *__restrict__ amax,
                                                        int32_t *__restrict__ exps)
{
    constexpr int es = FMT == HPLAI_FMT_E4M3 ? 1 : 2;
    const int64_t kq = k / 16;
    const int64_t t = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    const int64_t j = t / kq, r = (t % kq) * 16;
    if (j >= m) return;
    const int e = fmt_exponent<FMT>(amax, j / tb);
    if (exps != nullptr && r == 0 && (j % tb) == 0) exps[j / tb] = e;
    const float4 *s4 = reinterpret_cast<const float4 *>(src + j * lds + r);
    float v[16];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 x = s4[q];
        v[4 * q] = x.x;
        v[4 * q + 1] = x.y;
        v[4 * q + 2] = x.z;
        v[4 * q + 3] = x.w;
    }
    uint8_t *d = static_cast<uint8_t *>(dst) + (PACKED ? packed_off(j, es * r, es * ldd) : (j * ldd + r) * es);
    store_q16<FMT>(d, v, e);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int launch_getrf(int n, float *A, int64_t lda, int *info, hipStream_t s)
{
    if (n <= 0) return HPLAI_OK;
    const bool mfma_ok = n % 32 == 0 && lda % 4 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0;
    if (n <= 512 && mfma_ok)
        hipLaunchKernelGGL(getrf_mfma_kernel, dim3(1), dim3(kGetrfThreads), 0, s, n, A, lda, info);
    else if (n <= 512)
        hipLaunchKernelGGL((getrf_kernel<512, 32>), dim3(1), dim3(kGetrfThreads), 0, s, n, A, lda, info);
    else
        hipLaunchKernelGGL((getrf_kernel<1024, 16>), dim3(1), dim3(kGetrfThreads), 0, s, n, A, lda, info);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

bool trsm_strip_ok(int b, const float *amax, int tb);
int launch_trsm_l_strip(int64_t m, int b, const float *U, int64_t ldu, const float *inv, float *L,
                        int64_t ldl, float *amax, int tb, hipStream_t s);
int launch_trsm_u_strip(int b, int64_t n, const float *Lkk, int64_t ldl, const float *inv, float *U,
                        int64_t ldu, float *amax, int tb, hipStream_t s);

// Strip TRSMs need the diagonal sub-block inverses; compute them into a
// temporary when the caller did not supply them.
template <typename F>
static int with_inv(int b, const float *LU, int64_t ld, const float *inv, hipStream_t s, F &&f)
{
    if (inv != nullptr) return f(inv);
    float *tmp = nullptr;
    HPLAI_HIP(hipMallocAsync(reinterpret_cast<void **>(&tmp), sizeof(float) * diag_inv_floats(b), s));
    int rc = launch_diag_inv(b, LU, ld, tmp, s);
    if (rc == HPLAI_OK) rc = f(tmp);
    HPLAI_HIP(hipFreeAsync(tmp, s));
    return rc;
}

int launch_trsm_l(int64_t m, int b, const float *U, int64_t ldu, float *L, int64_t ldl, float *amax,
                  int tb, const float *inv, hipStream_t s)
{
    if (m <= 0 || b <= 0) return HPLAI_OK;
    if (trsm_strip_ok(b, amax, tb))
        return with_inv(b, U, ldu, inv, s, [&](const float *iv) {
            return launch_trsm_l_strip(m, b, U, ldu, iv, L, ldl, amax, tb, s);
        });
    hipLaunchKernelGGL(trsm_l_kernel, dim3(static_cast<unsigned>(ceil_div(m, 256))), dim3(256), 0, s,
                       m, b, U, ldu, L, ldl, amax, tb);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

int launch_trsm_u(int b, int64_t n, const float *Lkk, int64_t ldl, float *U, int64_t ldu,
                  float *amax, int tb, const float *inv, hipStream_t s)
{
    if (n <= 0 || b <= 0) return HPLAI_OK;
    if (trsm_strip_ok(b, amax, tb))
        return with_inv(b, Lkk, ldl, inv, s, [&](const float *iv) {
            return launch_trsm_u_strip(b, n, Lkk, ldl, iv, U, ldu, amax, tb, s);
        });
    hipLaunchKernelGGL(trsm_u_kernel, dim3(static_cast<unsigned>(ceil_div(n, 256))), dim3(256), 0, s,
                       b, n, Lkk, ldl, U, ldu, amax, tb);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

int launch_tile_amax(int side, int64_t m, int64_t k, int tb, const float *src, int64_t lds,
                     float *amax, hipStream_t s)
{
    if (m <= 0 || k <= 0) return HPLAI_OK;
    hipLaunchKernelGGL(tile_amax_kernel, dim3(static_cast<unsigned>(ceil_div(m, 256))), dim3(256), 0,
                       s, side, m, k, tb, src, lds, amax);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

int launch_fill_f32(float *p, int64_t cnt, float v, hipStream_t s)
{
    if (cnt <= 0) return HPLAI_OK;
    hipLaunchKernelGGL(fill_f32_kernel, dim3(static_cast<unsigned>(ceil_div(cnt, 256))), dim3(256), 0,
                       s, p, cnt, v);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

template <int FMT, bool PACKED>
static int quant_dispatch(int side, int64_t m, int64_t k, int tb, const float *src, int64_t lds,
                          void *dst, int64_t ldd, const float *amax, int32_t *exps, hipStream_t s)
{
    constexpr int es = FMT == HPLAI_FMT_E4M3 ? 1 : 2;
    const bool v16 = k % 16 == 0 && lds % 4 == 0 && reinterpret_cast<uintptr_t>(src) % 16 == 0 &&
                     reinterpret_cast<uintptr_t>(dst) % 16 == 0 && (ldd * es) % 16 == 0;
    if (side == 0 && v16) {
        dim3 grid(static_cast<unsigned>(ceil_div(m, 64)), static_cast<unsigned>(ceil_div(k, 64)));
        hipLaunchKernelGGL((quant_l16_kernel<FMT, PACKED>), grid, dim3(256), 0, s, m, k, tb, src, lds, dst, ldd,
                           amax, exps);
    } else if (side == 0) {
        dim3 grid(static_cast<unsigned>(ceil_div(m, 64)), static_cast<unsigned>(ceil_div(k, 64)));
        hipLaunchKernelGGL((quant_l_kernel<FMT, PACKED>), grid, dim3(256), 0, s, m, k, tb, src, lds, dst, ldd,
                           amax, exps);
    } else if (v16) {
        const int64_t thr = m * (k / 16);
        hipLaunchKernelGGL((quant_u16_kernel<FMT, PACKED>), dim3(static_cast<unsigned>(ceil_div(thr, 256))),
                           dim3(256), 0, s, m, k, tb, src, lds, dst, ldd, amax, exps);
    } else {
        const int64_t thr = m * ceil_div(k, 4);
        hipLaunchKernelGGL((quant_u_kernel<FMT, PACKED>), dim3(static_cast<unsigned>(ceil_div(thr, 256))),
                           dim3(256), 0, s, m, k, tb, src, lds, dst, ldd, amax, exps);
    }
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

template <int FMT>
static int quant_dispatch(bool packed, int side, int64_t m, int64_t k, int tb, const float *src, int64_t lds,
                          void *dst, int64_t ldd, const float *amax, int32_t *exps, hipStream_t s)
{
    return packed ? quant_dispatch<FMT, true>(side, m, k, tb, src, lds, dst, ldd, amax, exps, s)
                  : quant_dispatch<FMT, false>(side, m, k, tb, src, lds, dst, ldd, amax, exps, s);
}

int launch_quantize(int fmt, int side, int64_t m, int64_t k, int tb, const float *src, int64_t lds,
                    void *dst, int64_t ldd, const float *amax, int32_t *exps, hipStream_t s, bool packed)
{
    if (m <= 0 || k <= 0) return HPLAI_OK;
    switch (fmt) {
    case HPLAI_FMT_F16:
        return quant_dispatch<HPLAI_FMT_F16>(packed, side, m, k, tb, src, lds, dst, ldd, amax, exps, s);
    case HPLAI_FMT_F16_TILE:
        return quant_dispatch<HPLAI_FMT_F16_TILE>(packed, side, m, k, tb, src, lds, dst, ldd, amax, exps, s);
    case HPLAI_FMT_E4M3:
        return quant_dispatch<HPLAI_FMT_E4M3>(packed, side, m, k, tb, src, lds, dst, ldd, amax, exps, s);
    default:
        set_error("quantize: unknown format %d", fmt);
        return HPLAI_EINVAL;
    }
}

}  // namespace hplai
