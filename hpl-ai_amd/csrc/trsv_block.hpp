// trsv_block.hpp — one workgroup solves a triangular system of order <= 1024
// with FP32 factor entries and an FP64 vector (the diagonal-block step of
// panel_trsvL / panel_trsvU, panel_trsv.hpp:310-758).  Shared by the
// blocked triangular solves; the batched solver runs the same steps for
// several right-hand sides at once (batch.hip).
//
// Thread t owns rows t + 256 k and keeps their x in registers.  The triangle
// is walked in 32-column strips; every thread holds its rows' 32 entries of
// the current strip in registers and prefetches the next strip while the
// current one is used.  The 32 x 32 diagonal sub-block is solved by the
// half-wave that owns its rows: x_c is broadcast with v_readlane
// (wave-uniform lane index) and the other lanes subtract l_ic * x_c — 32
// dependent steps of a few cycles each.  The strip's final x goes through
// LDS to the threads owning the rows further on.  Non-unit diagonals are
// applied as a multiply by 1 / t_cc computed before the dependent chain (one
// rounding different from a division).
#pragma once
#include "device_common.hpp"

namespace hplai {

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// strip loader: dst[k][c] = element (row t + 256 k, column col0 + c) of the
// block, clamped in-bounds; element (r, c) is M[c*lda + r], or M[r*lda + c]
// for the transposed factor (TRANS)
template <int RPT, bool TRANS>
__device__ __forceinline__ void trsv_load_strip(const float *__restrict__ M, int64_t lda, int bs,
                                                int col0, int tid, float (&dst)[RPT][32])
{
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = min(tid + 256 * k, bs - 1);
#pragma unroll
        for (int c = 0; c < 32; ++c) {
            const int64_t col = min(col0 + c, bs - 1);
            dst[k][c] = TRANS ? M[static_cast<int64_t>(r) * lda + col] : M[col * lda + r];
        }
    }
}

// Solves T x = x in place for the bs x bs triangle T (lower or upper, unit
// or not; bs <= 256 * RPT) with one 256-thread workgroup.  xs: 32 doubles of
// LDS.  Thread t reads and finally writes back only its own rows of x
// (t + 256 k), after the last workgroup barrier.
template <int RPT, bool UPPER, bool UNIT, bool TRANS>
__device__ __forceinline__ void trsv_block(int bs, const float *__restrict__ T, int64_t lda,
                                           double *__restrict__ x, double *xs)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double xr[RPT];
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = tid + 256 * k;
        xr[k] = r < bs ? x[r] : 0.0;
    }
    const int nsb = (bs + 31) / 32;
    constexpr bool PF = RPT <= 2;   // register room to prefetch the next strip
    float cur[RPT][32], nxt[PF ? RPT : 1][32];
    trsv_load_strip<RPT, TRANS>(T, lda, bs, (UPPER ? nsb - 1 : 0) * 32, tid, cur);
    for (int step = 0; step < nsb; ++step) {
        const int q = UPPER ? nsb - 1 - step : step;
        const int sb = q * 32;
        const int w = bs - sb < 32 ? bs - sb : 32;
        if constexpr (PF)
            if (step + 1 < nsb)
                trsv_load_strip<RPT, TRANS>(T, lda, bs, (UPPER ? q - 1 : q + 1) * 32, tid, nxt);
        // ---- the owner half-wave solves the 32 x 32 diagonal sub-block ----
        const int kq = sb / 256, owner = (sb % 256) / 64, half = ((sb % 256) / 32) & 1;
        if (wave == owner) {
            const int i = lane - 32 * half;                // row sb + i for the owning lanes
            const bool mine = i >= 0 && i < w;
            double v = 0.0, rinv = 1.0;
            float rowv[32];
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (k == kq) {
                    v = xr[k];
#pragma unroll
                    for (int c = 0; c < 32; ++c) rowv[c] = cur[k][c];
                }
            if (!UNIT) {
                float dg = 1.0f;
#pragma unroll
                for (int c = 0; c < 32; ++c)
                    if (c == i) dg = rowv[c];
                rinv = 1.0 / static_cast<double>(dg);
            }
            if (!UPPER) {
#pragma unroll
                for (int c = 0; c < 32; ++c) {
                    if (c < w) {
                        if (!UNIT && i == c) v *= rinv;
                        const double xc = readlane_f64(v, 32 * half + c);
                        if (mine && i > c) v = fma(-static_cast<double>(rowv[c]), xc, v);
                    }
                }
            } else {
#pragma unroll
                for (int c = 31; c >= 0; --c) {
                    if (c < w) {
                        if (!UNIT && i == c) v *= rinv;
                        const double xc = readlane_f64(v, 32 * half + c);
                        if (mine && i < c) v = fma(-static_cast<double>(rowv[c]), xc, v);
                    }
                }
            }
            if (mine) xs[i] = v;
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (k == kq && mine) xr[k] = v;
        }
        __syncthreads();
        // ---- rows further on subtract the strip's contribution ----
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + 256 * k;
            const bool further = UPPER ? (r < sb) : (r >= sb + w && r < bs);
            if (further) {
                double acc = 0.0;
#pragma unroll
                for (int c = 0; c < 32; ++c)
                    if (c < w) acc = fma(static_cast<double>(cur[k][c]), xs[c], acc);
                xr[k] -= acc;
            }
        }
        __syncthreads();
        if (step + 1 < nsb) {
            if constexpr (PF) {
#pragma unroll
                for (int k = 0; k < RPT; ++k)
#pragma unroll
                    for (int c = 0; c < 32; ++c) cur[k][c] = nxt[k][c];
            } else {
                trsv_load_strip<RPT, TRANS>(T, lda, bs, (UPPER ? q - 1 : q + 1) * 32, tid, cur);
            }
        }
    }
#pragma unroll
    for (int k = 0; k < RPT; ++k) {
        const int r = tid + 256 * k;
        if (r < bs) x[r] = xr[k];
    }
}


}  // namespace hplai
