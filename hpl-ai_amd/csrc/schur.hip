// schur.hip — the low-precision Schur trailing update on CDNA4 MFMA.
//
//   C(M x N, FP32, column-major) -= sum_s 2^-(eL_s + eU_s) * Lq_s * Uq_s
//
// Replaces gemmschur -> hgemm/shgemm (schur_updator.hpp:9-47, fp16sim.hpp:197-233,
// kernels/avx2.cpp:82-302).  Both operands are K-contiguous ("TN"): row i of
// Lq is Lq[i*ldl + k], column j of Uq is Uq[j*ldu + k], so every MFMA lane
// reads 16 contiguous bytes.  FP32 accumulation in the MFMA accumulators.
//
// FP16 : v_mfma_f32_32x32x16_f16.  Per-slice tile exponents (HPLAI_FMT_F16_TILE)
//        are applied by rescaling the accumulators by an exact power of two when
//        the exponent changes between K slices.
// E4M3 : v_mfma_scale_f32_32x32x64_f8f6f4 with the tile exponents as E8M0 block
//        scales, so the products come out unscaled.
//
// Workgroup: 256 threads (4 waves, 2 x 2), tile 128 x 128, each wave 64 x 64 as
// 2 x 2 MFMA blocks of 32 x 32.  128 bytes of K per row per stage for both
// formats (64 FP16 / 128 FP8), double-buffered in LDS with a 16-byte row pad.
#include "device_common.hpp"
#include "internal.hpp"

namespace hplai {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef int intx8_t __attribute__((ext_vector_type(8)));
typedef int intx4_t __attribute__((ext_vector_type(4)));

constexpr int kBM = 128, kBN = 128, kThreads = 256;
constexpr int kRowBytes = 128;               // K bytes per row per stage
constexpr int kLdsRow = kRowBytes + 16;      // padded LDS row (bank spread)

template <int FMT>
struct SchurFmt {
    static constexpr int es = (FMT == HPLAI_FMT_E4M3) ? 1 : 2;   // bytes per element
    static constexpr int bk = kRowBytes / es;                     // K elements per stage
};

// Hardware places workgroup w on XCD w % 8; give each XCD a contiguous range
// of logical tiles so neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int w, int nwg)
{
    const int xcd = w & 7, idx = w >> 3;
    const int q = nwg >> 3, r = nwg & 7;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

template <int FMT, bool VEC_C>
__global__ __launch_bounds__(kThreads) void schur_kernel(
    int64_t M, int64_t N, int nslices, int kslice, const uint8_t *__restrict__ Lq, int64_t ldl_b,
    const int32_t *__restrict__ eL, int64_t ldel, const uint8_t *__restrict__ Uq, int64_t ldu_b,
    const int32_t *__restrict__ eU, int64_t ldeu, int tb, float *__restrict__ C, int64_t ldc,
    int tiles_m, int tiles_n)
{
    using F = SchurFmt<FMT>;
    __shared__ __attribute__((aligned(16))) uint8_t As[2][kBM * kLdsRow];
    __shared__ __attribute__((aligned(16))) uint8_t Bs[2][kBN * kLdsRow];

    // ---- tile coordinates: XCD remap, then groups of 8 tile-rows ----
    const int nwg = tiles_m * tiles_n;
    const int logical = xcd_remap(static_cast<int>(blockIdx.x), nwg);
    constexpr int G = 8;
    const int per_group = G * tiles_n;
    const int first_m = (logical / per_group) * G;
    const int gsize = tiles_m - first_m < G ? tiles_m - first_m : G;
    const int in_group = logical % per_group;
    const int tm = first_m + in_group % gsize;
    const int tn = in_group / gsize;
    const int64_t m0 = static_cast<int64_t>(tm) * kBM;
    const int64_t n0 = static_cast<int64_t>(tn) * kBN;

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave & 1, wn = wave >> 1;
    const int lrow = lane & 31, lgrp = lane >> 5;

    // ---- global -> register staging: 4 x 16 B per operand per thread ----
    intx4_t ra[4], rb[4];
    const int64_t kbytes_total = static_cast<int64_t>(nslices) * kslice * F::es;
    const int nstages = static_cast<int>(kbytes_total / kRowBytes);

    auto gload = [&](int stage) {
        const int64_t kb = static_cast<int64_t>(stage) * kRowBytes;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = tid + kThreads * q;
            const int row = c >> 3, part = c & 7;
            const int64_t gm = m0 + row, gn = n0 + row;
            ra[q] = gm < M ? *reinterpret_cast<const intx4_t *>(Lq + gm * ldl_b + kb + part * 16)
                           : intx4_t{0, 0, 0, 0};
            rb[q] = gn < N ? *reinterpret_cast<const intx4_t *>(Uq + gn * ldu_b + kb + part * 16)
                           : intx4_t{0, 0, 0, 0};
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int c = tid + kThreads * q;
            const int row = c >> 3, part = c & 7;
            *reinterpret_cast<intx4_t *>(&As[buf][row * kLdsRow + part * 16]) = ra[q];
            *reinterpret_cast<intx4_t *>(&Bs[buf][row * kLdsRow + part * 16]) = rb[q];
        }
    };

    floatx16_t acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int row_tile = tb > 0 ? static_cast<int>(m0 / tb) : 0;
    const int col_tile = tb > 0 ? static_cast<int>(n0 / tb) : 0;
    const int stages_per_slice = kslice * F::es / kRowBytes;
    int e_cur = 0;          // FP16 tile mode: exponent the accumulators are scaled by
    int scale_a = 127, scale_b = 127;   // E4M3: E8M0 block scales (127 = 2^0)

    gload(0);
    sstore(0);
    __syncthreads();
    for (int ks = 0; ks < nstages; ++ks) {
        const int buf = ks & 1;
        if (ks + 1 < nstages) gload(ks + 1);
        if (ks % stages_per_slice == 0 && eL != nullptr) {
            const int s = ks / stages_per_slice;
            const int el = eL[s * ldel + row_tile], eu = eU[s * ldeu + col_tile];
            if (FMT == HPLAI_FMT_E4M3) {
                scale_a = 127 - el;
                scale_b = 127 - eu;
            } else if (FMT == HPLAI_FMT_F16_TILE) {
                const int e_new = el + eu;
                if (ks == 0) {
                    e_cur = e_new;
                } else if (e_new != e_cur) {
                    const int d = e_new - e_cur;
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j)
#pragma unroll
                            for (int r = 0; r < 16; ++r) acc[i][j][r] = ldexp_fast(acc[i][j][r], d);
                    e_cur = e_new;
                }
            }
        }
        const uint8_t *Ab = As[buf];
        const uint8_t *Bb = Bs[buf];
        if (FMT == HPLAI_FMT_E4M3) {
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                intx8_t a[2], b[2];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const uint8_t *p = Ab + (wm * 64 + i * 32 + lrow) * kLdsRow + kk * 64 + lgrp * 32;
                    intx4_t lo = *reinterpret_cast<const intx4_t *>(p);
                    intx4_t hi = *reinterpret_cast<const intx4_t *>(p + 16);
                    a[i] = intx8_t{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                }
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const uint8_t *p = Bb + (wn * 64 + j * 32 + lrow) * kLdsRow + kk * 64 + lgrp * 32;
                    intx4_t lo = *reinterpret_cast<const intx4_t *>(p);
                    intx4_t hi = *reinterpret_cast<const intx4_t *>(p + 16);
                    b[j] = intx8_t{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                            a[i], b[j], acc[i][j], 0, 0, 0, scale_a, 0, scale_b);
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                half8_t a[2], b[2];
#pragma unroll
                for (int i = 0; i < 2; ++i)
                    a[i] = *reinterpret_cast<const half8_t *>(
                        Ab + (wm * 64 + i * 32 + lrow) * kLdsRow + kk * 32 + lgrp * 16);
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    b[j] = *reinterpret_cast<const half8_t *>(
                        Bb + (wn * 64 + j * 32 + lrow) * kLdsRow + kk * 32 + lgrp * 16);
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
        if (ks + 1 < nstages) sstore(buf ^ 1);
        __syncthreads();
    }

    // ---- epilogue: C -= acc (FP16 tile mode: acc * 2^-e_cur) ----
    // acc[i][j][4*q + r] holds row 8q + 4*lgrp + r, column lrow of block (i, j)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int64_t gc = n0 + wn * 64 + j * 32 + lrow;
            if (gc >= N) continue;
            float *ccol = C + gc * ldc;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t gr = m0 + wm * 64 + i * 32 + 8 * q + 4 * lgrp;
                float v[4];
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    v[r] = (FMT == HPLAI_FMT_F16_TILE) ? ldexp_fast(acc[i][j][4 * q + r], -e_cur)
                                                        : acc[i][j][4 * q + r];
                if (VEC_C && gr + 3 < M) {
                    float4 *p = reinterpret_cast<float4 *>(ccol + gr);
                    float4 c4 = *p;
                    c4.x -= v[0];
                    c4.y -= v[1];
                    c4.z -= v[2];
                    c4.w -= v[3];
                    *p = c4;
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (gr + r < M) ccol[gr + r] -= v[r];
                }
            }
        }
}

template <int FMT>
static int schur_dispatch(int64_t m, int64_t n, int nslices, int kslice, const void *Lq,
                          int64_t ldl, const int32_t *eL, int64_t ldel, const void *Uq, int64_t ldu,
                          const int32_t *eU, int64_t ldeu, int tb, float *C, int64_t ldc,
                          hipStream_t s)
{
    using F = SchurFmt<FMT>;
    const int64_t tiles_m = ceil_div(m, kBM), tiles_n = ceil_div(n, kBN);
    if (tiles_m * tiles_n >= (int64_t(1) << 31)) {
        set_error("schur_update: %lld x %lld tiles exceed one launch", (long long)tiles_m,
                  (long long)tiles_n);
        return HPLAI_EINVAL;
    }
    const bool vec = (reinterpret_cast<uintptr_t>(C) % 16 == 0) && (ldc % 4 == 0);
    const dim3 grid(static_cast<unsigned>(tiles_m * tiles_n));
    const uint8_t *L8 = static_cast<const uint8_t *>(Lq);
    const uint8_t *U8 = static_cast<const uint8_t *>(Uq);
    if (vec)
        hipLaunchKernelGGL((schur_kernel<FMT, true>), grid, dim3(kThreads), 0, s, m, n, nslices,
                           kslice, L8, ldl * F::es, eL, ldel, U8, ldu * F::es, eU, ldeu, tb, C, ldc,
                           static_cast<int>(tiles_m), static_cast<int>(tiles_n));
    else
        hipLaunchKernelGGL((schur_kernel<FMT, false>), grid, dim3(kThreads), 0, s, m, n, nslices,
                           kslice, L8, ldl * F::es, eL, ldel, U8, ldu * F::es, eU, ldeu, tb, C, ldc,
                           static_cast<int>(tiles_m), static_cast<int>(tiles_n));
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

// Validation lives here so the C ABI and the solver share it.
int launch_schur(int fmt, int64_t m, int64_t n, int nslices, int kslice, const void *Lq,
                 int64_t ldl, const int32_t *eL, int64_t ldel, const void *Uq, int64_t ldu,
                 const int32_t *eU, int64_t ldeu, int tb, float *C, int64_t ldc, int num_cu,
                 hipStream_t s)
{
    (void)num_cu;
    if (m <= 0 || n <= 0 || nslices <= 0) return HPLAI_OK;
    const int es = fmt == HPLAI_FMT_E4M3 ? 1 : 2;
    const int bk = kRowBytes / es;
    HPLAI_CHECK(fmt == HPLAI_FMT_F16 || fmt == HPLAI_FMT_F16_TILE || fmt == HPLAI_FMT_E4M3,
                "schur_update: unknown format %d", fmt);
    HPLAI_CHECK(kslice > 0 && kslice % bk == 0,
                "schur_update: kslice %d must be a positive multiple of %d", kslice, bk);
    HPLAI_CHECK((ldl * es) % 16 == 0 && (ldu * es) % 16 == 0,
                "schur_update: ldl/ldu must be multiples of 16 bytes");
    HPLAI_CHECK(ldl >= static_cast<int64_t>(nslices) * kslice &&
                    ldu >= static_cast<int64_t>(nslices) * kslice,
                "schur_update: ldl/ldu smaller than K");
    HPLAI_CHECK(reinterpret_cast<uintptr_t>(Lq) % 16 == 0 && reinterpret_cast<uintptr_t>(Uq) % 16 == 0,
                "schur_update: Lq/Uq must be 16-byte aligned");
    HPLAI_CHECK(ldc >= m, "schur_update: ldc < m");
    if (fmt != HPLAI_FMT_F16) {
        HPLAI_CHECK(eL != nullptr && eU != nullptr, "schur_update: exponent arrays required");
        HPLAI_CHECK(tb > 0 && tb % kBM == 0, "schur_update: tile %d must be a multiple of %d", tb, kBM);
    }
    switch (fmt) {
    case HPLAI_FMT_F16:
        return schur_dispatch<HPLAI_FMT_F16>(m, n, nslices, kslice, Lq, ldl, nullptr, 0, Uq, ldu,
                                             nullptr, 0, 0, C, ldc, s);
    case HPLAI_FMT_F16_TILE:
        return schur_dispatch<HPLAI_FMT_F16_TILE>(m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu,
                                                  eU, ldeu, tb, C, ldc, s);
    default:
        return schur_dispatch<HPLAI_FMT_E4M3>(m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu, eU,
                                              ldeu, tb, C, ldc, s);
    }
}

}  // namespace hplai
