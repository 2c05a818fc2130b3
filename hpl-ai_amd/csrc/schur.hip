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
// Two kernels, both 256 threads (4 waves, 2 x 2) with each wave computing a
// quarter of the tile as 32 x 32 MFMA blocks:
//  - schur_dma_kernel: 256 x 256 tiles for the large trailing updates;
//    operands are streamed global -> LDS by direct loads, four LDS stages;
//  - schur_kernel: 128 x 128 tiles (register-staged, double-buffered LDS) for
//    thin updates and for the per-tile FP16 form.
// Either way the kernel's HBM traffic per launch is (M + N) K bytes of
// operands plus 8 M N bytes of C (read and written once).

#include <type_traits>

#include "device_common.hpp"
#include "internal.hpp"

namespace hplai {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float floatx16_t __attribute__((ext_vector_type(16)));
typedef int intx8_t __attribute__((ext_vector_type(8)));
typedef int intx4_t __attribute__((ext_vector_type(4)));

constexpr int kThreads = 256;
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

// C is read once and written once per update, so its accesses carry the
// non-temporal hint: they stream past the caches instead of evicting the
// operand panels that the other workgroups of the XCD are still reading
// (e4m3, 32768^2 x 2048: 3.56 -> 2.85 ms; x 8192: 8.29 -> 8.14 ms).
typedef float floatx4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float4 load_c_nt(const float *p)
{
    const floatx4_t v = __builtin_nontemporal_load(reinterpret_cast<const floatx4_t *>(p));
    return float4{v[0], v[1], v[2], v[3]};
}

__device__ __forceinline__ void store_c_nt(float *p, const float4 &c)
{
    const floatx4_t v = {c.x, c.y, c.z, c.w};
    __builtin_nontemporal_store(v, reinterpret_cast<floatx4_t *>(p));
}

// C -= acc for JB accumulator blocks side by side (columns col0 + 32 j ..),
// rows row0 .. row0 + 31.  acc[j][4 q + r] is row 8 q + 4 lgrp + r, column
// lrow of block j.  With VEC_C every lane loads its 4 JB float4 first
// (addresses outside C are redirected to C itself and the values dropped),
// so the loads are in flight together; a row group cut by M (M % 4 != 0)
// takes a scalar path.
template <int JB, bool VEC_C>
__device__ __forceinline__ void epilogue_block_row(float *__restrict__ C, int64_t ldc, int64_t M, int64_t N,
                                                   int64_t row0, int64_t col0, int lrow, int lgrp,
                                                   const floatx16_t (&acc)[JB])
{
    if constexpr (VEC_C) {
        float4 cv[JB][4];
#pragma unroll
        for (int j = 0; j < JB; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t gc = col0 + j * 32 + lrow, gr = row0 + 8 * q + 4 * lgrp;
                const bool full = gc < N && gr + 3 < M;
                cv[j][q] = load_c_nt(full ? C + gc * ldc + gr : C);
            }
#pragma unroll
        for (int j = 0; j < JB; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t gc = col0 + j * 32 + lrow, gr = row0 + 8 * q + 4 * lgrp;
                if (gc < N && gr + 3 < M) {
                    float4 c4 = cv[j][q];
                    c4.x -= acc[j][4 * q + 0];
                    c4.y -= acc[j][4 * q + 1];
                    c4.z -= acc[j][4 * q + 2];
                    c4.w -= acc[j][4 * q + 3];
                    store_c_nt(C + gc * ldc + gr, c4);
                } else if (gc < N) {
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (gr + r < M) C[gc * ldc + gr + r] -= acc[j][4 * q + r];
                }
            }
    } else {
#pragma unroll
        for (int j = 0; j < JB; ++j) {
            const int64_t gc = col0 + j * 32 + lrow;
            if (gc >= N) continue;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t gr = row0 + 8 * q + 4 * lgrp;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (gr + r < M) C[gc * ldc + gr + r] -= acc[j][4 * q + r];
            }
        }
    }
}

// The same for a block row inside C (whole tile, 16-byte aligned C): the
// MFMA layout gives each lane 4 rows of one column, so a direct float4 access
// touches 32 columns per instruction with 32 bytes each; transposed through
// the LDS scratch T (32 JB columns x 36 floats, private to the wave), 8 lanes
// cover one column's 128 bytes and each instruction 8 whole columns.
constexpr int kEpiTS = 36;   // LDS column stride in floats (16-byte aligned)

template <int JB>
__device__ __forceinline__ void epilogue_block_row_lds(float *__restrict__ C, int64_t ldc, int64_t row0,
                                                       int64_t col0, int lane, float *T,
                                                       const floatx16_t (&acc)[JB])
{
    const int lrow = lane & 31, lgrp = lane >> 5;
    const int cl = lane >> 3, rq = (lane & 7) * 4;
#pragma unroll
    for (int j = 0; j < JB; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 o = {acc[j][4 * q], acc[j][4 * q + 1], acc[j][4 * q + 2], acc[j][4 * q + 3]};
            *reinterpret_cast<float4 *>(T + (j * 32 + lrow) * kEpiTS + 8 * q + 4 * lgrp) = o;
        }
    float *Cb = C + col0 * ldc + row0 + rq;
    constexpr int PASSES = JB * 32 / 8;
#pragma unroll
    for (int h = 0; h < PASSES / 8; ++h) {
        float4 cv[8];
#pragma unroll
        for (int p = 0; p < 8; ++p)
            cv[p] = load_c_nt(Cb + static_cast<int64_t>((8 * h + p) * 8 + cl) * ldc);
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const float4 a = *reinterpret_cast<const float4 *>(T + ((8 * h + p) * 8 + cl) * kEpiTS + rq);
            float4 c4 = cv[p];
            c4.x -= a.x;
            c4.y -= a.y;
            c4.z -= a.z;
            c4.w -= a.w;
            store_c_nt(Cb + static_cast<int64_t>((8 * h + p) * 8 + cl) * ldc, c4);
        }
    }
}

// The same with the trailing matrix in FP16 (16-bit trailing storage): C
// holds cscale times the partial sums, C := fp16(C - cscale * acc), each lane
// reading and writing 4 rows (8 bytes) of a column.
template <int JB>
__device__ __forceinline__ void epilogue_block_row_lds_h(uint16_t *__restrict__ C, int64_t ldc, int64_t row0,
                                                         int64_t col0, int lane, float *T,
                                                         const floatx16_t (&acc)[JB], float cscale)
{
    const int lrow = lane & 31, lgrp = lane >> 5;
    const int cl = lane >> 3, rq = (lane & 7) * 4;
#pragma unroll
    for (int j = 0; j < JB; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 o = {acc[j][4 * q], acc[j][4 * q + 1], acc[j][4 * q + 2], acc[j][4 * q + 3]};
            *reinterpret_cast<float4 *>(T + (j * 32 + lrow) * kEpiTS + 8 * q + 4 * lgrp) = o;
        }
    uint16_t *Cb = C + col0 * ldc + row0 + rq;
    constexpr int PASSES = JB * 32 / 8;
#pragma unroll
    for (int h = 0; h < PASSES / 8; ++h) {
        uint2 cv[8];
#pragma unroll
        for (int p = 0; p < 8; ++p)
            cv[p] = *reinterpret_cast<const uint2 *>(Cb + static_cast<int64_t>((8 * h + p) * 8 + cl) * ldc);
#pragma unroll
        for (int p = 0; p < 8; ++p) {
            const float4 a = *reinterpret_cast<const float4 *>(T + ((8 * h + p) * 8 + cl) * kEpiTS + rq);
            const float c0 = f16_bits_to_f32(cv[p].x & 0xffffu) - cscale * a.x;
            const float c1 = f16_bits_to_f32(cv[p].x >> 16) - cscale * a.y;
            const float c2 = f16_bits_to_f32(cv[p].y & 0xffffu) - cscale * a.z;
            const float c3 = f16_bits_to_f32(cv[p].y >> 16) - cscale * a.w;
            uint2 o;
            o.x = f32_to_f16_bits(c0) | (static_cast<uint32_t>(f32_to_f16_bits(c1)) << 16);
            o.y = f32_to_f16_bits(c2) | (static_cast<uint32_t>(f32_to_f16_bits(c3)) << 16);
            *reinterpret_cast<uint2 *>(Cb + static_cast<int64_t>((8 * h + p) * 8 + cl) * ldc) = o;
        }
    }
}

// EPI: 0 = element-wise C access (unaligned C), 1 = float4 per accumulator
// row block with ragged edges, 2 = whole tiles only (M, N multiples of the
// tile), C transposed through LDS (epilogue_block_row_lds), 3 = as 2 with an
// FP16 C scaled by cscale (epilogue_block_row_lds_h).
template <int FMT, int BM, int BN, int EPI, bool PACKED>
__global__ __launch_bounds__(kThreads) void schur_kernel(
    int64_t M, int64_t N, int nslices, int kslice, const uint8_t *__restrict__ Lq, int64_t ldl_b,
    const int32_t *__restrict__ eL, int64_t ldel, const uint8_t *__restrict__ Uq, int64_t ldu_b,
    const int32_t *__restrict__ eU, int64_t ldeu, int tb, void *__restrict__ Cv, int64_t ldc,
    int tiles_m, int tiles_n, float cscale)
{
    float *__restrict__ C = static_cast<float *>(Cv);
    using F = SchurFmt<FMT>;
    constexpr int IB = BM / 64, JB = BN / 64;          // 32 x 32 blocks per wave
    constexpr int QA = BM * 8 / kThreads, QB = BN * 8 / kThreads;   // 16-byte chunks per thread
    __shared__ __attribute__((aligned(16))) uint8_t As[2][BM * kLdsRow];
    __shared__ __attribute__((aligned(16))) uint8_t Bs[2][BN * kLdsRow];

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
    const int64_t m0 = static_cast<int64_t>(tm) * BM;
    const int64_t n0 = static_cast<int64_t>(tn) * BN;

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave & 1, wn = wave >> 1;
    const int lrow = lane & 31, lgrp = lane >> 5;

    // ---- global -> register staging: QA / QB x 16 B per thread ----
    // chunk c of a stage (128 bytes of K for every row): row layout, 8 lanes
    // per row's 128 contiguous bytes; packed, the stage is two 64-byte blocks
    // of 16 KB-strided stage blocks, so 4 lanes per row-half and consecutive
    // lanes on consecutive rows keep every wave load on one contiguous range
    static_assert(BM == BN, "one chunk map serves both operand tiles");
    auto chunk_row = [&](int c) { return PACKED ? (c & (BM * 4 - 1)) >> 2 : c >> 3; };
    auto chunk_part = [&](int c) { return PACKED ? ((c / (BM * 4)) << 2) | (c & 3) : c & 7; };
    intx4_t ra[QA], rb[QB];
    const int64_t kbytes_total = static_cast<int64_t>(nslices) * kslice * F::es;
    const int nstages = static_cast<int>(kbytes_total / kRowBytes);

    auto gload = [&](int stage) {
        const int64_t kb = static_cast<int64_t>(stage) * kRowBytes;
        // rows past the edge re-read the last valid row: they feed only
        // accumulator rows / columns the epilogue never stores, so no select
        // is needed and the loads' wait stays at the LDS store after the MFMAs
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int c = tid + kThreads * q;
            const int row = chunk_row(c), part = chunk_part(c);
            const int64_t gm = m0 + row;
            const int64_t rm = gm < M ? gm : M - 1;
            ra[q] = *reinterpret_cast<const intx4_t *>(
                Lq + (PACKED ? packed_off(rm, kb + part * 16, ldl_b) : rm * ldl_b + kb + part * 16));
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int c = tid + kThreads * q;
            const int row = chunk_row(c), part = chunk_part(c);
            const int64_t gn = n0 + row;
            const int64_t rn = gn < N ? gn : N - 1;
            rb[q] = *reinterpret_cast<const intx4_t *>(
                Uq + (PACKED ? packed_off(rn, kb + part * 16, ldu_b) : rn * ldu_b + kb + part * 16));
        }
    };
    auto sstore = [&](int buf) {
#pragma unroll
        for (int q = 0; q < QA; ++q) {
            const int c = tid + kThreads * q;
            *reinterpret_cast<intx4_t *>(&As[buf][chunk_row(c) * kLdsRow + chunk_part(c) * 16]) = ra[q];
        }
#pragma unroll
        for (int q = 0; q < QB; ++q) {
            const int c = tid + kThreads * q;
            *reinterpret_cast<intx4_t *>(&Bs[buf][chunk_row(c) * kLdsRow + chunk_part(c) * 16]) = rb[q];
        }
    };

    floatx16_t acc[IB][JB];
#pragma unroll
    for (int i = 0; i < IB; ++i)
#pragma unroll
        for (int j = 0; j < JB; ++j)
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
        // the last stage re-loads itself and stores into the idle buffer, so
        // the loop body is one basic block the scheduler can interleave
        gload(ks + 1 < nstages ? ks + 1 : ks);
        if (ks % stages_per_slice == 0 && eL != nullptr) {
            const int s = ks / stages_per_slice;
            const int el = eL[s * ldel + row_tile], eu = eU[s * ldeu + col_tile];
            if (FMT == HPLAI_FMT_E4M3) {
                scale_a = e8m0_scale(el);
                scale_b = e8m0_scale(eu);
            } else if (FMT == HPLAI_FMT_F16_TILE) {
                const int e_new = el + eu;
                if (ks == 0) {
                    e_cur = e_new;
                } else if (e_new != e_cur) {
                    const int d = e_new - e_cur;
#pragma unroll
                    for (int i = 0; i < IB; ++i)
#pragma unroll
                        for (int j = 0; j < JB; ++j)
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
                intx8_t a[IB], b[JB];
#pragma unroll
                for (int i = 0; i < IB; ++i) {
                    const uint8_t *p = Ab + (wm * (BM / 2) + i * 32 + lrow) * kLdsRow + kk * 64 + lgrp * 32;
                    intx4_t lo = *reinterpret_cast<const intx4_t *>(p);
                    intx4_t hi = *reinterpret_cast<const intx4_t *>(p + 16);
                    a[i] = intx8_t{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                }
#pragma unroll
                for (int j = 0; j < JB; ++j) {
                    const uint8_t *p = Bb + (wn * (BN / 2) + j * 32 + lrow) * kLdsRow + kk * 64 + lgrp * 32;
                    intx4_t lo = *reinterpret_cast<const intx4_t *>(p);
                    intx4_t hi = *reinterpret_cast<const intx4_t *>(p + 16);
                    b[j] = intx8_t{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
                }
#pragma unroll
                for (int i = 0; i < IB; ++i)
#pragma unroll
                    for (int j = 0; j < JB; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                            a[i], b[j], acc[i][j], 0, 0, 0, scale_a, 0, scale_b);
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                half8_t a[IB], b[JB];
#pragma unroll
                for (int i = 0; i < IB; ++i)
                    a[i] = *reinterpret_cast<const half8_t *>(
                        Ab + (wm * (BM / 2) + i * 32 + lrow) * kLdsRow + kk * 32 + lgrp * 16);
#pragma unroll
                for (int j = 0; j < JB; ++j)
                    b[j] = *reinterpret_cast<const half8_t *>(
                        Bb + (wn * (BN / 2) + j * 32 + lrow) * kLdsRow + kk * 32 + lgrp * 16);
#pragma unroll
                for (int i = 0; i < IB; ++i)
#pragma unroll
                    for (int j = 0; j < JB; ++j)
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[i], b[j], acc[i][j], 0, 0, 0);
            }
        }
        // keep the prefetch's LDS stores (and their wait on the global loads)
        // after this stage's MFMAs; buf ^ 1 was last read before the previous
        // barrier, so storing into it is safe on every stage
        __builtin_amdgcn_sched_barrier(0);
        sstore(buf ^ 1);
        __syncthreads();
    }

    // ---- epilogue: C -= acc (FP16 tile mode: acc * 2^-e_cur) ----
    // acc[i][j][4*q + r] holds row 8q + 4*lgrp + r, column lrow of block (i, j)
    if (FMT == HPLAI_FMT_F16_TILE) {
#pragma unroll
        for (int i = 0; i < IB; ++i)
#pragma unroll
            for (int j = 0; j < JB; ++j)
#pragma unroll
                for (int r = 0; r < 16; ++r) acc[i][j][r] = ldexp_fast(acc[i][j][r], -e_cur);
    }
    if constexpr (EPI >= 2) {
        static_assert(4 * (BN / 2) * kEpiTS * 4 <= 2 * BM * kLdsRow, "epilogue scratch fits the A stages");
        __syncthreads();   // every wave is done with the operand stages
        float *T = reinterpret_cast<float *>(&As[0][0]) + wave * (BN / 2) * kEpiTS;
#pragma unroll
        for (int i = 0; i < IB; ++i) {
            if constexpr (EPI == 3)
                epilogue_block_row_lds_h<JB>(static_cast<uint16_t *>(Cv), ldc, m0 + wm * (BM / 2) + i * 32,
                                             n0 + wn * (BN / 2), lane, T, acc[i], cscale);
            else
                epilogue_block_row_lds<JB>(C, ldc, m0 + wm * (BM / 2) + i * 32, n0 + wn * (BN / 2), lane, T,
                                           acc[i]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < IB; ++i)
            epilogue_block_row<JB, EPI == 1>(C, ldc, M, N, m0 + wm * (BM / 2) + i * 32, n0 + wn * (BN / 2), lrow,
                                             lgrp, acc[i]);
    }
}

// ---------------------------------------------------------------------------
// Large-update kernel: 256 x 256 tile, operands streamed into LDS by the
// direct global->LDS loads (global_load_lds_dwordx4), four LDS stages of 64 K
// bytes per row, three of them in flight while the fourth feeds the MFMAs.
// One stage is 32 KB (256 rows of Lq + 256 rows of Uq); a wave-wide load
// fills 16 rows (1 KB) with lane l writing bytes [16 l, 16 l + 16).  The
// 16-byte chunks of a row are XOR-swizzled (chunk c of row r sits at
// position c ^ ((r >> 2) & 3)) so the 32 rows an MFMA operand read touches
// fall in distinct banks; the swizzle is applied by choosing which global
// chunk each lane fetches.  The waits are explicit: the compiler neither
// orders nor counts the LDS side of these loads.
// ---------------------------------------------------------------------------
constexpr int kDmaTile = 256;
constexpr int kDmaStageBytes = 64;                                  // K bytes per row per stage
constexpr int kDmaStages = 4;
constexpr int kDmaStage = 2 * kDmaTile * kDmaStageBytes;           // 32 KB
constexpr int kDmaLoadsPerStage = 2 * kDmaTile * kDmaStageBytes / (kThreads * 16);   // 8 per thread
// e4m3 exponents of one launch staged in LDS (slices per launch; longer
// updates take the 128 x 128 kernel)
constexpr int kDmaMaxSlices = 2048;

// EPI: 0 = element-wise C access (unaligned C), 1 = float4 per accumulator
// row block with ragged edges, 2 = whole tiles only (M, N multiples of 256),
// C transposed through LDS for coalesced access.
template <int FMT, int EPI, bool PACKED>
__global__ __launch_bounds__(kThreads) void schur_dma_kernel(
    int64_t M, int64_t N, int nslices, int kslice, const uint8_t *__restrict__ Lq, int64_t ldl_b,
    const int32_t *__restrict__ eL, int64_t ldel, const uint8_t *__restrict__ Uq, int64_t ldu_b,
    const int32_t *__restrict__ eU, int64_t ldeu, int tb, void *__restrict__ Cv, int64_t ldc,
    int tiles_m, int tiles_n, float cscale)
{
    float *__restrict__ C = static_cast<float *>(Cv);
    static_assert(FMT != HPLAI_FMT_F16_TILE, "per-tile FP16 rescales accumulators; not supported here");
    static_assert(kDmaLoadsPerStage == 8, "four Lq and four Uq loads per thread per stage");
    using F = SchurFmt<FMT>;
    constexpr int BM = kDmaTile, BN = kDmaTile, SB = kDmaStageBytes;
    // the operand stages, then (e4m3) the exponents of the launch's slices;
    // one array, so that the compiler sees a single LDS object written by the
    // direct loads
    constexpr int kExpBytes = FMT == HPLAI_FMT_E4M3 ? 2 * kDmaMaxSlices * 4 : 0;
    __shared__ __attribute__((aligned(1024))) uint8_t S[kDmaStages * kDmaStage + kExpBytes];
    int32_t *const SE = reinterpret_cast<int32_t *>(S + kDmaStages * kDmaStage);

    const int nwg = tiles_m * tiles_n;
    const int logical = xcd_remap(static_cast<int>(blockIdx.x), nwg);
    constexpr int G = 8;
    const int per_group = G * tiles_n;
    const int first_m = (logical / per_group) * G;
    const int gsize = tiles_m - first_m < G ? tiles_m - first_m : G;
    const int in_group = logical % per_group;
    const int tm = first_m + in_group % gsize;
    const int tn = in_group / gsize;
    const int64_t m0 = static_cast<int64_t>(tm) * BM;
    const int64_t n0 = static_cast<int64_t>(tn) * BN;

    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = wave & 1, wn = wave >> 1;
    const int lrow = lane & 31, lgrp = lane >> 5;

    // ---- per-thread sources: load q of this wave fills rows (4 q + wave) * 16 + lane / 4 ----
    // rows past the edge re-read the last valid row; they only feed
    // accumulator rows / columns the epilogue never stores
    const uint8_t *srcA[4], *srcB[4];
    {
        const int chunk = (lane & 3) ^ ((lane >> 4) & 3);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = (4 * q + wave) * 16 + (lane >> 2);
            const int64_t gm = m0 + row < M ? m0 + row : M - 1;
            const int64_t gn = n0 + row < N ? n0 + row : N - 1;
            srcA[q] = Lq + (PACKED ? packed_off(gm, chunk * 16, ldl_b) : gm * ldl_b + chunk * 16);
            srcB[q] = Uq + (PACKED ? packed_off(gn, chunk * 16, ldu_b) : gn * ldu_b + chunk * 16);
        }
    }
    const int nstages = static_cast<int>(static_cast<int64_t>(nslices) * kslice * F::es / SB);
    const int stages_per_slice = kslice * F::es / SB;

    // stage `st` of K into LDS buffer `buf`
    auto issue = [&](int st, int buf) {
        // next K stage: SB bytes on in the row layout, one 16 KB block on when packed
        const int64_t kb = static_cast<int64_t>(st) * (PACKED ? 256 * SB : SB);
        uint8_t *dst = S + buf * kDmaStage + wave * 1024;
#pragma unroll
        for (int q = 0; q < 4; ++q)
            __builtin_amdgcn_global_load_lds(srcA[q] + kb, dst + q * 4096, 16, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            __builtin_amdgcn_global_load_lds(srcB[q] + kb, dst + BM * SB + q * 4096, 16, 0, 0);
    };

    floatx16_t acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.0f;

    const int row_tile = tb > 0 ? static_cast<int>(m0 / tb) : 0;
    const int col_tile = tb > 0 ? static_cast<int>(n0 / tb) : 0;
    // the tile's exponents of every slice go to LDS before the operand
    // stream starts: read in the loop they are ordered with the operand
    // reads, where a scalar load would make every wait on those reads also
    // wait for it (and a scalar-cache miss behind the operand stream costs
    // more than a stage).  (Staged after the prologue's operand loads
    // instead, the LDS writes wait for all four stages, not just the first:
    // measured slower.)
    if constexpr (FMT == HPLAI_FMT_E4M3) {
        for (int t = tid; t < nslices; t += kThreads) {
            SE[t] = eL[t * ldel + row_tile];
            SE[kDmaMaxSlices + t] = eU[t * ldeu + col_tile];
        }
        __syncthreads();
    }

    // operand registers of one stage; two sets alternate so the reads of
    // stage ks + 1 overlap the second half of stage ks's MFMAs
    struct Ops {
        typename std::conditional<FMT == HPLAI_FMT_E4M3, intx8_t, half8_t>::type a[2][4], b[2][4];
        int sa, sb;   // E4M3 block scales of the stage
    };
    constexpr int KK = FMT == HPLAI_FMT_E4M3 ? 1 : 2;   // MFMA K steps per stage

    // operand read offsets inside a stage (swizzled chunk positions)
    const int f = (lrow >> 2) & 3;
    const int a_base = (wm * (BM / 2) + lrow) * SB;
    const int b_base = BM * SB + (wn * (BN / 2) + lrow) * SB;

    // slice bookkeeping of the stage being read next (E4M3 scales change per slice)
    int rd_slice = 0, rd_in_slice = 0;
    auto read = [&](int buf, Ops &o) {
        const uint8_t *St = S + buf * kDmaStage;
        if constexpr (FMT == HPLAI_FMT_E4M3) {
            const int off_lo = 16 * ((2 * lgrp) ^ f), off_hi = 16 * ((2 * lgrp + 1) ^ f);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const uint8_t *p = St + a_base + i * 32 * SB;
                const intx4_t lo = *reinterpret_cast<const intx4_t *>(p + off_lo);
                const intx4_t hi = *reinterpret_cast<const intx4_t *>(p + off_hi);
                o.a[0][i] = intx8_t{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const uint8_t *p = St + b_base + j * 32 * SB;
                const intx4_t lo = *reinterpret_cast<const intx4_t *>(p + off_lo);
                const intx4_t hi = *reinterpret_cast<const intx4_t *>(p + off_hi);
                o.b[0][j] = intx8_t{lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < 2; ++kk) {
                const int off = 16 * ((2 * kk + lgrp) ^ f);
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    o.a[kk][i] = *reinterpret_cast<const half8_t *>(St + a_base + i * 32 * SB + off);
#pragma unroll
                for (int j = 0; j < 4; ++j)
                    o.b[kk][j] = *reinterpret_cast<const half8_t *>(St + b_base + j * 32 * SB + off);
            }
        }
    };
    // E4M3 scales: the exponents of the stage after next are read from LDS one
    // step early and consumed right after the workgroup barrier
    int pe_l = 0, pe_u = 0;
    auto fetch_exps = [&]() {
        if constexpr (FMT == HPLAI_FMT_E4M3) {
            // fetches past the last stage (never used) keep to the last slice
            const int sl = rd_slice < nslices ? rd_slice : nslices - 1;
            pe_l = SE[sl];
            pe_u = SE[kDmaMaxSlices + sl];
        }
        if (++rd_in_slice == stages_per_slice) {
            rd_in_slice = 0;
            ++rd_slice;
        }
    };
    auto take_exps = [&](Ops &o) {
        o.sa = e8m0_scale(pe_l);
        o.sb = e8m0_scale(pe_u);
    };
    // MFMAs of accumulator rows i0, i0 + 1
    auto mma = [&](const Ops &o, int i0) {
#pragma unroll
        for (int kk = 0; kk < KK; ++kk)
#pragma unroll
            for (int i = i0; i < i0 + 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    if constexpr (FMT == HPLAI_FMT_E4M3)
                        acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
                            o.a[kk][i], o.b[kk][j], acc[i][j], 0, 0, 0, o.sa, 0, o.sb);
                    else
                        acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.a[kk][i], o.b[kk][j],
                                                                           acc[i][j], 0, 0, 0);
                }
    };
    // One stage: first half of the MFMAs, then (once stage ks + 1 has landed
    // and every wave has its stage-ks operands in registers) refill the
    // stage-ks buffer with stage ks + 4 and read stage ks + 1 while the second
    // half runs.  At that wait the loads of stages ks + 2 and ks + 3 (8 each
    // per thread) may still be in flight.
    // (The last stage also reads the clamped copy in the next buffer; the
    // result is never used.)  The scheduling barriers keep the MFMAs, which
    // touch no memory, on their side of the workgroup barrier.
    auto step = [&](int ks, const Ops &cur, Ops &nxt) {
        mma(cur, 0);
        __builtin_amdgcn_sched_barrier(0);
        asm volatile("s_waitcnt vmcnt(16)\n\ts_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
        take_exps(nxt);
        fetch_exps();
        __builtin_amdgcn_sched_barrier(0);
        read((ks + 1) & (kDmaStages - 1), nxt);
        const int ld = ks + kDmaStages;
        issue(ld < nstages ? ld : nstages - 1, ld & (kDmaStages - 1));
        mma(cur, 2);
        // operand reads first, spread between the MFMAs, then the refill loads
        constexpr int NM = 8 * KK;   // MFMAs in this half
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);   // 4 LDS reads
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // 1 MFMA
        }
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            __builtin_amdgcn_sched_group_barrier(0x010, 2, 0);   // 2 refill loads
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);   // 1 MFMA
        }
        __builtin_amdgcn_sched_group_barrier(0x008, NM - 8, 0);
        __builtin_amdgcn_sched_barrier(0);
    };

    // prologue: stages 0..3 (clamped; extra copies land in buffers nobody reads)
#pragma unroll
    for (int st = 0; st < kDmaStages; ++st) issue(st < nstages ? st : nstages - 1, st);
    Ops r0, r1;
    fetch_exps();               // stage 0
    take_exps(r0);
    fetch_exps();               // stage 1
    asm volatile("s_waitcnt vmcnt(24)\n\ts_barrier" ::: "memory");
    read(0, r0);
    // nstages is even (K per launch is a multiple of 128 bytes)
    for (int ks = 0; ks < nstages; ks += 2) {
        step(ks, r0, r1);
        step(ks + 1, r1, r0);
    }
    // drain the clamped tail loads before the workgroup can exit
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");

    // ---- epilogue: C -= acc ----
    if constexpr (EPI >= 2) {
        __syncthreads();   // every wave is done with the operand stages
        float *T = reinterpret_cast<float *>(S) + wave * (BN / 2) * kEpiTS;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if constexpr (EPI == 3)
                epilogue_block_row_lds_h<4>(static_cast<uint16_t *>(Cv), ldc, m0 + wm * (BM / 2) + i * 32,
                                            n0 + wn * (BN / 2), lane, T, acc[i], cscale);
            else
                epilogue_block_row_lds<4>(C, ldc, m0 + wm * (BM / 2) + i * 32, n0 + wn * (BN / 2), lane, T,
                                          acc[i]);
        }
    } else {
        // one accumulator row block (16 float4 per lane) at a time with all its
        // loads in flight together; ragged edges element by element
#pragma unroll
        for (int i = 0; i < 4; ++i)
            epilogue_block_row<4, EPI == 1>(C, ldc, M, N, m0 + wm * (BM / 2) + i * 32, n0 + wn * (BN / 2), lrow,
                                            lgrp, acc[i]);
    }
}

// c16: C is the FP16 trailing storage (whole tiles, 8-byte aligned), C := fp16(C - cscale * L U)
template <int FMT, bool PACKED>
static int schur_dma_launch(int64_t m, int64_t n, int nslices, int kslice, const void *Lq, int64_t ldl,
                            const int32_t *eL, int64_t ldel, const void *Uq, int64_t ldu,
                            const int32_t *eU, int64_t ldeu, int tb, void *C, int64_t ldc, hipStream_t s,
                            bool c16 = false, float cscale = 1.0f)
{
    using F = SchurFmt<FMT>;
    const int64_t tiles_m = ceil_div(m, kDmaTile), tiles_n = ceil_div(n, kDmaTile);
    if (tiles_m * tiles_n >= (int64_t(1) << 31)) {
        set_error("schur_update: %lld x %lld tiles exceed one launch", (long long)tiles_m,
                  (long long)tiles_n);
        return HPLAI_EINVAL;
    }
    const bool vec = (reinterpret_cast<uintptr_t>(C) % 16 == 0) && (ldc % 4 == 0);
    const dim3 grid(static_cast<unsigned>(tiles_m * tiles_n));
    const uint8_t *L8 = static_cast<const uint8_t *>(Lq);
    const uint8_t *U8 = static_cast<const uint8_t *>(Uq);
    const bool whole = vec && m % kDmaTile == 0 && n % kDmaTile == 0;
    if (c16)
        hipLaunchKernelGGL((schur_dma_kernel<FMT, 3, PACKED>), grid, dim3(kThreads), 0, s, m, n, nslices, kslice,
                           L8, ldl * F::es, eL, ldel, U8, ldu * F::es, eU, ldeu, tb, C, ldc,
                           static_cast<int>(tiles_m), static_cast<int>(tiles_n), cscale);
    else if (whole)
        hipLaunchKernelGGL((schur_dma_kernel<FMT, 2, PACKED>), grid, dim3(kThreads), 0, s, m, n, nslices, kslice,
                           L8, ldl * F::es, eL, ldel, U8, ldu * F::es, eU, ldeu, tb, C, ldc,
                           static_cast<int>(tiles_m), static_cast<int>(tiles_n), 1.0f);
    else if (vec)
        hipLaunchKernelGGL((schur_dma_kernel<FMT, 1, PACKED>), grid, dim3(kThreads), 0, s, m, n, nslices, kslice,
                           L8, ldl * F::es, eL, ldel, U8, ldu * F::es, eU, ldeu, tb, C, ldc,
                           static_cast<int>(tiles_m), static_cast<int>(tiles_n), 1.0f);
    else
        hipLaunchKernelGGL((schur_dma_kernel<FMT, 0, PACKED>), grid, dim3(kThreads), 0, s, m, n, nslices, kslice,
                           L8, ldl * F::es, eL, ldel, U8, ldu * F::es, eU, ldeu, tb, C, ldc,
                           static_cast<int>(tiles_m), static_cast<int>(tiles_n), 1.0f);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

template <int FMT, int BM, int BN, bool PACKED>
static int schur_launch(int64_t m, int64_t n, int nslices, int kslice, const void *Lq, int64_t ldl,
                        const int32_t *eL, int64_t ldel, const void *Uq, int64_t ldu,
                        const int32_t *eU, int64_t ldeu, int tb, void *C, int64_t ldc, hipStream_t s,
                        bool c16 = false, float cscale = 1.0f)
{
    using F = SchurFmt<FMT>;
    const int64_t tiles_m = ceil_div(m, BM), tiles_n = ceil_div(n, BN);
    if (tiles_m * tiles_n >= (int64_t(1) << 31)) {
        set_error("schur_update: %lld x %lld tiles exceed one launch", (long long)tiles_m,
                  (long long)tiles_n);
        return HPLAI_EINVAL;
    }
    const bool vec = (reinterpret_cast<uintptr_t>(C) % 16 == 0) && (ldc % 4 == 0);
    const dim3 grid(static_cast<unsigned>(tiles_m * tiles_n));
    const uint8_t *L8 = static_cast<const uint8_t *>(Lq);
    const uint8_t *U8 = static_cast<const uint8_t *>(Uq);
    const bool whole = vec && m % BM == 0 && n % BN == 0;
    if (c16)
        hipLaunchKernelGGL((schur_kernel<FMT, BM, BN, 3, PACKED>), grid, dim3(kThreads), 0, s, m, n, nslices,
                           kslice, L8, ldl * F::es, eL, ldel, U8, ldu * F::es, eU, ldeu, tb, C, ldc,
                           static_cast<int>(tiles_m), static_cast<int>(tiles_n), cscale);
    else if (whole)
        hipLaunchKernelGGL((schur_kernel<FMT, BM, BN, 2, PACKED>), grid, dim3(kThreads), 0, s, m, n, nslices,
                           kslice, L8, ldl * F::es, eL, ldel, U8, ldu * F::es, eU, ldeu, tb, C, ldc,
                           static_cast<int>(tiles_m), static_cast<int>(tiles_n), 1.0f);
    else if (vec)
        hipLaunchKernelGGL((schur_kernel<FMT, BM, BN, 1, PACKED>), grid, dim3(kThreads), 0, s, m, n, nslices,
                           kslice, L8, ldl * F::es, eL, ldel, U8, ldu * F::es, eU, ldeu, tb, C, ldc,
                           static_cast<int>(tiles_m), static_cast<int>(tiles_n), 1.0f);
    else
        hipLaunchKernelGGL((schur_kernel<FMT, BM, BN, 0, PACKED>), grid, dim3(kThreads), 0, s, m, n, nslices,
                           kslice, L8, ldl * F::es, eL, ldel, U8, ldu * F::es, eU, ldeu, tb, C, ldc,
                           static_cast<int>(tiles_m), static_cast<int>(tiles_n), 1.0f);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

// The streamed 256 x 256 kernel for updates with at least two tiles per CU;
// the 128 x 128 kernel for thin ones.  With per-tile exponents every
// workgroup tile must lie inside one tb x tb tile, so the large tiles also
// need tb % 256 == 0.  The per-tile FP16 form rescales its accumulators
// between slices and keeps the 128 x 128 tile.
bool schur_large_tiles(int fmt, int64_t m, int64_t n, int tb, int num_cu)
{
    if (fmt == HPLAI_FMT_F16_TILE) return false;
    if (fmt == HPLAI_FMT_E4M3 && tb % kDmaTile != 0) return false;
    if (num_cu <= 0) num_cu = 256;
    return ceil_div(m, kDmaTile) * ceil_div(n, kDmaTile) >= 2 * num_cu;
}

template <int FMT, bool PACKED>
static int schur_dispatch(int64_t m, int64_t n, int nslices, int kslice, const void *Lq, int64_t ldl,
                          const int32_t *eL, int64_t ldel, const void *Uq, int64_t ldu,
                          const int32_t *eU, int64_t ldeu, int tb, void *C, int64_t ldc, int num_cu,
                          hipStream_t s, bool c16 = false, float cscale = 1.0f)
{
    if constexpr (FMT != HPLAI_FMT_F16_TILE) {
        // the FP16 trailing storage takes the streamed kernel only for whole 256 x 256 tiles
        if (schur_large_tiles(FMT, m, n, tb, num_cu) && (!c16 || (m % kDmaTile == 0 && n % kDmaTile == 0)) &&
            (FMT != HPLAI_FMT_E4M3 || nslices <= kDmaMaxSlices))
            return schur_dma_launch<FMT, PACKED>(m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu, eU, ldeu,
                                         tb, C, ldc, s, c16, cscale);
    }
    return schur_launch<FMT, 128, 128, PACKED>(m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu, eU, ldeu, tb,
                                       C, ldc, s, c16, cscale);
}

// Validation lives here so the C ABI and the solver share it.
static int schur_update(int fmt, int64_t m, int64_t n, int nslices, int kslice, const void *Lq,
                        int64_t ldl, const int32_t *eL, int64_t ldel, const void *Uq, int64_t ldu,
                        const int32_t *eU, int64_t ldeu, int tb, void *C, int64_t ldc, int num_cu,
                        hipStream_t s, bool packed, bool c16, float cscale)
{
    if (m <= 0 || n <= 0 || nslices <= 0) return HPLAI_OK;
    if (num_cu <= 0) num_cu = 256;
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
        HPLAI_CHECK(tb > 0 && tb % 128 == 0, "schur_update: tile %d must be a multiple of 128", tb);
    }
    if (packed)
        HPLAI_CHECK((ldl * es) % 64 == 0 && (ldu * es) % 64 == 0,
                    "schur_update: packed operands need rows of whole 64-byte stages");
    if (c16)
        HPLAI_CHECK(m % 128 == 0 && n % 128 == 0 && ldc % 4 == 0 && reinterpret_cast<uintptr_t>(C) % 8 == 0 &&
                        fmt != HPLAI_FMT_F16_TILE,
                    "schur_update: FP16 trailing storage needs whole 128 x 128 tiles, ldc %% 4 == 0, an "
                    "8-byte aligned C and the FP16 or e4m3 format");
    switch (fmt) {
    case HPLAI_FMT_F16:
        return packed ? schur_dispatch<HPLAI_FMT_F16, true>(m, n, nslices, kslice, Lq, ldl, nullptr, 0, Uq, ldu,
                                                            nullptr, 0, 0, C, ldc, num_cu, s, c16, cscale)
                      : schur_dispatch<HPLAI_FMT_F16, false>(m, n, nslices, kslice, Lq, ldl, nullptr, 0, Uq, ldu,
                                                             nullptr, 0, 0, C, ldc, num_cu, s, c16, cscale);
    case HPLAI_FMT_F16_TILE:
        return packed ? schur_dispatch<HPLAI_FMT_F16_TILE, true>(m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq,
                                                                 ldu, eU, ldeu, tb, C, ldc, num_cu, s, c16, cscale)
                      : schur_dispatch<HPLAI_FMT_F16_TILE, false>(m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq,
                                                                  ldu, eU, ldeu, tb, C, ldc, num_cu, s, c16, cscale);
    default:
        return packed ? schur_dispatch<HPLAI_FMT_E4M3, true>(m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu,
                                                             eU, ldeu, tb, C, ldc, num_cu, s, c16, cscale)
                      : schur_dispatch<HPLAI_FMT_E4M3, false>(m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu,
                                                              eU, ldeu, tb, C, ldc, num_cu, s, c16, cscale);
    }
}


int launch_schur(int fmt, int64_t m, int64_t n, int nslices, int kslice, const void *Lq,
                 int64_t ldl, const int32_t *eL, int64_t ldel, const void *Uq, int64_t ldu,
                 const int32_t *eU, int64_t ldeu, int tb, float *C, int64_t ldc, int num_cu,
                 hipStream_t s, bool packed)
{
    return schur_update(fmt, m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu, eU, ldeu, tb, C, ldc, num_cu, s,
                        packed, false, 1.0f);
}

int launch_schur_c16(int fmt, int64_t m, int64_t n, int nslices, int kslice, const void *Lq,
                     int64_t ldl, const int32_t *eL, int64_t ldel, const void *Uq, int64_t ldu,
                     const int32_t *eU, int64_t ldeu, int tb, uint16_t *C, int64_t ldc, float cscale,
                     int num_cu, hipStream_t s, bool packed)
{
    return schur_update(fmt, m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu, eU, ldeu, tb, C, ldc, num_cu, s,
                        packed, true, cscale);
}

}  // namespace hplai
