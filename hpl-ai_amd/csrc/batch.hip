// batch.hip — many independent HPL-AI systems of uneven sizes (n <= 1024,
// 1-4 right-hand sides) solved by one persistent launch.
//
// No reference counterpart: the reference solves one system per run.  Each
// system applies the reference's arithmetic to its own matrix: the LCG matrix
// of hpl_rand.hpp (element (i,j) = state 1+i+n*j from the system's seed,
// diagonal = FP64 row sum of |a_ij|, rhs r = column n+r of the stream), an
// unpivoted right-looking LU with FP32 panels and an FP16 Schur update with
// FP32 accumulation (the reference's plain method), and FP64 iterative
// refinement to HPL-AI harness < 16 (iterative_refinement.hpp:18-57) against
// the FP64 matrix, for every right-hand side.
//
// Scheduling: the grid is a fixed number of workgroups (one per CU); each
// workgroup takes the next system from a device counter (one vector atomic by
// one thread) until the list is exhausted, largest systems first.  One
// workgroup owns one system: generation, LU and refinement run without any
// grid-wide synchronisation, and every workgroup leaves the loop once the
// counter passes the end.
//
// LU of one system, panel width NB (16, 32 or 64; chosen per size bucket by
// the autotuner): (a) the n-k x NB column panel is factored in LDS (FP32,
// one barrier per column); (b) U12 = L11^-1 A12 in FP32; (c) L21 and U12 are
// rounded to FP16 into LDS (rows K-contiguous, 16-byte chunks XOR-swizzled)
// and the trailing matrix takes C -= L21 U12 with v_mfma_f32_32x32x16_f16
// (FP32 accumulation), each wave owning 32 x 32 tiles.
#include <algorithm>
#include <cfloat>
#include <chrono>
#include <vector>

#include "device_common.hpp"
#include "internal.hpp"
#include "trsv_block.hpp"

namespace hplai {

typedef _Float16 half8b_t __attribute__((ext_vector_type(8)));
typedef float floatx16b_t __attribute__((ext_vector_type(16)));

constexpr int kBT = 256;                  // threads per workgroup
constexpr int kBatchMaxN = 1024;
constexpr int kBatchLds = 132 * 1024;     // panel / FP16 operands + L11 copy
constexpr int kBuckets = 4;               // [32,128) [128,256) [256,512) [512,1024]

struct BatchSys {
    int32_t n, nrhs, nb, lda;             // lda = n rounded up to 4 (16-byte row groups)
    uint64_t x1;                          // lcg_init(seed)
    int64_t a32, a64, d, b, x, r;         // offsets into the float / double pools
};

struct BatchOut {
    double harness[4];
    int32_t steps[4];
    int32_t info;                         // 0, or k+1 for the first zero / non-finite pivot
    int32_t pad;
    uint64_t t0, t1;                      // wall clock at start / end of the system
};

__device__ __forceinline__ double block_max(double v, double *red)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off));
    if (lane == 0) red[wave] = v;
    __syncthreads();
    const double r = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
    __syncthreads();
    return r;
}

// ---- generation: A64 (FP64 values), A32 = (float)A64 with (float)d on the
// diagonal, d, and the right-hand sides (stream columns n .. n+nrhs-1) ----
__device__ __noinline__ void batch_generate(const BatchSys &S, float *A32, double *A64, double *d, double *B)
{
    const int n = S.n, lda = S.lda, tid = threadIdx.x;
    const int64_t total = static_cast<int64_t>(n) * (n + S.nrhs);
    // element e (column-major, continuing into the rhs columns) is state 1 + e;
    // thread t takes e = t + 256 q, advancing 256 states per step
    const LcgCoeff step = kLcg.p[8];      // (a, c)^256
    uint64_t s = lcg_jump(S.x1, static_cast<uint64_t>(tid));
    // (i, j) of element e, advanced with e (one division at the start)
    int64_t j = tid / n, i = tid - j * n;
    for (int64_t e = tid; e < total; e += kBT) {
        const double v = lcg_value(s);
        if (j < n) {
            A64[i + j * lda] = v;
            A32[i + j * lda] = static_cast<float>(v);
        } else {
            B[(j - n) * n + i] = v;
        }
        s = step.a * s + step.c;
        i += kBT;
        while (i >= n) {
            i -= n;
            ++j;
        }
    }
    __syncthreads();
    // calc_diag (hpl_rand.hpp:108-123): FP64, j ascending, j != i; the loads
    // of 16 columns are issued together ahead of their adds
    for (int i = tid; i < n; i += kBT) {
        const double *arow = A64 + i;
        double sum = 0.0;
        int j = 0;
        for (; j + 16 <= n; j += 16) {
            double a[16];
#pragma unroll
            for (int q = 0; q < 16; ++q) a[q] = arow[static_cast<int64_t>(j + q) * lda];
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if (j + q != i) sum += fabs(a[q]);
        }
        for (; j < n; ++j)
            if (j != i) sum += fabs(arow[static_cast<int64_t>(j) * lda]);
        d[i] = sum;
        A64[i + static_cast<int64_t>(i) * lda] = sum;
        A32[i + static_cast<int64_t>(i) * lda] = static_cast<float>(sum);
    }
    __syncthreads();
}

// swizzled position of the 16-byte chunk `c` of K-row `r` (rows of NB halves)
template <int NB>
__device__ __forceinline__ int chunk_pos(int r, int c)
{
    constexpr int CH = NB * 2 / 16;       // chunks per row: 2, 4, 8
    return c ^ ((r >> 2) & (CH - 1));
}

// LDS bytes the LU needs: panel (n x NB FP32), L21 and U12 in FP16 (rows
// padded to 32) and the L11 copy
__host__ __device__ constexpr int64_t batch_lu_lds(int n, int nb)
{
    return (static_cast<int64_t>(n) * nb * 4 + 15) / 16 * 16 + 2 * ((n + 31) / 32 * 32) * nb * 2 +
           static_cast<int64_t>(nb) * nb * 4;
}

// ---- unpivoted LU of the n x n FP32 matrix A (leading dimension lda, a
// multiple of 4), panel width NB ----
template <int NB>
__device__ __noinline__ int batch_lu(int n, int lda, float *A, uint8_t *lds)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int RB = NB * 2;            // bytes of one FP16 K-row
    float *P = reinterpret_cast<float *>(lds);                       // mp x NB, column-major
    float *L11 = reinterpret_cast<float *>(lds + kBatchLds - NB * NB * 4);
    int info = 0;
    for (int k = 0; k < n; k += NB) {
        const int w = n - k < NB ? n - k : NB;
        const int mp = n - k;
        // (a) panel into LDS, 4 columns at a time: the (up to 16) loads of a
        // thread are issued together ahead of their LDS stores
        constexpr int RPT = (kBatchMaxN + kBT - 1) / kBT;   // panel rows per thread (mp <= n)
        for (int c0 = 0; c0 < w; c0 += 4) {
            float v[4][RPT];
#pragma unroll
            for (int cc = 0; cc < 4; ++cc)
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int c = c0 + cc, r = tid + q * kBT;
                    v[cc][q] = (c < w && r < mp) ? A[(k + r) + static_cast<int64_t>(k + c) * lda] : 0.0f;
                }
#pragma unroll
            for (int cc = 0; cc < 4; ++cc)
#pragma unroll
                for (int q = 0; q < RPT; ++q) {
                    const int c = c0 + cc, r = tid + q * kBT;
                    if (c < w && r < mp) P[c * mp + r] = v[cc][q];
                }
        }
        __syncthreads();
        // (a') unblocked LU of the tall panel (getrf_nopiv_small order per column)
        for (int c = 0; c < w; ++c) {
            const float piv = P[c * mp + c];
            if (!(piv != 0.0f && isfinite(piv))) {
                if (info == 0) info = k + c + 1;
            }
            for (int r = c + 1 + tid; r < mp; r += kBT) {
                const float l = P[c * mp + r] / piv;
                P[c * mp + r] = l;
                for (int c2 = c + 1; c2 < w; ++c2) P[c2 * mp + r] -= l * P[c2 * mp + c];
            }
            __syncthreads();
        }
        // (a'') factors back to A; keep L11 for the row block
        for (int c = 0; c < w; ++c)
            for (int r = tid; r < mp; r += kBT) A[(k + r) + static_cast<int64_t>(k + c) * lda] = P[c * mp + r];
        for (int e = tid; e < w * w; e += kBT) L11[e] = P[(e / w) * mp + (e % w)];   // L11[c*w + r]
        const int m = n - k - w;
        if (m <= 0) break;
        __syncthreads();
        // (c, L side) L21 rows -> FP16 K-rows in LDS, behind the panel
        const int mt = (m + 31) / 32 * 32;                  // rows padded to the tile
        uint8_t *Lh = lds + (static_cast<int64_t>(mp) * NB * 4 + 15) / 16 * 16;
        uint8_t *Uh = Lh + static_cast<int64_t>(mt) * RB;
        for (int i = tid; i < mt; i += kBT) {               // row of L21 (0-based)
            const int r = w + i;                            // its panel row
#pragma unroll
            for (int c = 0; c < NB / 8; ++c) {
                half8b_t h;
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int col = c * 8 + e;
                    h[e] = (i < m && col < w) ? static_cast<_Float16>(P[col * mp + r]) : static_cast<_Float16>(0.0f);
                }
                *reinterpret_cast<half8b_t *>(Lh + i * RB + chunk_pos<NB>(i, c) * 16) = h;
            }
        }
        // (b) U12 = L11^-1 A12, one column per thread; FP32 to A, FP16 to LDS
        for (int j = tid; j < mt; j += kBT) {
            float u[NB];
            if (j < m) {
                const float *col = A + k + static_cast<int64_t>(k + w + j) * lda;
#pragma unroll
                for (int i = 0; i < NB; ++i) u[i] = i < w ? col[i] : 0.0f;
#pragma unroll
                for (int i = 1; i < NB; ++i)
#pragma unroll
                    for (int t = 0; t < NB; ++t)
                        if (t < i && i < w) u[i] -= L11[t * w + i] * u[t];
                float *colw = A + k + static_cast<int64_t>(k + w + j) * lda;
#pragma unroll
                for (int i = 0; i < NB; ++i)
                    if (i < w) colw[i] = u[i];
            } else {
#pragma unroll
                for (int i = 0; i < NB; ++i) u[i] = 0.0f;
            }
#pragma unroll
            for (int c = 0; c < NB / 8; ++c) {
                half8b_t h;
#pragma unroll
                for (int e = 0; e < 8; ++e) h[e] = static_cast<_Float16>(u[c * 8 + e]);
                *reinterpret_cast<half8b_t *>(Uh + j * RB + chunk_pos<NB>(j, c) * 16) = h;
            }
        }
        __syncthreads();
        // (c) trailing update: 32 x 32 tiles, four per wave at a time so that
        // sixteen 16-byte loads per lane are in flight; rows go in aligned
        // groups of 4 (lda % 4 == 0, so a group never leaves the allocation)
        const int T = mt / 32;
        const int lrow = lane & 31, lg = lane >> 5;
        for (int t0 = wave * 4; t0 < T * T; t0 += 16) {
            float4 cv[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = t0 + u;
                const int ti = t % T, tj = t / T;
                const int gc = tj * 32 + lrow;
                const float *ccol = A + (k + w) + static_cast<int64_t>(k + w + (gc < m ? gc : 0)) * lda;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int gr = ti * 32 + 8 * q + 4 * lg;
                    cv[u][q] = (t < T * T && gc < m && gr < m) ? *reinterpret_cast<const float4 *>(ccol + gr)
                                                               : float4{0.0f, 0.0f, 0.0f, 0.0f};
                }
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int t = t0 + u;
                if (t >= T * T) break;
                const int ti = t % T, tj = t / T;
                floatx16b_t acc;
#pragma unroll
                for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll
                for (int kk = 0; kk < NB / 16; ++kk) {
                    const int ra = ti * 32 + lrow, rb = tj * 32 + lrow;
                    const int ch = kk * 2 + lg;             // 16-byte chunk: halves [8 ch, 8 ch + 8)
                    const half8b_t a = *reinterpret_cast<const half8b_t *>(Lh + ra * RB + chunk_pos<NB>(ra, ch) * 16);
                    const half8b_t b = *reinterpret_cast<const half8b_t *>(Uh + rb * RB + chunk_pos<NB>(rb, ch) * 16);
                    acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, acc, 0, 0, 0);
                }
                // acc[4 q + r] is row 8 q + 4 lg + r, column lrow of the tile
                const int gc = tj * 32 + lrow;
                if (gc < m) {
                    float *ccol = A + (k + w) + static_cast<int64_t>(k + w + gc) * lda;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const int gr = ti * 32 + 8 * q + 4 * lg;
                        if (gr < m) {
                            float4 c4 = cv[u][q];
                            c4.x -= acc[4 * q + 0];
                            c4.y -= acc[4 * q + 1];
                            c4.z -= acc[4 * q + 2];
                            c4.w -= acc[4 * q + 3];
                            *reinterpret_cast<float4 *>(ccol + gr) = c4;
                        }
                    }
                }
            }
        }
        __syncthreads();
    }
    __syncthreads();
    return info;
}

// The triangular solve of trsv_block (trsv_block.hpp) for the nr <= 4
// right-hand sides of a system at once (columns x + rr bs): every strip of
// the triangle is loaded once for all of them, and each right-hand side
// takes exactly trsv_block's operations in its order.  xs: 4 x 32 doubles
// of LDS.
template <int RPT, bool UPPER, bool UNIT>
__device__ __noinline__ void batch_trsv_multi(int bs, int lda, const float *__restrict__ T,
                                              double *__restrict__ x, int nr, double *xs)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    double xr[4][RPT];
#pragma unroll
    for (int rr = 0; rr < 4; ++rr)
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + 256 * k;
            xr[rr][k] = (rr < nr && r < bs) ? x[static_cast<int64_t>(rr) * bs + r] : 0.0;
        }
    const int nsb = (bs + 31) / 32;
    constexpr bool PF = RPT <= 2;
    float cur[RPT][32], nxt[PF ? RPT : 1][32];
    trsv_load_strip<RPT, false>(T, lda, bs, (UPPER ? nsb - 1 : 0) * 32, tid, cur);
    for (int step = 0; step < nsb; ++step) {
        const int q = UPPER ? nsb - 1 - step : step;
        const int sb = q * 32;
        const int w = bs - sb < 32 ? bs - sb : 32;
        if constexpr (PF)
            if (step + 1 < nsb) trsv_load_strip<RPT, false>(T, lda, bs, (UPPER ? q - 1 : q + 1) * 32, tid, nxt);
        const int kq = sb / 256, owner = (sb % 256) / 64, half = ((sb % 256) / 32) & 1;
        if (wave == owner) {
            const int i = lane - 32 * half;
            const bool mine = i >= 0 && i < w;
            double v[4] = {0.0, 0.0, 0.0, 0.0}, rinv = 1.0;
            float rowv[32];
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (k == kq) {
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) v[rr] = xr[rr][k];
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
#pragma unroll
            for (int cc = 0; cc < 32; ++cc) {
                const int c = UPPER ? 31 - cc : cc;
                if (c < w) {
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) {
                        if (rr < nr) {
                            if (!UNIT && i == c) v[rr] *= rinv;
                            const double xc = readlane_f64(v[rr], 32 * half + c);
                            if (mine && (UPPER ? i < c : i > c))
                                v[rr] = fma(-static_cast<double>(rowv[c]), xc, v[rr]);
                        }
                    }
                }
            }
            if (mine)
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) xs[rr * 32 + i] = v[rr];
#pragma unroll
            for (int k = 0; k < RPT; ++k)
                if (k == kq && mine)
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr) xr[rr][k] = v[rr];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + 256 * k;
            const bool further = UPPER ? (r < sb) : (r >= sb + w && r < bs);
            if (further) {
#pragma unroll
                for (int rr = 0; rr < 4; ++rr) {
                    if (rr < nr) {
                        double acc = 0.0;
#pragma unroll
                        for (int c = 0; c < 32; ++c)
                            if (c < w) acc = fma(static_cast<double>(cur[k][c]), xs[rr * 32 + c], acc);
                        xr[rr][k] -= acc;
                    }
                }
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
                trsv_load_strip<RPT, false>(T, lda, bs, (UPPER ? q - 1 : q + 1) * 32, tid, cur);
            }
        }
    }
#pragma unroll
    for (int rr = 0; rr < 4; ++rr)
#pragma unroll
        for (int k = 0; k < RPT; ++k) {
            const int r = tid + 256 * k;
            if (rr < nr && r < bs) x[static_cast<int64_t>(rr) * bs + r] = xr[rr][k];
        }
}

template <bool UPPER, bool UNIT>
__device__ void batch_trsv_all(int n, int lda, const float *A, double *x, int nr, double *xs)
{
    if (n <= 256)
        batch_trsv_multi<1, UPPER, UNIT>(n, lda, A, x, nr, xs);
    else if (n <= 512)
        batch_trsv_multi<2, UPPER, UNIT>(n, lda, A, x, nr, xs);
    else
        batch_trsv_multi<4, UPPER, UNIT>(n, lda, A, x, nr, xs);
}

// ---- FP64 refinement of every right-hand side (iterative_refinement.hpp:18-57) ----
__device__ __noinline__ void batch_refine(const BatchSys &S, const float *LU, const double *A64, const double *d,
                             const double *B, double *X, double *R, double *xl, double *xs, double *red,
                             BatchOut &o, int maxit)
{
    // every right-hand side refines on its own (its own harness, steps and
    // stop), but the residuals of all that still refine come from one pass
    // over A64 per step: each element is read once and feeds up to 4 FMAs in
    // the same order as a pass per right-hand side would
    const int n = S.n, lda = S.lda, tid = threadIdx.x, nr = S.nrhs;
    double dm = 0.0;
    for (int i = tid; i < n; i += kBT) dm = fmax(dm, fabs(d[i]));
    const double norma = 2.0 * block_max(dm, red);          // hpl_infnorm, panel_norm.hpp:78-84
    double normb[4] = {0.0, 0.0, 0.0, 0.0}, h[4] = {-1.0, -1.0, -1.0, -1.0};
    int steps[4] = {0, 0, 0, 0};
    bool done[4] = {false, false, false, false};
    for (int rr = 0; rr < nr; ++rr) {
        const double *b = B + static_cast<int64_t>(rr) * n;
        double *x = X + static_cast<int64_t>(rr) * n;
        double bm = 0.0;
        for (int i = tid; i < n; i += kBT) {
            x[i] = b[i] / d[i];                              // x0 = b ./ diag(A)
            bm = fmax(bm, fabs(b[i]));
        }
        normb[rr] = block_max(bm, red);
    }
    for (int it = 0; it < maxit; ++it) {
        // r = b - A x in FP64 for the right-hand sides still refining (x staged in LDS)
        for (int rr = 0; rr < nr; ++rr)
            if (!done[rr])
                for (int i = tid; i < n; i += kBT) xl[rr * n + i] = X[static_cast<int64_t>(rr) * n + i];
        __syncthreads();
        double xm[4] = {0.0, 0.0, 0.0, 0.0}, rm[4] = {0.0, 0.0, 0.0, 0.0};
        for (int i = tid; i < n; i += kBT) {
            double acc[4];
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) acc[rr] = rr < nr ? B[static_cast<int64_t>(rr) * n + i] : 0.0;
            // j ascending; the loads of 16 columns are issued together
            const double *arow = A64 + i;
            int j = 0;
            for (; j + 16 <= n; j += 16) {
                double a[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) a[q] = arow[static_cast<int64_t>(j + q) * lda];
#pragma unroll
                for (int q = 0; q < 16; ++q)
#pragma unroll
                    for (int rr = 0; rr < 4; ++rr)
                        if (rr < nr) acc[rr] = fma(-a[q], xl[rr * n + j + q], acc[rr]);
            }
            for (; j < n; ++j) {
                const double a = arow[static_cast<int64_t>(j) * lda];
#pragma unroll
                for (int rr = 0; rr < 4; ++rr)
                    if (rr < nr) acc[rr] = fma(-a, xl[rr * n + j], acc[rr]);
            }
#pragma unroll
            for (int rr = 0; rr < 4; ++rr) {
                if (rr < nr && !done[rr]) {
                    R[static_cast<int64_t>(rr) * n + i] = acc[rr];
                    rm[rr] = fmax(rm[rr], fabs(acc[rr]));
                    xm[rr] = fmax(xm[rr], fabs(xl[rr * n + i]));
                }
            }
        }
        bool all = true;
        for (int rr = 0; rr < nr; ++rr) {
            if (done[rr]) continue;
            const double normr = block_max(rm[rr], red);
            const double normx = block_max(xm[rr], red);
            h[rr] = normr / (norma * normx + normb[rr]) * 1. / (static_cast<double>(n) * DBL_EPSILON / 2);
            steps[rr] = it + 1;
            done[rr] = h[rr] < 16.0;
            all = all && done[rr];
        }
        if (all) break;
        // corrections of all right-hand sides in one pass over each triangle
        // (the ones already done are solved too and not used)
        batch_trsv_all<false, true>(n, lda, LU, R, nr, xs);
        batch_trsv_all<true, false>(n, lda, LU, R, nr, xs);
        for (int rr = 0; rr < nr; ++rr) {
            if (done[rr]) continue;
            const double *r = R + static_cast<int64_t>(rr) * n;
            double *x = X + static_cast<int64_t>(rr) * n;
            for (int i = tid; i < n; i += kBT) x[i] += r[i];
        }
        __syncthreads();
    }
    if (tid == 0)
        for (int rr = 0; rr < nr; ++rr) {
            o.harness[rr] = h[rr];
            o.steps[rr] = steps[rr];
        }
    __syncthreads();
}

__global__ __launch_bounds__(kBT) void batch_kernel(const BatchSys *__restrict__ sys,
                                                    const int32_t *__restrict__ order, int count,
                                                    float *__restrict__ fpool, double *__restrict__ dpool,
                                                    BatchOut *__restrict__ out, int32_t *counter, int maxit)
{
    extern __shared__ __attribute__((aligned(16))) uint8_t lds[];
    __shared__ double xs[4 * 32];
    __shared__ double red[4];
    __shared__ int32_t next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int q = next;
        __syncthreads();
        if (q >= count) break;
        const int sid = order[q];
        const BatchSys S = sys[sid];
        const uint64_t t0 = wall_clock64();
        float *A32 = fpool + S.a32;
        double *A64 = dpool + S.a64, *d = dpool + S.d, *B = dpool + S.b, *X = dpool + S.x,
               *R = dpool + S.r;
        batch_generate(S, A32, A64, d, B);
        int info;
        if (S.nb == 16)
            info = batch_lu<16>(S.n, S.lda, A32, lds);
        else if (S.nb == 64)
            info = batch_lu<64>(S.n, S.lda, A32, lds);
        else
            info = batch_lu<32>(S.n, S.lda, A32, lds);
        BatchOut o;
        o.info = info;
        if (info == 0)
            batch_refine(S, A32, A64, d, B, X, R, reinterpret_cast<double *>(lds), xs, red, o, maxit);
        __syncthreads();
        if (threadIdx.x == 0) {
            BatchOut &g = out[sid];
            for (int r = 0; r < 4; ++r) {
                g.harness[r] = (info == 0 && r < S.nrhs) ? o.harness[r] : -1.0;
                g.steps[r] = (info == 0 && r < S.nrhs) ? o.steps[r] : 0;
            }
            g.info = info;
            g.t0 = t0;
            g.t1 = wall_clock64();
        }
    }
}

__global__ void batch_reset_kernel(int32_t *counter) { *counter = 0; }

}  // namespace hplai

// ===========================================================================
// host side
// ===========================================================================
using namespace hplai;

struct hplai_batch_s {
    hplai_ctx ctx = nullptr;
    int count = 0;
    std::vector<int32_t> n, nrhs;
    std::vector<uint64_t> seed;
    std::vector<BatchSys> sys;
    std::vector<int32_t> order;           // largest first
    int32_t nb_table[kBuckets] = {32, 32, 32, 32};
    BatchSys *d_sys = nullptr;
    int32_t *d_order = nullptr, *d_counter = nullptr;   // d_counter[0]: one launch; [1 + s]: system s alone
    int32_t *d_ident = nullptr;                          // systems in input order
    hipStream_t streams[4] = {};                          // baseline: per-system launches on 4 streams
    float *fpool = nullptr;
    double *dpool = nullptr;
    BatchOut *d_out = nullptr;
    std::vector<BatchOut> out;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_maxit = -1;
    double clock_khz = 100000.0;
};

namespace {

int bucket_of(int n) { return n < 128 ? 0 : n < 256 ? 1 : n < 512 ? 2 : 3; }

bool nb_fits(int nb, int n) { return batch_lu_lds(n, nb) <= kBatchLds; }

int upload_sys(hplai_batch b)
{
    for (int s = 0; s < b->count; ++s) {
        int nb = b->nb_table[bucket_of(b->n[s])];
        while (nb > 16 && !nb_fits(nb, b->n[s])) nb /= 2;
        b->sys[s].nb = nb;
    }
    HPLAI_HIP(hipMemcpy(b->d_sys, b->sys.data(), sizeof(BatchSys) * b->count, hipMemcpyHostToDevice));
    return HPLAI_OK;
}

int launch_batch(hplai_batch b, const int32_t *d_order, int count, int grid, int maxit, hipStream_t s,
                 int32_t *counter = nullptr)
{
    if (counter == nullptr) counter = b->d_counter;
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(1), 0, s, counter);
    hipLaunchKernelGGL(batch_kernel, dim3(grid), dim3(kBT), kBatchLds, s, b->d_sys, d_order, count, b->fpool,
                       b->dpool, b->d_out, counter, maxit);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

}  // namespace

extern "C" {

int hplai_batch_create(hplai_ctx ctx, int count, const int32_t *n, const int32_t *nrhs, const uint64_t *seed,
                       hplai_batch *out)
{
    HPLAI_CHECK(ctx != nullptr && out != nullptr && n != nullptr && nrhs != nullptr && seed != nullptr,
                "hplai_batch_create: NULL argument");
    HPLAI_CHECK(count > 0, "hplai_batch_create: count %d", count);
    for (int s = 0; s < count; ++s) {
        HPLAI_CHECK(n[s] >= 1 && n[s] <= kBatchMaxN, "hplai_batch_create: system %d has n=%d (1..%d)", s, n[s],
                    kBatchMaxN);
        HPLAI_CHECK(nrhs[s] >= 1 && nrhs[s] <= 4, "hplai_batch_create: system %d has %d rhs (1..4)", s,
                    nrhs[s]);
    }
    HPLAI_HIP(hipSetDevice(ctx->device));
    hplai_batch b = new hplai_batch_s;
    b->ctx = ctx;
    b->count = count;
    b->n.assign(n, n + count);
    b->nrhs.assign(nrhs, nrhs + count);
    b->seed.assign(seed, seed + count);
    b->sys.resize(count);
    int64_t fo = 0, dof = 0;
    auto up = [](int64_t v) { return (v + 63) / 64 * 64; };
    for (int s = 0; s < count; ++s) {
        BatchSys &S = b->sys[s];
        S.n = n[s];
        S.nrhs = nrhs[s];
        S.nb = 32;
        S.lda = (n[s] + 3) / 4 * 4;
        const int64_t nn = static_cast<int64_t>(S.lda) * n[s];
        S.x1 = lcg_init(seed[s]);
        S.a32 = fo;
        fo += up(nn);
        S.a64 = dof;
        dof += up(nn);
        S.d = dof;
        dof += up(n[s]);
        S.b = dof;
        dof += up(static_cast<int64_t>(n[s]) * nrhs[s]);
        S.x = dof;
        dof += up(static_cast<int64_t>(n[s]) * nrhs[s]);
        S.r = dof;
        dof += up(static_cast<int64_t>(n[s]) * nrhs[s]);
    }
    b->order.resize(count);
    for (int s = 0; s < count; ++s) b->order[s] = s;
    std::stable_sort(b->order.begin(), b->order.end(), [&](int a, int c) { return b->n[a] > b->n[c]; });
    b->out.resize(count);
    int rc = HPLAI_OK;
    auto alloc = [&](void **p, size_t bytes) {
        if (rc == HPLAI_OK && hipMalloc(p, bytes) != hipSuccess) rc = hip_fail(hipGetLastError(), "hplai_batch_create: hipMalloc");
    };
    alloc(reinterpret_cast<void **>(&b->d_sys), sizeof(BatchSys) * count);
    alloc(reinterpret_cast<void **>(&b->d_order), sizeof(int32_t) * count);
    alloc(reinterpret_cast<void **>(&b->d_counter), sizeof(int32_t) * (count + 1));
    alloc(reinterpret_cast<void **>(&b->d_ident), sizeof(int32_t) * count);
    alloc(reinterpret_cast<void **>(&b->fpool), sizeof(float) * fo);
    alloc(reinterpret_cast<void **>(&b->dpool), sizeof(double) * dof);
    alloc(reinterpret_cast<void **>(&b->d_out), sizeof(BatchOut) * count);
    if (rc == HPLAI_OK &&
        hipMemcpy(b->d_order, b->order.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice) != hipSuccess)
        rc = hip_fail(hipGetLastError(), "hplai_batch_create: copy");
    if (rc == HPLAI_OK) {
        std::vector<int32_t> ident(count);
        for (int q = 0; q < count; ++q) ident[q] = q;
        if (hipMemcpy(b->d_ident, ident.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice) != hipSuccess)
            rc = hip_fail(hipGetLastError(), "hplai_batch_create: copy");
    }
    for (int q = 0; q < 4 && rc == HPLAI_OK; ++q)
        if (hipStreamCreateWithFlags(&b->streams[q], hipStreamNonBlocking) != hipSuccess)
            rc = hip_fail(hipGetLastError(), "hplai_batch_create: streams");
    if (rc == HPLAI_OK) rc = upload_sys(b);
    if (rc == HPLAI_OK &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(batch_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, kBatchLds) != hipSuccess)
        rc = hip_fail(hipGetLastError(), "hplai_batch_create: LDS size");
    int khz = 0;
    if (rc == HPLAI_OK && hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, ctx->device) == hipSuccess &&
        khz > 0)
        b->clock_khz = khz;
    if (rc != HPLAI_OK) {
        hplai_batch_destroy(b);
        return rc;
    }
    *out = b;
    return HPLAI_OK;
}

int hplai_batch_destroy(hplai_batch b)
{
    if (b == nullptr) return HPLAI_OK;
    (void)hipSetDevice(b->ctx->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    (void)hipFree(b->d_sys);
    (void)hipFree(b->d_order);
    (void)hipFree(b->d_counter);
    (void)hipFree(b->d_ident);
    for (int q = 0; q < 4; ++q)
        if (b->streams[q]) (void)hipStreamDestroy(b->streams[q]);
    (void)hipFree(b->fpool);
    (void)hipFree(b->dpool);
    (void)hipFree(b->d_out);
    delete b;
    return HPLAI_OK;
}

int hplai_batch_set_nb(hplai_batch b, const int32_t nb[4])
{
    HPLAI_CHECK(b != nullptr && nb != nullptr, "hplai_batch_set_nb: NULL argument");
    for (int k = 0; k < kBuckets; ++k)
        HPLAI_CHECK(nb[k] == 16 || nb[k] == 32 || nb[k] == 64, "hplai_batch_set_nb: nb %d not in {16,32,64}",
                    nb[k]);
    for (int k = 0; k < kBuckets; ++k) b->nb_table[k] = nb[k];
    // a captured graph stays valid: it reads the descriptors through the same pointer
    return upload_sys(b);
}

int hplai_batch_get_nb(hplai_batch b, int32_t nb[4])
{
    HPLAI_CHECK(b != nullptr && nb != nullptr, "hplai_batch_get_nb: NULL argument");
    for (int k = 0; k < kBuckets; ++k) nb[k] = b->nb_table[k];
    return HPLAI_OK;
}

int hplai_batch_run(hplai_batch b, int mode, int maxit, hplai_batch_stats *st)
{
    HPLAI_CHECK(b != nullptr && st != nullptr, "hplai_batch_run: NULL argument");
    HPLAI_CHECK(mode >= 0 && mode <= 4, "hplai_batch_run: mode %d", mode);
    if (maxit <= 0) maxit = 50;
    HPLAI_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->side;
    const int grid = b->ctx->num_cu > 0 ? b->ctx->num_cu : 256;
    HPLAI_HIP(hipStreamSynchronize(s));
    const auto t0 = std::chrono::steady_clock::now();
    int rc = HPLAI_OK;
    if (mode == 0) {
        rc = launch_batch(b, b->d_order, b->count, grid, maxit, s);
    } else if (mode == 1) {
        if (b->exec == nullptr || b->graph_maxit != maxit) {
            if (b->exec) (void)hipGraphExecDestroy(b->exec);
            if (b->graph) (void)hipGraphDestroy(b->graph);
            b->exec = nullptr;
            b->graph = nullptr;
            HPLAI_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            rc = launch_batch(b, b->d_order, b->count, grid, maxit, s);
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(s, &g);
            if (rc == HPLAI_OK && e != hipSuccess) rc = hip_fail(e, "hplai_batch_run: capture");
            if (rc == HPLAI_OK) {
                b->graph = g;
                const hipError_t e2 = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
                if (e2 != hipSuccess) rc = hip_fail(e2, "hplai_batch_run: instantiate");
                b->graph_maxit = maxit;
            } else if (g) {
                (void)hipGraphDestroy(g);
            }
        }
        if (rc == HPLAI_OK) HPLAI_HIP(hipGraphLaunch(b->exec, s));
    } else if (mode == 2) {
        // baseline: one launch (one workgroup) per system, in the same order
        for (int q = 0; q < b->count && rc == HPLAI_OK; ++q)
            rc = launch_batch(b, b->d_order + q, 1, 1, maxit, s);
    } else if (mode == 3) {
        // baseline: one launch per system, round-robin over four streams (the
        // process has four hardware queues), each launch with its own counter
        for (int q = 0; q < b->count && rc == HPLAI_OK; ++q)
            rc = launch_batch(b, b->d_order + q, 1, 1, maxit, b->streams[q & 3], b->d_counter + 1 + q);
        for (int q = 0; q < 4; ++q) HPLAI_HIP(hipStreamSynchronize(b->streams[q]));
    } else {
        // baseline: one launch of one workgroup per system in input order (a
        // plain batched grid: no largest-first order, no persistent workgroups)
        rc = launch_batch(b, b->d_ident, b->count, b->count, maxit, s);
    }
    HPLAI_HIP(hipStreamSynchronize(s));
    const auto t1 = std::chrono::steady_clock::now();
    if (rc) return rc;
    HPLAI_HIP(hipMemcpy(b->out.data(), b->d_out, sizeof(BatchOut) * b->count, hipMemcpyDeviceToHost));
    std::vector<double> us(b->count);
    int failed = 0, max_steps = 0;
    double max_h = 0.0;
    for (int s2 = 0; s2 < b->count; ++s2) {
        const BatchOut &o = b->out[s2];
        us[s2] = static_cast<double>(o.t1 - o.t0) / b->clock_khz * 1e3;
        bool ok = o.info == 0;
        for (int r = 0; r < b->nrhs[s2]; ++r) {
            if (!(o.harness[r] >= 0.0 && o.harness[r] < 16.0)) ok = false;
            max_h = std::max(max_h, o.harness[r]);
            max_steps = std::max(max_steps, o.steps[r]);
        }
        if (!ok) ++failed;
    }
    std::vector<double> sorted = us;
    std::sort(sorted.begin(), sorted.end());
    auto pct = [&](double p) { return sorted[std::min<size_t>(sorted.size() - 1, static_cast<size_t>(p * sorted.size()))]; };
    st->seconds = std::chrono::duration<double>(t1 - t0).count();
    st->systems_per_s = b->count / st->seconds;
    st->p50_us = pct(0.50);
    st->p99_us = pct(0.99);
    st->max_us = sorted.back();
    st->failed = failed;
    st->max_harness = max_h;
    st->max_steps = max_steps;
    return HPLAI_OK;
}

int hplai_batch_results(hplai_batch b, double *harness, int32_t *steps, int32_t *info, double *time_us)
{
    HPLAI_CHECK(b != nullptr, "hplai_batch_results: NULL batch");
    for (int s = 0; s < b->count; ++s) {
        const BatchOut &o = b->out[s];
        for (int r = 0; r < 4; ++r) {
            if (harness) harness[4 * s + r] = o.harness[r];
            if (steps) steps[4 * s + r] = o.steps[r];
        }
        if (info) info[s] = o.info;
        if (time_us) time_us[s] = static_cast<double>(o.t1 - o.t0) / b->clock_khz * 1e3;
    }
    return HPLAI_OK;
}

int hplai_batch_solution(hplai_batch b, int s, double *x_host)
{
    HPLAI_CHECK(b != nullptr && x_host != nullptr && s >= 0 && s < b->count, "hplai_batch_solution: bad argument");
    HPLAI_HIP(hipSetDevice(b->ctx->device));
    const BatchSys &S = b->sys[s];
    HPLAI_HIP(hipMemcpy(x_host, b->dpool + S.x, sizeof(double) * S.n * S.nrhs, hipMemcpyDeviceToHost));
    return HPLAI_OK;
}

int hplai_batch_autotune(hplai_batch b, int maxit, int32_t nb_out[4])
{
    HPLAI_CHECK(b != nullptr, "hplai_batch_autotune: NULL batch");
    if (maxit <= 0) maxit = 50;
    HPLAI_HIP(hipSetDevice(b->ctx->device));
    hipStream_t s = b->ctx->side;
    const int grid = b->ctx->num_cu > 0 ? b->ctx->num_cu : 256;
    const int cands[3] = {16, 32, 64};
    int32_t best[kBuckets];
    for (int k = 0; k < kBuckets; ++k) best[k] = b->nb_table[k];
    int32_t *d_sub = nullptr;
    HPLAI_HIP(hipMalloc(&d_sub, sizeof(int32_t) * b->count));
    int rc = HPLAI_OK;
    for (int k = 0; k < kBuckets && rc == HPLAI_OK; ++k) {
        // up to 2 systems per CU of this bucket, largest first
        std::vector<int32_t> sub;
        int maxn = 0;
        for (int q = 0; q < b->count && static_cast<int>(sub.size()) < 2 * grid; ++q)
            if (bucket_of(b->n[b->order[q]]) == k) {
                sub.push_back(b->order[q]);
                maxn = std::max(maxn, b->n[b->order[q]]);
            }
        if (sub.empty()) continue;
        if (hipMemcpy(d_sub, sub.data(), sizeof(int32_t) * sub.size(), hipMemcpyHostToDevice) != hipSuccess) {
            rc = hip_fail(hipGetLastError(), "hplai_batch_autotune: copy");
            break;
        }
        double best_t = 1e300;
        for (int c = 0; c < 3 && rc == HPLAI_OK; ++c) {
            const int nb = cands[c];
            if (!nb_fits(nb, maxn)) continue;
            for (int32_t sid : sub) b->sys[sid].nb = nb;
            if (hipMemcpy(b->d_sys, b->sys.data(), sizeof(BatchSys) * b->count, hipMemcpyHostToDevice) != hipSuccess) {
                rc = hip_fail(hipGetLastError(), "hplai_batch_autotune: copy");
                break;
            }
            // one untimed pass, then the timed one
            double t = 0.0;
            for (int pass = 0; pass < 2 && rc == HPLAI_OK; ++pass) {
                const auto t0 = std::chrono::steady_clock::now();
                rc = launch_batch(b, d_sub, static_cast<int>(sub.size()), grid, maxit, s);
                if (rc == HPLAI_OK && hipStreamSynchronize(s) != hipSuccess)
                    rc = hip_fail(hipGetLastError(), "hplai_batch_autotune: run");
                t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            }
            if (rc == HPLAI_OK && t < best_t) {
                best_t = t;
                best[k] = nb;
            }
        }
    }
    (void)hipFree(d_sub);
    if (rc) return rc;
    for (int k = 0; k < kBuckets; ++k) b->nb_table[k] = best[k];
    if (nb_out)
        for (int k = 0; k < kBuckets; ++k) nb_out[k] = best[k];
    return upload_sys(b);
}

}  // extern "C"
