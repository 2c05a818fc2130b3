// trsm.hip — panel TRSMs of the blocked LU with the whole panel height in LDS.
//
// One workgroup owns a strip of 32 columns (U panel) or 32 rows (L panel) of
// the panel (64 for the triangular products when b % 64 != 0) and all
// b <= 512 entries of its other dimension, in LDS.
//  * b % 32 == 0: the triangles of the factored diagonal block are inverted
//    in full once per block (tri_inv_kernel, after the 32 x 32 sub-block
//    inverses of diag_inv32_kernel), and every strip is a triangular matrix
//    product with the inverse on the FP32 MFMA (trsm_l_tri_kernel,
//    trsm_u_tri_kernel).
//  * otherwise: the strip is solved 32 at a time: the 32 x 32 diagonal
//    sub-block through its inverse, then a rank-32 update of the rest of the
//    strip (trsm_l_strip_kernel, trsm_u_strip_kernel).
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
// MFMA forms (b % 32 == 0, b <= 512).  The diagonal block's triangles are
// inverted in full once per factored block (tri_inv_kernel); each panel solve
// is then a triangular matrix product (TRMM) with that inverse,
//   L := A Uinv,  U := Linv A,
// whose 32 x 32 output blocks are independent, so the four waves of a
// workgroup run without barriers between them instead of walking the
// 32-wide substitution chain.  Products run on v_mfma_f32_32x32x2_f32 (FP32
// inputs and accumulation).  The strip (32 rows or columns, all b of the
// other dimension) is in LDS; the inverse is read from global memory
// (L2-resident) in MFMA operand order, 16 consecutive floats per lane.
// MFMA layout: A operand lane -> (row lane % 32, k 2 kk + lane / 32), B
// operand lane -> (k 2 kk + lane / 32, column lane % 32); output acc[4 q + r]
// is row 8 q + 4 (lane / 32) + r, column lane % 32.
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

// Inverse layout after the nb = b / 32 sub-block inverses (kInvBlock each):
//   LinvP: block (I, J) at (I nb + J) 1024, [lane 16 + kk] = Linv(32 I + lane % 32, 32 J + 2 kk + lane / 32)
//          (A-operand order for U := Linv A)
//   UinvP: block (I, J) at nb^2 1024 + (J nb + I) 1024,
//          [lane 16 + kk] = Uinv(32 I + 2 kk + lane / 32, 32 J + lane % 32)
//          (B-operand order for L := A Uinv)
// Blocks above the diagonal of Linv and below the diagonal of Uinv are never
// written or read.
__device__ __forceinline__ void tri_store(float *tri, int nb, bool upper, int I, int J, int rho, int gam, float v)
{
    if (upper)
        tri[static_cast<int64_t>(nb) * nb * 1024 + (static_cast<int64_t>(J) * nb + I) * 1024 +
            (gam + 32 * (rho & 1)) * 16 + (rho >> 1)] = v;
    else
        tri[(static_cast<int64_t>(I) * nb + J) * 1024 + (rho + 32 * (gam & 1)) * 16 + (gam >> 1)] = v;
}

// One workgroup per 32-column block j of Linv (blockIdx.y = 0) or Uinv
// (blockIdx.y = 1).  From L Linv = I, top to bottom:
//   Linv(i, j) = -Dinv_i sum_{k=j}^{i-1} L(i, k) Linv(k, j),   Linv(j, j) = Dinv_j;
// from U Uinv = I, bottom to top:
//   Uinv(i, j) = -Einv_i sum_{k=i+1}^{j} U(i, k) Uinv(k, j),   Uinv(j, j) = Einv_j,
// with Dinv, Einv the 32 x 32 sub-block inverses.  The four waves split the
// sum over k and add their partial products in LDS.
__global__ __launch_bounds__(256) void tri_inv_kernel(int b, const float *__restrict__ LU, int64_t lda,
                                                      const float *__restrict__ inv32, float *__restrict__ tri)
{
    __shared__ float W[kPanelMax * kXS];   // W[r * kXS + c] = inverse(r, 32 j + c)
    __shared__ float P[4 * 1024];          // per-wave partial sums, lane-major
    __shared__ float S[kStrip * kXS];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, lh = lane >> 5;
    const int nb = b / kStrip, j = blockIdx.x;
    const bool upper = blockIdx.y != 0;
    const int half = upper ? 1024 : 0;
    {
        const float *dj = inv32 + static_cast<int64_t>(j) * kInvBlock + half;   // row-major 32 x 32
        for (int e = tid; e < 1024; e += 256) {
            const int r = e / 32, c = e % 32;
            const float v = dj[e];
            W[(32 * j + r) * kXS + c] = v;
            tri_store(tri, nb, upper, j, j, r, c, v);
        }
    }
    __syncthreads();
    const int steps = upper ? j : nb - 1 - j;
    // the L (U) blocks of a step do not depend on the step before it: each
    // wave loads its (at most four, nb <= 16) blocks of step st + 1 while step
    // st is reduced and multiplied by the sub-block inverse
    auto load_blocks = [&](int st, float (&av)[4][16]) {
        const int i = upper ? j - st : j + st;
        const int klo = upper ? i + 1 : j, khi = upper ? j : i - 1;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = klo + wave + 4 * u;
            if (k <= khi) {
                // A = L(i, k) or U(i, k): element (32 i + l32, 32 k + 2 kk + lh)
                const float *src = LU + static_cast<int64_t>(32 * k + lh) * lda + 32 * i + l32;
#pragma unroll
                for (int kk = 0; kk < 16; ++kk) av[u][kk] = src[static_cast<int64_t>(2 * kk) * lda];
            }
        }
    };
    float avc[4][16], avn[4][16];
    if (steps >= 1) load_blocks(1, avc);
    for (int st = 1; st <= steps; ++st) {
        const int i = upper ? j - st : j + st;
        const int klo = upper ? i + 1 : j, khi = upper ? j : i - 1;
        if (st < steps) load_blocks(st + 1, avn);
        floatx16t_t acc;
#pragma unroll
        for (int e = 0; e < 16; ++e) acc[e] = 0.0f;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = klo + wave + 4 * u;
            if (k <= khi) {
#pragma unroll
                for (int kk = 0; kk < 16; ++kk)
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(avc[u][kk], W[(32 * k + 2 * kk + lh) * kXS + l32],
                                                               acc, 0, 0, 0);
            }
        }
#pragma unroll
        for (int e = 0; e < 16; ++e) P[wave * 1024 + lane * 16 + e] = acc[e];
        __syncthreads();
        if (wave == 0) {
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int e = lane * 16 + 4 * q + r;
                    S[(8 * q + 4 * lh + r) * kXS + l32] = P[e] + P[1024 + e] + P[2048 + e] + P[3072 + e];
                }
        }
        __syncthreads();
        if (wave == 0) {
            float av[16];
            load_k_pairs(inv32 + static_cast<int64_t>(i) * kInvBlock + half + l32 * 32, lh, av);
            floatx16t_t r2;
#pragma unroll
            for (int e = 0; e < 16; ++e) r2[e] = 0.0f;
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
                r2 = __builtin_amdgcn_mfma_f32_32x32x2f32(av[kk], S[(2 * kk + lh) * kXS + l32], r2, 0, 0, 0);
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int rho = 8 * q + 4 * lh + r;
                    const float v = -r2[4 * q + r];
                    W[(32 * i + rho) * kXS + l32] = v;
                    tri_store(tri, nb, upper, i, j, rho, l32, v);
                }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) avc[u][kk] = avn[u][kk];
    }
}

// Snake assignment of output blocks to the eight waves, largest first: block
// nb - 1 - p goes to wave snake(p), so every wave gets about the same number
// of 32 x 32 x 32 products (block q costs q + 1 of them; with nb = 16 every
// wave gets 17).
__device__ __forceinline__ int tri_snake(int p) { return ((p >> 3) & 1) ? 7 - (p & 7) : (p & 7); }

constexpr int kTriThreads = 512;

// Strip width S (rows of the L panel or columns of the U panel per
// workgroup): S = 64 shares every inverse operand between two 32-wide halves
// (half the operand traffic per flop) but its LDS strip (b x 65 floats) allows
// one workgroup per CU, so a strip's loads and its products cannot overlap;
// S = 32 halves the strip (b x 33 floats) so two workgroups share a CU.
template <int S> struct TriStrip {
    static constexpr int XS = S + 1;          // LDS stride (odd: rows spread over the banks)
    static constexpr int NH = S / 32;         // 32-wide halves
};

// Strip loads into LDS: all of a thread's 16-byte loads are issued at once
// (one memory latency per strip), and the 32 lanes of a half-wave write 32
// distinct banks: each half-wave covers 8 row chunks i (4 rows each) of 4
// columns j, which land on bank (4 i + j + const) mod 32 (XS = 1 mod 32).
// L strip (S rows x b columns, column-major, 16-byte aligned): X[c * XS + r].
template <int S>
__device__ __forceinline__ void load_strip_l(const float *__restrict__ L, int64_t ldl, int64_t i0, int b,
                                             float *X, int wave, int lane)
{
    constexpr int XS = TriStrip<S>::XS;
    constexpr int CPI = 256 / S;                 // columns per wave instruction (S / 4 row chunks each)
    // S = 64: a half-wave takes row chunks 0..7 (lanes 0-31) or 8..15 of 4 columns
    // S = 32: lanes 0-31 and 32-63 take columns 0..3 and 4..7 (8 row chunks each)
    const int i = S == 64 ? (lane & 7) + 8 * (lane >> 5) : (lane & 7);
    const int j = S == 64 ? ((lane >> 3) & 3) : (lane >> 3);
    const int per_wave = b / (8 * CPI);          // instructions per wave (b / CPI in all)
    float4 v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        if (u < per_wave) {
            const int c = CPI * (wave * per_wave + u) + j;
            v[u] = *reinterpret_cast<const float4 *>(L + static_cast<int64_t>(c) * ldl + i0 + 4 * i);
        }
    }
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        if (u < per_wave) {
            float *d = X + (CPI * (wave * per_wave + u) + j) * XS + 4 * i;
            d[0] = v[u].x;
            d[1] = v[u].y;
            d[2] = v[u].z;
            d[3] = v[u].w;
        }
    }
}

// the last strip of the panel (nr < S rows; rows past nr read as zero)
template <int S>
__device__ __forceinline__ void load_strip_l_ragged(const float *__restrict__ L, int64_t ldl, int64_t i0, int nr,
                                                    int b, float *X, int tid)
{
    constexpr int XS = TriStrip<S>::XS;
    const int tot = b * S;
    for (int base = 0; base < tot; base += kTriThreads * 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = min(base + u * kTriThreads + tid, tot - 1);
            const int c = idx / S, r = idx % S;
            v[u] = L[static_cast<int64_t>(c) * ldl + i0 + min(r, nr - 1)];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = base + u * kTriThreads + tid;
            const int c = idx / S, r = idx % S;
            if (idx < tot) X[c * XS + r] = r < nr ? v[u] : 0.0f;
        }
    }
}

// U strip (b rows x S columns, column-major, 16-byte aligned): X[r * XS + c];
// columns past nc read as zero.  Wave w loads columns (S / 8) w .. + S / 8 - 1.
template <int S>
__device__ __forceinline__ void load_strip_u(const float *__restrict__ U, int64_t ldu, int64_t j0, int nc, int b,
                                             float *X, int wave, int lane)
{
    constexpr int XS = TriStrip<S>::XS;
    constexpr int CPW = S / 8;                   // columns per wave
    // S = 64: 8 columns x 8 row chunks per instruction; S = 32: 4 columns x 16 row chunks
    const int i = S == 64 ? (lane & 7) : (lane & 7) + 8 * (lane >> 5);
    const int j = S == 64 ? (lane >> 3) : ((lane >> 3) & 3);
    constexpr int RPI = 256 / CPW;               // rows per instruction
    const int c = CPW * wave + j;
    const float *col = U + (j0 + min(c, nc - 1)) * ldu;
    const int groups = b / RPI;
    float4 v[16];
#pragma unroll
    for (int u = 0; u < 16; ++u)
        if (u < groups) v[u] = *reinterpret_cast<const float4 *>(col + RPI * u + 4 * i);
    const bool in = c < nc;
#pragma unroll
    for (int u = 0; u < 16; ++u) {
        if (u < groups) {
            float *d = X + (RPI * u + 4 * i) * XS + c;
            d[0] = in ? v[u].x : 0.0f;
            d[XS] = in ? v[u].y : 0.0f;
            d[2 * XS] = in ? v[u].z : 0.0f;
            d[3 * XS] = in ? v[u].w : 0.0f;
        }
    }
}

template <int S>
__device__ __forceinline__ void load_strip_u_scalar(const float *__restrict__ U, int64_t ldu, int64_t j0, int nc,
                                                    int b, float *X, int tid)
{
    constexpr int XS = TriStrip<S>::XS;
    const int tot = b * S;
    for (int base = 0; base < tot; base += kTriThreads * 16) {
        float v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = min(base + u * kTriThreads + tid, tot - 1);
            const int c = idx / b, r = idx % b;
            v[u] = U[(j0 + min(c, nc - 1)) * ldu + r];
        }
#pragma unroll
        for (int u = 0; u < 16; ++u) {
            const int idx = base + u * kTriThreads + tid;
            const int c = idx / b, r = idx % b;
            if (idx < tot) X[r * XS + c] = c < nc ? v[u] : 0.0f;
        }
    }
}

// L(m x b) := L Ukk^-1 (update_left_panels, panel_trf.hpp:143-157) as
// L := L Uinv: output block jb = sum_{kb <= jb} X(:, kb) Uinv(kb, jb).  A
// workgroup owns S rows (S / 32 halves sharing every Uinv operand).
template <int S, bool VEC>
__global__ __launch_bounds__(kTriThreads) void trsm_l_tri_kernel(int64_t m, int b, const float *__restrict__ tri,
                                                                 float *__restrict__ L, int64_t ldl,
                                                                 float *__restrict__ amax, int tb)
{
    constexpr int XS = TriStrip<S>::XS, NH = TriStrip<S>::NH;
    __shared__ float X[kPanelMax * XS];   // X[c * XS + r]: column c, row i0 + r
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, lh = lane >> 5;
    const int64_t i0 = static_cast<int64_t>(blockIdx.x) * S;
    const int nr = m - i0 < S ? static_cast<int>(m - i0) : S;
    if (VEC && nr == S)
        load_strip_l<S>(L, ldl, i0, b, X, wave, lane);
    else
        load_strip_l_ragged<S>(L, ldl, i0, nr, b, X, tid);
    __syncthreads();
    const int nb = b / kStrip;
    const float *Up = tri + static_cast<int64_t>(nb) * nb * 1024;
    float vmax[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) vmax[h] = 0.0f;
    for (int p = 0; p < nb; ++p) {
        if (tri_snake(p) != wave) continue;
        const int jb = nb - 1 - p;
        const float4 *bp = reinterpret_cast<const float4 *>(Up + static_cast<int64_t>(jb) * nb * 1024) + lane * 4;
        floatx16t_t acc[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[h][e] = 0.0f;
        // the Uinv operand two blocks ahead (L2 latency under a full GPU)
        float4 cur[4], nxt[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) cur[e] = bp[e];
#pragma unroll
        for (int e = 0; e < 4; ++e) nxt[e] = bp[static_cast<int64_t>(min(1, jb)) * 256 + e];
        for (int kb = 0; kb <= jb; ++kb) {
            float4 nx2[4];
            const float4 *bn = bp + static_cast<int64_t>(min(kb + 2, jb)) * 256;
#pragma unroll
            for (int e = 0; e < 4; ++e) nx2[e] = bn[e];
            float xv[NH][16];
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) {
                const float *xr = X + (32 * kb + 2 * kk + lh) * XS + l32;
#pragma unroll
                for (int h = 0; h < NH; ++h) xv[h][kk] = xr[32 * h];
            }
            // keeps the operand loads where they are: without it the compiler
            // re-issues each global load right before its MFMA and waits on it
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
            const float bv[16] = {cur[0].x, cur[0].y, cur[0].z, cur[0].w, cur[1].x, cur[1].y, cur[1].z, cur[1].w,
                                  cur[2].x, cur[2].y, cur[2].z, cur[2].w, cur[3].x, cur[3].y, cur[3].z, cur[3].w};
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
#pragma unroll
                for (int h = 0; h < NH; ++h)
                    acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(xv[h][kk], bv[kk], acc[h], 0, 0, 0);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                cur[e] = nxt[e];
                nxt[e] = nx2[e];
            }
        }
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            float *dst = L + static_cast<int64_t>(32 * jb + l32) * ldl + i0 + 32 * h + 4 * lh;
            if (VEC && nr == S) {
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float4 o = {acc[h][4 * q], acc[h][4 * q + 1], acc[h][4 * q + 2], acc[h][4 * q + 3]};
                    *reinterpret_cast<float4 *>(dst + 8 * q) = o;
                    vmax[h] = fmaxf(vmax[h], fmaxf(fmaxf(fabsf(o.x), fabsf(o.y)), fmaxf(fabsf(o.z), fabsf(o.w))));
                }
            } else {
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        if (32 * h + 8 * q + 4 * lh + r < nr) {
                            dst[8 * q + r] = acc[h][4 * q + r];
                            vmax[h] = fmaxf(vmax[h], fabsf(acc[h][4 * q + r]));
                        }
            }
        }
    }
    // the halves may fall in different tiles when tb % 64 == 32
#pragma unroll
    for (int h = 0; h < NH; ++h) tile_max_reduce(amax, (i0 + 32 * h) / tb, vmax[h], 32 * h < nr);
}

// U(b x n) := Lkk^-1 U (update_right_panels, panel_trf.hpp:158-171) as
// U := Linv U: output block rb = sum_{kb <= rb} Linv(rb, kb) X(kb, :).  A
// workgroup owns S columns (S / 32 halves sharing every Linv operand).
template <int S, bool VEC>
__global__ __launch_bounds__(kTriThreads) void trsm_u_tri_kernel(int b, int64_t n, const float *__restrict__ tri,
                                                                 float *__restrict__ U, int64_t ldu,
                                                                 float *__restrict__ amax, int tb)
{
    constexpr int XS = TriStrip<S>::XS, NH = TriStrip<S>::NH;
    __shared__ float X[kPanelMax * XS];   // X[r * XS + c]: row r of column j0 + c
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, lh = lane >> 5;
    const int64_t j0 = static_cast<int64_t>(blockIdx.x) * S;
    const int nc = n - j0 < S ? static_cast<int>(n - j0) : S;
    if (VEC)
        load_strip_u<S>(U, ldu, j0, nc, b, X, wave, lane);
    else
        load_strip_u_scalar<S>(U, ldu, j0, nc, b, X, tid);
    __syncthreads();
    const int nb = b / kStrip;
    float vmax[NH];
#pragma unroll
    for (int h = 0; h < NH; ++h) vmax[h] = 0.0f;
    for (int p = 0; p < nb; ++p) {
        if (tri_snake(p) != wave) continue;
        const int rb = nb - 1 - p;
        const float4 *ap = reinterpret_cast<const float4 *>(tri + static_cast<int64_t>(rb) * nb * 1024) + lane * 4;
        floatx16t_t acc[NH];
#pragma unroll
        for (int h = 0; h < NH; ++h)
#pragma unroll
            for (int e = 0; e < 16; ++e) acc[h][e] = 0.0f;
        // the Linv operand two blocks ahead (L2 latency under a full GPU)
        float4 cur[4], nxt[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) cur[e] = ap[e];
#pragma unroll
        for (int e = 0; e < 4; ++e) nxt[e] = ap[static_cast<int64_t>(min(1, rb)) * 256 + e];
        for (int kb = 0; kb <= rb; ++kb) {
            float4 nx2[4];
            const float4 *an = ap + static_cast<int64_t>(min(kb + 2, rb)) * 256;
#pragma unroll
            for (int e = 0; e < 4; ++e) nx2[e] = an[e];
            float xv[NH][16];
#pragma unroll
            for (int kk = 0; kk < 16; ++kk) {
                const float *xr = X + (32 * kb + 2 * kk + lh) * XS + l32;
#pragma unroll
                for (int h = 0; h < NH; ++h) xv[h][kk] = xr[32 * h];
            }
            // keeps the operand loads where they are: without it the compiler
            // re-issues each global load right before its MFMA and waits on it
            asm volatile("" ::: "memory");
            __builtin_amdgcn_sched_barrier(0);
            const float av[16] = {cur[0].x, cur[0].y, cur[0].z, cur[0].w, cur[1].x, cur[1].y, cur[1].z, cur[1].w,
                                  cur[2].x, cur[2].y, cur[2].z, cur[2].w, cur[3].x, cur[3].y, cur[3].z, cur[3].w};
#pragma unroll
            for (int kk = 0; kk < 16; ++kk)
#pragma unroll
                for (int h = 0; h < NH; ++h)
                    acc[h] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[kk], xv[h][kk], acc[h], 0, 0, 0);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                cur[e] = nxt[e];
                nxt[e] = nx2[e];
            }
        }
#pragma unroll
        for (int h = 0; h < NH; ++h) {
            if (32 * h + l32 >= nc) continue;
            float *dst = U + (j0 + 32 * h + l32) * ldu + 32 * rb + 4 * lh;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (VEC) {
                    const float4 o = {acc[h][4 * q], acc[h][4 * q + 1], acc[h][4 * q + 2], acc[h][4 * q + 3]};
                    *reinterpret_cast<float4 *>(dst + 8 * q) = o;
                } else {
#pragma unroll
                    for (int r = 0; r < 4; ++r) dst[8 * q + r] = acc[h][4 * q + r];
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) vmax[h] = fmaxf(vmax[h], fabsf(acc[h][4 * q + r]));
            }
        }
    }
#pragma unroll
    for (int h = 0; h < NH; ++h) tile_max_reduce(amax, (j0 + 32 * h) / tb, vmax[h], 32 * h < nc);
}

// The strip kernels need b <= 512 and, when tile maxima are collected, every
// strip inside one tile (tb % 32 == 0).
bool trsm_strip_ok(int b, const float *amax, int tb)
{
    return b <= kPanelMax && (amax == nullptr || (tb % kStrip) == 0);
}

static bool tri_ok(int b) { return b % kStrip == 0 && b <= kPanelMax; }

int64_t diag_inv_floats(int b)
{
    const int64_t nb = ceil_div(b, kStrip);
    return nb * kInvBlock + (tri_ok(b) ? 2 * nb * nb * 1024 : 0);
}

int launch_diag_inv(int b, const float *LU, int64_t lda, float *inv, hipStream_t s)
{
    const int nb = static_cast<int>(ceil_div(b, kStrip));
    hipLaunchKernelGGL(diag_inv32_kernel, dim3(static_cast<unsigned>(nb)), dim3(256), 0, s, b, LU, lda, inv);
    HPLAI_HIP(hipGetLastError());
    if (tri_ok(b)) {
        hipLaunchKernelGGL(tri_inv_kernel, dim3(static_cast<unsigned>(nb), 2), dim3(256), 0, s, b, LU, lda, inv,
                           inv + static_cast<int64_t>(nb) * kInvBlock);
        HPLAI_HIP(hipGetLastError());
    }
    return HPLAI_OK;
}

// x_out := Tinv x_in for one factored diagonal block (the IR's diagonal-block
// solve as a product with the stored inverse): Linv (UPPER = false, the unit
// lower solve) or Uinv (UPPER = true).  Workgroup I computes rows
// 32 I .. 32 I + 31; its 256 threads are 32 rows x 8 column-block groups,
// FP32 inverse entries times the FP64 vector summed in FP64, the 8 partial
// sums of a row added in a fixed order.  Out of place: every workgroup reads
// the whole input segment.
template <bool UPPER>
__global__ __launch_bounds__(256) void tri_trsv_kernel(int b, const float *__restrict__ tri,
                                                       const double *__restrict__ xin, double *__restrict__ xout)
{
    __shared__ double xs[kPanelMax];
    __shared__ double part[8][kStrip];
    const int tid = threadIdx.x, rho = tid & 31, p = tid >> 5;
    const int nb = b / kStrip, I = blockIdx.x;
    for (int c = tid; c < b; c += 256) xs[c] = xin[c];
    __syncthreads();
    double acc = 0.0;
    if (!UPPER) {
        for (int J = p; J <= I; J += 8) {
            const float *blk = tri + (static_cast<int64_t>(I) * nb + J) * 1024;
            const float4 *ev = reinterpret_cast<const float4 *>(blk + rho * 16);          // gamma = 2 kk
            const float4 *od = reinterpret_cast<const float4 *>(blk + (rho + 32) * 16);   // gamma = 2 kk + 1
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float4 a = ev[e], c = od[e];
                const double *xj = xs + 32 * J + 8 * e;
                acc = fma(static_cast<double>(a.x), xj[0], acc);
                acc = fma(static_cast<double>(c.x), xj[1], acc);
                acc = fma(static_cast<double>(a.y), xj[2], acc);
                acc = fma(static_cast<double>(c.y), xj[3], acc);
                acc = fma(static_cast<double>(a.z), xj[4], acc);
                acc = fma(static_cast<double>(c.z), xj[5], acc);
                acc = fma(static_cast<double>(a.w), xj[6], acc);
                acc = fma(static_cast<double>(c.w), xj[7], acc);
            }
        }
    } else {
        const float *Up = tri + static_cast<int64_t>(nb) * nb * 1024;
        for (int J = I + p; J < nb; J += 8) {
            const float *blk = Up + (static_cast<int64_t>(J) * nb + I) * 1024 + 32 * 16 * (rho & 1) + (rho >> 1);
            const double *xj = xs + 32 * J;
#pragma unroll
            for (int g = 0; g < 32; ++g) acc = fma(static_cast<double>(blk[g * 16]), xj[g], acc);
        }
    }
    part[p][rho] = acc;
    __syncthreads();
    if (p == 0) {
        double y = part[0][rho];
#pragma unroll
        for (int q = 1; q < 8; ++q) y += part[q][rho];
        xout[32 * I + rho] = y;
    }
}

bool tri_inverse_ok(int b) { return tri_ok(b); }

int64_t tri_inverse_offset(int b) { return ceil_div(b, kStrip) * kInvBlock; }

int launch_tri_trsv(bool upper, int b, const float *tri, const double *xin, double *xout, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(b / kStrip));
    if (upper)
        hipLaunchKernelGGL(tri_trsv_kernel<true>, grid, dim3(256), 0, s, b, tri, xin, xout);
    else
        hipLaunchKernelGGL(tri_trsv_kernel<false>, grid, dim3(256), 0, s, b, tri, xin, xout);
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

static bool aligned16(const float *p, int64_t ld)
{
    return (ld % 4) == 0 && (reinterpret_cast<uintptr_t>(p) % 16) == 0;
}

// Strip width of the triangular-product TRSMs: 32 (two workgroups per CU,
// 10% faster at b = 512, m = 188416) where the strip loaders allow it.
static int tri_strip_width(int b) { return b % 64 == 0 ? 32 : 64; }

template <int S>
static void tri_l_launch(int64_t m, int b, const float *tri, float *L, int64_t ldl, float *amax, int tb,
                         hipStream_t s)
{
    const dim3 g(static_cast<unsigned>(ceil_div(m, S)));
    if (aligned16(L, ldl))
        hipLaunchKernelGGL((trsm_l_tri_kernel<S, true>), g, dim3(kTriThreads), 0, s, m, b, tri, L, ldl, amax, tb);
    else
        hipLaunchKernelGGL((trsm_l_tri_kernel<S, false>), g, dim3(kTriThreads), 0, s, m, b, tri, L, ldl, amax, tb);
}

template <int S>
static void tri_u_launch(int b, int64_t n, const float *tri, float *U, int64_t ldu, float *amax, int tb,
                         hipStream_t s)
{
    const dim3 g(static_cast<unsigned>(ceil_div(n, S)));
    if (aligned16(U, ldu))
        hipLaunchKernelGGL((trsm_u_tri_kernel<S, true>), g, dim3(kTriThreads), 0, s, b, n, tri, U, ldu, amax, tb);
    else
        hipLaunchKernelGGL((trsm_u_tri_kernel<S, false>), g, dim3(kTriThreads), 0, s, b, n, tri, U, ldu, amax, tb);
}

int launch_trsm_l_strip(int64_t m, int b, const float *U, int64_t ldu, const float *inv, float *L,
                        int64_t ldl, float *amax, int tb, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(ceil_div(m, kStrip)));
    if (tri_ok(b)) {
        const float *tri = inv + ceil_div(b, kStrip) * kInvBlock;
        if (tri_strip_width(b) == 32)
            tri_l_launch<32>(m, b, tri, L, ldl, amax, tb, s);
        else
            tri_l_launch<64>(m, b, tri, L, ldl, amax, tb, s);
    } else {
        hipLaunchKernelGGL(trsm_l_strip_kernel, grid, dim3(256), 0, s, m, b, U, ldu, inv, L, ldl, amax, tb);
    }
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

int launch_trsm_u_strip(int b, int64_t n, const float *Lkk, int64_t ldl, const float *inv, float *U,
                        int64_t ldu, float *amax, int tb, hipStream_t s)
{
    const dim3 grid(static_cast<unsigned>(ceil_div(n, kStrip)));
    if (tri_ok(b)) {
        const float *tri = inv + ceil_div(b, kStrip) * kInvBlock;
        if (tri_strip_width(b) == 32)
            tri_u_launch<32>(b, n, tri, U, ldu, amax, tb, s);
        else
            tri_u_launch<64>(b, n, tri, U, ldu, amax, tb, s);
    } else {
        hipLaunchKernelGGL(trsm_u_strip_kernel, grid, dim3(256), 0, s, b, n, Lkk, ldl, inv, U, ldu, amax, tb);
    }
    HPLAI_HIP(hipGetLastError());
    return HPLAI_OK;
}

}  // namespace hplai
