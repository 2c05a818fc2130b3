// device_common.hpp — device helpers shared by the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hplai {

// ---------------------------------------------------------------------------
// LCG of hpl_rand.hpp:8-57: x <- a x + c mod 2^64.  A coefficient pair (a, c)
// is an affine map; composition l*r = (l.a r.a, l.a r.c + l.c) applies r first.
// ---------------------------------------------------------------------------
constexpr uint64_t kLcgA = 6364136223846793005ULL;
constexpr uint64_t kLcgC = 1ULL;

struct LcgCoeff {
    uint64_t a, c;
};

__host__ __device__ constexpr LcgCoeff lcg_mul(LcgCoeff l, LcgCoeff r)
{
    return LcgCoeff{l.a * r.a, l.a * r.c + l.c};
}

// p[k] = (a, c)^(2^k): the squares used by jump-ahead.
struct LcgTable {
    LcgCoeff p[64];
};

constexpr LcgTable make_lcg_table()
{
    LcgTable t{};
    LcgCoeff c{kLcgA, kLcgC};
    for (int k = 0; k < 64; ++k) {
        t.p[k] = c;
        c = lcg_mul(c, c);
    }
    return t;
}

__host__ inline LcgCoeff lcg_pow_host(uint64_t n)
{
    constexpr LcgTable t = make_lcg_table();
    LcgCoeff r{1, 0};
    for (int k = 0; n; ++k, n >>= 1)
        if (n & 1) r = lcg_mul(r, t.p[k]);
    return r;
}

// Per-translation-unit copy of the jump table in constant memory.
static __constant__ LcgTable kLcg = make_lcg_table();

// State after `idx` further steps from x: compose the table squares for the
// set bits of idx (all factors are powers of one map, so order is free).
__device__ __forceinline__ uint64_t lcg_jump(uint64_t x, uint64_t idx)
{
    for (int k = 0; idx; ++k, idx >>= 1)
        if (idx & 1) x = kLcg.p[k].a * x + kLcg.p[k].c;
    return x;
}

// Initial state x1 = a*seed + c (RandStat::initialize, hpl_rand.hpp:38-43).
__host__ __device__ inline uint64_t lcg_init(uint64_t seed) { return kLcgA * seed + kLcgC; }

// Value map (hpl_rand.hpp:50-52): (int64)x * 0x1.fffffffffffffP-65 in [-0.5, 0.5).
// int64 -> double is correctly rounded (one rounding in the final add of the
// hi/lo split) and the product is an IEEE multiply.
__device__ __forceinline__ double lcg_value(uint64_t s)
{
    return __dmul_rn(static_cast<double>(static_cast<int64_t>(s)), 0x1.fffffffffffffP-65);
}

// ---------------------------------------------------------------------------
// Low-precision helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float f16_bits_to_f32(uint32_t h)
{
    return static_cast<float>(__builtin_bit_cast(_Float16, static_cast<uint16_t>(h)));
}

__device__ __forceinline__ uint16_t f32_to_f16_bits(float x)
{
    _Float16 h = static_cast<_Float16>(x);   // v_cvt_f16_f32, round to nearest even
    return __builtin_bit_cast(uint16_t, h);
}

// Pack 4 floats to 4 OCP e4m3 bytes (v_cvt_pk_fp8_f32, round to nearest even).
__device__ __forceinline__ uint32_t f32x4_to_e4m3x4(float a, float b, float c, float d)
{
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(a, b, 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(c, d, w, true);
    return static_cast<uint32_t>(w);
}

// Largest e with amax * 2^e <= fmt_max, fmt_max = (1 + mr/2^23) * 2^er; exact
// integer arithmetic on the FP32 bits.  Returns 0 for amax == 0.
__host__ __device__ inline int tile_exponent(float amax, uint32_t mr_frac, int er)
{
    if (!(amax > 0.0f)) return 0;
    uint32_t u = __builtin_bit_cast(uint32_t, amax);
    int E;
    uint32_t frac;
    if (u >= 0x00800000u) {
        E = static_cast<int>(u >> 23) - 127;
        frac = u & 0x7fffffu;
    } else {
        int lz = __builtin_clz(u) - 8;
        E = -126 - lz;
        frac = (u << lz) & 0x7fffffu;
    }
    return er - E - (frac > mr_frac ? 1 : 0);
}
constexpr uint32_t kE4m3MrFrac = 0x600000u;  // 448 = 1.75 * 2^8
constexpr int kE4m3Er = 8;
constexpr uint32_t kF16MrFrac = 0x7fe000u;   // 65504 = (2 - 2^-10) * 2^15
constexpr int kF16Er = 15;

// e4m3 exponents become the E8M0 block scales 127 - e of the scaled MFMA, so
// they stay in [-127, 127]: a tile whose largest magnitude is below about
// 2^-119 (e.g. subnormal) takes e = 127 and quantizes to (near) zero instead
// of producing the NaN scale byte 0xFF.
__host__ __device__ inline int clamp_e4m3_exponent(int e) { return e > 127 ? 127 : (e < -127 ? -127 : e); }
__host__ __device__ inline int e4m3_exponent(float amax)
{
    return clamp_e4m3_exponent(tile_exponent(amax, kE4m3MrFrac, kE4m3Er));
}
// E8M0 scale byte of exponent e (caller-supplied exponents are clamped too)
__host__ __device__ inline int e8m0_scale(int e) { return 127 - clamp_e4m3_exponent(e); }

__device__ __forceinline__ float ldexp_fast(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }

// Panel-packed layout of the solver's quantized operand panels: rows in
// panels of 256, each panel's K bytes in stages of 64 bytes, and the
// 256 x 64-byte block of one (panel, stage) contiguous (16 KB, row r at
// r * 64), so one wave-wide 1 KB load covers 16 whole rows of a stage.  A
// panel occupies the bytes of 256 rows of the row layout with the same row
// length ld_b: a row offset that is a multiple of 256 moves the base by
// row * ld_b in both layouts, a K offset that is a multiple of 64 by
// 256 * kbyte here.
__host__ __device__ __forceinline__ int64_t packed_off(int64_t row, int64_t kbyte, int64_t ld_b)
{
    return (row >> 8) * (ld_b << 8) + (kbyte >> 6) * 16384 + (row & 255) * 64 + (kbyte & 63);
}

// atomic max of a non-negative float, ordered as its uint32 bit pattern
__device__ __forceinline__ void atomic_max_pos(float *p, float v)
{
    atomicMax(reinterpret_cast<unsigned int *>(p), __float_as_uint(v));
}

// Max of v over the lanes of the wave that share `tile`; one atomic per tile.
// Every lane of the wave must call it (inactive lanes pass active = false).
__device__ __forceinline__ void tile_max_reduce(float *amax, int64_t tile, float v, bool active)
{
    if (amax == nullptr) return;
    const int64_t t0 = __shfl(tile, 0);
    const bool uniform = __all(!active || tile == t0);
    if (uniform) {
        float m = active ? v : 0.0f;
        for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
        if (__lane_id() == 0 && __any(active)) atomic_max_pos(amax + t0, m);
    } else if (active) {
        atomic_max_pos(amax + tile, v);
    }
}

}  // namespace hplai

namespace hplai {

// In-place inverse of a lower-triangular 32 x 32 matrix held in LDS
// (M[r * 33 + c], non-zero diagonal) by recursive doubling:
//   [A 0; C B]^-1 = [A^-1 0; -B^-1 C A^-1, B^-1]
// with block sizes 1, 2, 4, 8, 16.  T is a 32 x 33 LDS scratch.  Every thread
// of the workgroup must call it (it contains barriers).
__device__ inline void lower_inv32(float *M, float *T, int tid, int nthr)
{
    for (int i = tid; i < 32; i += nthr) M[i * 33 + i] = 1.0f / M[i * 33 + i];
    __syncthreads();
    for (int s = 1; s < 32; s *= 2) {
        const int np = 16 / s, per = s * s;
        for (int e = tid; e < np * per; e += nthr) {      // T = C A^-1
            const int p = e / per, ij = e % per, i = ij / s, j = ij % s, base = p * 2 * s;
            float acc = 0.0f;
            for (int k = j; k < s; ++k) acc = fmaf(M[(base + s + i) * 33 + base + k], M[(base + k) * 33 + base + j], acc);
            T[(base + s + i) * 33 + base + j] = acc;
        }
        __syncthreads();
        for (int e = tid; e < np * per; e += nthr) {      // C = -B^-1 T
            const int p = e / per, ij = e % per, i = ij / s, j = ij % s, base = p * 2 * s;
            float acc = 0.0f;
            for (int k = 0; k <= i; ++k) acc = fmaf(M[(base + s + i) * 33 + base + s + k], T[(base + s + k) * 33 + base + j], acc);
            M[(base + s + i) * 33 + base + j] = -acc;
        }
        __syncthreads();
    }
}

}  // namespace hplai
