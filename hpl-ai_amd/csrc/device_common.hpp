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

__device__ __forceinline__ float ldexp_fast(float x, int e) { return __builtin_amdgcn_ldexpf(x, e); }

}  // namespace hplai
