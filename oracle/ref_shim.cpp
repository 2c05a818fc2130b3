// ref_shim.cpp — C entry points over the reference's own MPI/BLAS-free sources.
//
// TEST INFRASTRUCTURE ONLY.  Compiled (by oracle/Makefile, target `ref`) with
// -I pointing at the reference tree, together with the reference's
// otf_gemv.cpp and kernels/avx2.cpp, into oracle/_ref/libhplai_ref.so.  It
// exists so the oracle restatement can be checked against the reference's
// actual code for the parts that build without MPI or BLAS:
//   hpl_rand.hpp        (LCG, fill_one_panel_with_rand, calc_diag)
//   fp16sim.hpp + kernels/avx2.cpp  (F16C fp16, hgemm_opt, shgemm_opt)
//   otf_gemv.cpp        (otf_gemv_kernel, SVE emulated by svesim.hpp)
// Nothing here restates an algorithm; it only forwards arguments.
#include <cstdint>
#include <cstring>
#include "hpl_rand.hpp"
#include "fp16sim.hpp"

extern "C" void otf_gemv_kernel(int64_t n, int mb, int nb, double alpha,
                                double const* __restrict__ x, double* __restrict__ y, uint64_t seed);

extern "C" {

void ref_fill_panel_f32(int n, int i0, int j0, int nrow, int ncol, float* a, long lda,
                        uint64_t seed, int calc_diag_flag)
{
    fill_one_panel_with_rand<float>(n, i0, j0, nrow, ncol, a, (size_t)lda, seed, calc_diag_flag != 0);
}

void ref_fill_panel_f64(int n, int i0, int j0, int nrow, int ncol, double* a, long lda,
                        uint64_t seed, int calc_diag_flag)
{
    fill_one_panel_with_rand<double>(n, i0, j0, nrow, ncol, a, (size_t)lda, seed, calc_diag_flag != 0);
}

double ref_calc_diag(int i, int n, uint64_t seed) { return calc_diag(i, n, seed); }

double ref_mat_elem(int n, int i, int j, int seed) { return mat_elem(n, i, j, seed); }

void ref_coeff_pow(uint64_t n, uint64_t* a, uint64_t* c)
{
    RandCoeff p = RandCoeff::default_vals().pow(n);
    *a = p.a;
    *c = p.c;
}

uint16_t ref_f32_to_f16(float x)
{
    fp16 h(x);
    return h.x;
}

// C(m x n, FP16) = C + alpha * A B via hgemm (beta must be 1 to reach hgemm_opt)
void ref_hgemm(int m, int n, int k, float alpha, const uint16_t* a, int lda,
               const uint16_t* b, int ldb, float beta, uint16_t* c, int ldc)
{
    hgemm(m, n, k, alpha, reinterpret_cast<const fp16*>(a), lda,
          reinterpret_cast<const fp16*>(b), ldb, beta, reinterpret_cast<fp16*>(c), ldc);
}

void ref_shgemm(int m, int n, int k, float alpha, const uint16_t* a, int lda,
                const uint16_t* b, int ldb, float beta, float* c, int ldc)
{
    shgemm(m, n, k, alpha, reinterpret_cast<const fp16*>(a), lda,
           reinterpret_cast<const fp16*>(b), ldb, beta, c, ldc);
}

void ref_otf_gemv_kernel(int64_t n, int mb, int nb, double alpha, const double* x,
                         double* y, uint64_t seed_state)
{
    otf_gemv_kernel(n, mb, nb, alpha, x, y, seed_state);
}

}  // extern "C"
