"""Write the golden fixtures in tests/golden/ from the reference's own code.

Runs only where the reference tree exists (it needs oracle/_ref/libhplai_ref.so,
built by `make -C oracle ref` from the reference's hpl_rand.hpp, fp16sim.hpp,
kernels/avx2.cpp and otf_gemv.cpp compiled where they lie).  The outputs are
data (inputs and the reference's outputs) in .npz files, loaded with
allow_pickle=False.

    python tests/golden/make_golden.py
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import oracle as O  # noqa: E402

SEED = 42


def p(a):
    return ctypes.c_void_p(a.ctypes.data)


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libhplai_ref.so missing: run `make -C oracle ref`")

    # ---- LCG coefficients (hpl_rand.hpp:20-32) ----
    pows = np.array([0, 1, 2, 3, 60, 1200, 4096, 61440, 196608, 196608 * 196608, 2 ** 40 + 12345,
                     2 ** 63 + 7], dtype=np.uint64)
    pa = np.zeros_like(pows)
    pc = np.zeros_like(pows)
    for i, nn in enumerate(pows):
        a, c = ctypes.c_uint64(), ctypes.c_uint64()
        R.ref_coeff_pow(int(nn), ctypes.byref(a), ctypes.byref(c))
        pa[i], pc[i] = a.value, c.value

    # ---- generator blocks (fill_one_panel_with_rand<float>, calc_diag=true) ----
    blocks = [(1200, 0, 0, 60, 60), (1200, 60, 0, 60, 60), (1200, 540, 600, 60, 60),
              (1200, 1140, 1140, 60, 60), (4096, 512, 1024, 128, 96),
              (61440, 61000, 60900, 64, 40), (196608, 196544, 196480, 64, 32)]
    blk_data = []
    for (n, i0, j0, m, k) in blocks:
        A = np.zeros((k, m), dtype=np.float32)
        diag = 1 if (i0 == j0 and m == k) else 0
        R.ref_fill_panel_f32(n, i0, j0, m, k, p(A), m, SEED, diag)
        blk_data.append(A)

    # ---- diagonal and rhs (calc_diag, pcolvgen) ----
    n = 1200
    diag1200 = np.array([R.ref_calc_diag(i, n, SEED) for i in range(n)], dtype=np.float64)
    rhs1200 = np.zeros(n, dtype=np.float64)
    for blk in range(n // 60):
        R.ref_fill_panel_f64(n, 60 * blk, n, 60, 1, p(rhs1200[60 * blk:]), 1, SEED, 0)
    diag_rows_61440 = np.array([0, 1, 30000, 61439], dtype=np.int64)
    diag61440 = np.array([R.ref_calc_diag(int(i), 61440, SEED) for i in diag_rows_61440])

    # ---- FP16 conversion (F16C RNE via struct fp16, fp16sim.hpp:103-132) ----
    rng = np.random.default_rng(2024)
    f16_in = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(1e-3),
        (rng.standard_normal(20000) * 1e4).astype(np.float32),
        np.array([0.0, -0.0, 65504.0, 65519.99, 65520.0, 6.1035156e-05, 5.9604645e-08,
                  2.9802322e-08, 2.9802326e-08, 1e-10, np.inf, -np.inf], dtype=np.float32)])
    f16_out = np.array([R.ref_f32_to_f16(float(v)) for v in f16_in], dtype=np.uint16)

    # ---- hgemm_opt / shgemm_opt (kernels/avx2.cpp) known answers ----
    gm, gn, gk = 53, 40, 150      # k > 100 exercises the k-blocking; m % 16 != 0 the remainder
    ga = (rng.standard_normal((gk, gm)) * 0.05).astype(np.float16)   # ga[k, i], lda = gm
    gb = (rng.standard_normal((gn, gk)) * 0.05).astype(np.float16)   # gb[j, k], ldb = gk
    gc16 = (rng.standard_normal((gn, gm)) * 0.5).astype(np.float16)
    gc32 = (rng.standard_normal((gn, gm)) * 0.5).astype(np.float32)
    hout = gc16.copy()
    R.ref_hgemm(gm, gn, gk, -1.0, p(ga), gm, p(gb), gk, 1.0, p(hout), gm)
    sout = gc32.copy()
    R.ref_shgemm(gm, gn, gk, -1.0, p(ga), gm, p(gb), gk, 1.0, p(sout), gm)

    # ---- otf_gemv_kernel (otf_gemv.cpp:5-131) on an off-diagonal 60x60 block ----
    n = 1200
    ox = rng.standard_normal(60)
    oy = rng.standard_normal(60)
    oyout = oy.copy()
    state = O.lib().orc_lcg_state(SEED, 0 + n * 120)   # block (rows 0..59, cols 120..179)
    R.ref_otf_gemv_kernel(n, 60, 60, -1.0, p(ox), p(oyout), state)

    np.savez(os.path.join(HERE, "ref_kernels.npz"),
             pow_n=pows, pow_a=pa, pow_c=pc,
             blocks=np.array(blocks, dtype=np.int64),
             **{"block%d" % i: b for i, b in enumerate(blk_data)},
             diag1200=diag1200, rhs1200=rhs1200, diag_rows_61440=diag_rows_61440,
             diag61440=diag61440, f16_in=f16_in, f16_out=f16_out,
             gemm_shape=np.array([gm, gn, gk]), gemm_a=ga.view(np.uint16), gemm_b=gb.view(np.uint16),
             gemm_c16=gc16.view(np.uint16), gemm_c32=gc32, hgemm_out=hout.view(np.uint16),
             shgemm_out=sout, otf_x=ox, otf_y=oy, otf_out=oyout,
             otf_block=np.array([n, 0, 120, 60, 60], dtype=np.int64))
    print("wrote", os.path.join(HERE, "ref_kernels.npz"))


if __name__ == "__main__":
    main()
