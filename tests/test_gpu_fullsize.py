"""Full-size solves on one MI355X, checked through size-independent properties.

C2: n = 61440, b = 512, plain FP16 Schur update (the reference's precision).
C3: n = 196608, b = 512, e4m3 with per-512x512-tile power-of-two exponents,
    outer 16, look-ahead 1 (the bench configuration; 155 GB of FP32 factors).

The oracle's LU does not run at these sizes (the reference's CPU path would
take hours to days), so the solve is checked by what does not need it:
  - the refinement reaches harness < 16 (iterative_refinement.hpp:43-48) in
    at most the bench's number of residual evaluations;
  - sampled rows of the FP64 diagonal (calc_diag, hpl_rand.hpp:108-123) and
    the whole right-hand side are bit-exact against the oracle;
  - the step-0 residual b - A x0, x0 = b ./ d (iterative_refinement.hpp:33-41),
    computed through hplai_otf_gemv_f64 equals the solve's first trace entry
    and the oracle's full FP64 residual (orc_otf_gemv, A regenerated on the
    CPU) within the FP64 summation-order bound 8 sqrt(n) 2^-53 sum_j |a_ij x_j|;
  - sampled blocks of the kept factors reproduce the matrix from the
    quantized operands the Schur updates consumed (parity.block_residual):
    first blocks, blocks on both sides of an outer-block boundary, the
    middle, the last blocks (K up to n - b), each checked against
    |R| <= 2^-7 sigma + 2^-20 sqrt(m + 2) (|A| + S).  Every block spans two
    256-row packed operand panels and 8 64-byte K stages per slice.
"""
import numpy as np
import pytest

import parity as P

pytestmark = pytest.mark.gpu
SEED = 42

CONFIGS = [
    pytest.param("C2", 61440, 512, "f16", 16, 3, id="C2-n61440-f16"),
    pytest.param("C3", 196608, 512, "e4m3", 16, 3, id="C3-n196608-e4m3"),
]


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name,n,b,fmt,outer,max_steps", CONFIGS)
def test_fullsize_solve(ctx, oracle, name, n, b, fmt, outer, max_steps):
    import torch
    print("%s: n=%d b=%d %s outer %d" % (name, n, b, fmt, outer), flush=True)
    res = ctx.solve(n, b, fmt=fmt, outer=outer, lookahead=1, keep_factors=True)
    print("  %.3f s LU+IR, %d residuals, harness %.4g, trace %s" %
          (res["time_total"], res["ir_steps"], res["harness"], res["trace_harness"]), flush=True)
    try:
        assert res["converged"] and res["harness"] < 16.0, res["trace_harness"]
        assert res["ir_steps"] <= max_steps

        # generator: sampled diagonal rows and the whole right-hand side
        d = torch.empty(n, dtype=torch.float64, device="cuda")
        rhs = torch.empty(n, dtype=torch.float64, device="cuda")
        ctx.gen_diag(n, SEED, 0, n, d)
        ctx.gen_rhs(n, SEED, 0, n, rhs)
        torch.cuda.synchronize()
        dh, bh = d.cpu().numpy(), rhs.cpu().numpy()
        for i0 in (0, n // 2 - 32, n - 64):
            want = oracle.gen_diag(n, SEED, i0, 64)
            assert np.array_equal(dh[i0:i0 + 64].view(np.uint64), want.view(np.uint64)), i0
        assert np.array_equal(bh.view(np.uint64), oracle.gen_rhs(n, SEED).view(np.uint64))

        # step-0 residual through the C ABI, the solve's trace, the oracle
        x0 = bh / dh
        r = rhs.clone()
        ctx.otf_gemv(n, SEED, d, -1.0, torch.from_numpy(x0).cuda(), r)
        torch.cuda.synchronize()
        rg = r.cpu().numpy()
        r0 = float(np.max(np.abs(rg)))
        assert abs(r0 - res["trace_resid"][0]) <= 1e-14 * r0, (r0, res["trace_resid"][0])
        rr = oracle.otf_gemv(n, dh, -1.0, x0, bh)
        # both are FP64 sums of n terms in different orders: |dr_i| within
        # 8 sqrt(n) 2^-53 sum_j |a_ij x_j|, and sum_j |a_ij x_j| <= 2 d_i max|x0|
        # (d_i is the row's off-diagonal absolute sum).  r is ~100x smaller
        # than the terms it cancels, so a relative 1e-13 does not hold here.
        scale = 8.0 * np.sqrt(n) * 2.0 ** -53 * 2.0 * dh * np.max(np.abs(x0))
        ratio = float(np.max(np.abs(rg - rr) / scale))
        print("  step-0 residual %.16e, vs oracle: %.3g relative, %.3g of the bound" %
              (r0, np.max(np.abs(rg - rr)) / np.max(np.abs(rr)), ratio), flush=True)
        assert ratio < 1.0

        # sampled-block factor residual
        nb = n // b
        s = outer
        blocks = [(0, 0), (1, 0), (0, 1), (s, s - 1), (s - 1, s), (s, s), (s + 1, s), (s, s + 1),
                  (nb // 2 + 1, nb // 2), (nb // 2 - 2, nb // 2 + 3), (nb - 1, nb - 1), (nb - 1, nb - 2),
                  (nb - 2, nb - 1)]
        worst = 0.0
        for I, J in blocks:
            a = oracle.gen_block(n, I * b, J * b, b, b, SEED)
            R, S, sig, absA = P.block_residual(ctx.solution_block, a, I, J, b, fmt)
            ratio, rs = P.residual_ratio(R, S, sig, absA, min(I, J), I == J)
            print("  block (%d, %d): |R| / bound %.3g, |R| / sigma %.3g" % (I, J, ratio, rs), flush=True)
            worst = max(worst, ratio)
        assert worst < 1.0, worst
    finally:
        ctx.release()
