"""GPU tests of Higham's HPL-AI matrix (--highammat), a second input family.

Pins:
  generator   : blocks against the FP64 product L U (closed form), |error| <=
                2 FP32 ulps of the entry (the kernel rounds one FMA in FP32)
  residual    : y + s A x against numpy's FP64 product, relative 1e-13
  LU          : the unpivoted LU of L U is L and U, so the factors of a solve
                must be a below and b above the unit diagonal: relative 1e-4 in
                the first block column and row (FP32 getrf and TRSMs only), and
                over the whole matrix relative 1e-2 for the reference's signs,
                5e-2 for the exact-kappa signs, whose rounding errors double
                from one block column to the next (measured 4e-4, 8e-4, ...,
                2.9e-2 at n = 4096, b = 512: the growth that ||A^-1|| =
                (1 + alpha)^(n-1) allows); a sign or index mistake is O(1)
  solve       : harness < 16 for every format, both sign conventions
"""
import os
import re
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def exact_block(n, a, b, i0, j0, m, k):
    i = np.arange(i0, i0 + m, dtype=np.float64)[:, None]
    j = np.arange(j0, j0 + k, dtype=np.float64)[None, :]
    return np.where(i < j, b + a * b * i, np.where(i == j, 1 + a * b * i, a + a * b * j))


@pytest.mark.parametrize("sign", [1.0, -1.0])
@pytest.mark.parametrize("blk", [(0, 0, 300, 300), (1000, 2000, 500, 400), (2000, 100, 700, 300)])
def test_generator_blocks(ctx, sign, blk):
    import hplai
    import torch
    n = 3000
    alpha, beta = hplai.higham_coeffs(n)
    a, b = sign * alpha, sign * beta
    i0, j0, m, k = blk
    A = torch.full((k, m), float("nan"), dtype=torch.float32, device="cuda")   # column-major
    ctx.gen_higham(n, a, b, i0, j0, m, k, A, m)
    torch.cuda.synchronize()
    got = A.cpu().numpy().T.astype(np.float64)
    want = exact_block(n, a, b, i0, j0, m, k)
    ulp = np.spacing(np.abs(want).astype(np.float32)).astype(np.float64)
    assert np.all(np.abs(got - want) <= 2 * ulp), np.max(np.abs(got - want) / ulp)


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_residual_gemv(ctx, sign):
    import hplai
    import torch
    n = 3000
    alpha, beta = hplai.higham_coeffs(n)
    a, b = sign * alpha, sign * beta
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    A = exact_block(n, a, b, 0, 0, n, n)
    want = y - 0.75 * (A @ x)
    xd, yd = torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()
    ctx.higham_gemv(n, a, b, -0.75, xd, yd)
    torch.cuda.synchronize()
    got = yd.cpu().numpy()
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < 1e-13


@pytest.mark.parametrize("matrix,sign", [("higham", 1.0), ("higham_kappa", -1.0)])
def test_lu_factors_are_the_closed_form(ctx, matrix, sign):
    import hplai
    n, b = 4096, 512
    alpha, beta = hplai.higham_coeffs(n)
    res = ctx.solve(n, b, fmt="f16", matrix=matrix, keep_factors=True)
    lu = ctx.solution_factors(n).T.astype(np.float64)   # lu[i, j]
    ctx.release()
    assert res["converged"] and res["harness"] < 16.0, res
    tol = 1e-2 if sign > 0 else 5e-2
    low = np.tril(lu, -1)
    up = np.triu(lu, 1)
    mask_l = np.tril(np.ones((n, n), dtype=bool), -1)
    mask_u = np.triu(np.ones((n, n), dtype=bool), 1)
    err_l = np.where(mask_l, np.abs(low - sign * alpha), 0.0) / alpha
    err_u = np.where(mask_u, np.abs(up - sign * beta), 0.0) / beta
    assert err_l[:, :b].max() < 1e-4 and err_u[:b, :].max() < 1e-4, (err_l[:, :b].max(), err_u[:b, :].max())
    assert err_l.max() < tol and err_u.max() < tol, (err_l.max(), err_u.max())
    assert np.max(np.abs(np.diag(lu) - 1.0)) < 1e-3


@pytest.mark.parametrize("matrix", ["higham", "higham_kappa"])
@pytest.mark.parametrize("fmt,outer", [("f16", 1), ("f16t", 4), ("e4m3", 4)])
def test_solve_converges(ctx, matrix, fmt, outer):
    res = ctx.solve(8192, 512, fmt=fmt, outer=outer, matrix=matrix)
    print(matrix, fmt, "IR steps", res["ir_steps"], "harness", res["harness"])
    assert res["converged"] and res["harness"] < 16.0, (matrix, fmt, res["trace_harness"])


def test_driver_highammat():
    exe = os.path.join(ROOT, "hpl-ai_amd", "bin", "driver.out")
    out = subprocess.run([exe, "1200", "60", "1", "--highammat", "--no-timerdump"], capture_output=True,
                         text=True, timeout=120).stdout
    assert " hmat" in out, out
    h = [float(v) for v in re.findall(r"hpl-harness=([0-9.eE+-]+)", out)]
    assert h and h[-1] < 16.0, out


def test_solve_lazy_init(ctx):
    """-d: blocks get their Higham entries added when they join a panel."""
    res = ctx.solve(4096, 512, fmt="f16", outer=2, matrix="higham_kappa", lazy=True)
    assert res["converged"] and res["harness"] < 16.0, res["trace_harness"]
