"""GPU tests of the dense factor / multi-RHS solves and the autograd binding.

Tolerances:
  multi-RHS solves (plain and transposed) : relative 1e-12 against FP64
                                            substitution with the same FP32 factors
  hplai_factor_f32                        : bit-identical to the factors hplai_solve
                                            leaves (same kernels, same schedule)
  gradient vs FP64 torch.linalg.solve     : relative 1e-9 (both are FP64-accurate)
  gradient vs central differences (FP64)  : relative 1e-6 on 32 random entries
                                            (h = 1e-2; the loss is computed to ~1e-14)
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = 42


def _gen_A64(ctx, n):
    """The seeded diagonally dominant HPL-AI matrix as a dense FP64 tensor
    (off-diagonals are the FP32 LCG values, the diagonal is the FP64 row sum)."""
    import torch
    d = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.gen_diag(n, SEED, 0, n, d)
    A32 = torch.empty((n, n), dtype=torch.float32, device="cuda")   # column-major: A32[j, i] = a_ij
    ctx.gen_matrix(n, SEED, 0, 0, n, n, None, A32, n)
    A = A32.t().double().contiguous()
    A.diagonal().copy_(d)
    return A


def _np_factors(F, n):
    """L (unit lower) and U from the column-major FP32 factor tensor F[j, i]."""
    M = F.cpu().numpy().T.astype(np.float64)
    return np.tril(M, -1) + np.eye(n), np.triu(M)


@pytest.mark.parametrize("n,b,k", [(1200, 60, 3), (2048, 512, 5), (3000, 1000, 2)])
def test_lu_solve_plain_and_transposed(ctx, n, b, k):
    import torch
    A = _gen_A64(ctx, n)
    F = A.t().contiguous().float()
    ctx.factor(n, b, F, n, fmt="f16")
    L, U = _np_factors(F, n)
    rng = np.random.default_rng(1)
    Bn = rng.standard_normal((n, k))
    for trans in (False, True):
        Xc = torch.from_numpy(np.ascontiguousarray(Bn.T)).cuda()     # column j at Xc[j]
        ctx.lu_solve(n, k, b, F, n, Xc, n, trans=trans)
        torch.cuda.synchronize()
        if trans:   # (LU)^-T = L^-T U^-T
            ref = np.linalg.solve(L.T, np.linalg.solve(U.T, Bn))
        else:
            ref = np.linalg.solve(U, np.linalg.solve(L, Bn))
        got = Xc.cpu().numpy().T
        assert np.max(np.abs(got - ref)) / np.max(np.abs(ref)) < 1e-12, trans


@pytest.mark.parametrize("fmt,n,b", [("f16", 1200, 60), ("e4m3", 4096, 512)])
def test_factor_equals_solve_factors(ctx, fmt, n, b):
    import torch
    res = ctx.solve(n, b, fmt=fmt, keep_factors=True)
    assert res["converged"]
    lu_solve = ctx.solution_factors(n)           # lu[j, i] layout
    ctx.release()
    A = _gen_A64(ctx, n)
    F = A.t().contiguous().float()
    ctx.factor(n, b, F, n, fmt=fmt)
    torch.cuda.synchronize()
    assert np.array_equal(F.cpu().numpy().view(np.uint32), np.ascontiguousarray(lu_solve).view(np.uint32))


@pytest.mark.parametrize("m,n,k,ta,tb,beta", [
    (100, 70, 33, False, False, 0.0), (257, 129, 1000, True, True, 0.5),
    (1000, 1000, 64, True, False, 0.0),          # the gradient's shape (rank-64 product)
    (64, 8192, 8192, False, False, 1.0),         # the residuals: K split over workgroups
    (64, 2048, 2048, False, True, 1.0), (300, 200, 0, False, False, 2.0)])
def test_gemm_f64_vs_fp64(ctx, m, n, k, ta, tb, beta):
    """hplai_gemm_f64 against an FP64 product on the CPU: |C - ref| <=
    (k + 2) 2^-52 (|alpha| |op(A)| |op(B)| + |beta C0|), the FP64 summation
    bound (the library adds in a different order); a second call gives the same
    bits (a split K is summed in a fixed order)."""
    import torch
    rng = np.random.default_rng(m + n + k)
    alpha = -1.0
    Ash = (k, m) if ta else (m, k)
    Bsh = (n, k) if tb else (k, n)
    A = rng.standard_normal(Ash)
    B = rng.standard_normal(Bsh)
    C0 = rng.standard_normal((m, n))
    opA = A.T if ta else A
    opB = B.T if tb else B
    ref = alpha * (opA @ opB) + beta * C0
    bound = (k + 2) * 2.0 ** -52 * (abs(alpha) * (np.abs(opA) @ np.abs(opB)) + abs(beta) * np.abs(C0)) + 1e-300
    # column-major device copies: a column-major r x c matrix is the row-major (c, r) tensor of its transpose
    dA = torch.from_numpy(np.ascontiguousarray(A.T)).cuda()
    dB = torch.from_numpy(np.ascontiguousarray(B.T)).cuda()
    outs = []
    for _ in range(2):
        dC = torch.from_numpy(np.ascontiguousarray(C0.T)).cuda()
        ctx.gemm_f64(ta, tb, m, n, k, alpha, dA, Ash[0], dB, Bsh[0], beta, dC, m)
        torch.cuda.synchronize()
        outs.append(dC.cpu().numpy().T)
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))
    ratio = np.max(np.abs(outs[0] - ref) / bound)
    assert ratio < 1.0, ratio


def test_factor_zero_pivot(ctx):
    import hplai
    import torch
    n = 256
    F = torch.eye(n, dtype=torch.float32, device="cuda")
    F[7, 7] = 0.0
    with pytest.raises(hplai.HplaiError) as e:
        ctx.factor(n, 128, F, n, fmt="f16")
    assert e.value.code == hplai.EPIVOT and "column 7" in str(e.value)


def test_autograd_solve_n8192(ctx):
    """x = solve(A, B), n = 8192, B 8192 x 64, loss sum(W * X)."""
    import torch
    from hplai import autograd
    n, k = 8192, 64
    A = _gen_A64(ctx, n)
    g = torch.Generator(device="cpu").manual_seed(2024)
    B = torch.randn((n, k), generator=g, dtype=torch.float64).cuda()
    W = torch.randn((n, k), generator=g, dtype=torch.float64).cuda()
    Ar = A.clone().requires_grad_(True)
    Br = B.clone().requires_grad_(True)
    info = {}
    X = autograd.solve(Ar, Br, ctx=ctx, b=512, fmt="f16", tol=1.0, min_steps=1, info=info)
    loss = (W * X).sum()
    loss.backward()
    assert info["forward_harness"] < 1.0 and info["backward_harness"] < 1.0
    print("IR steps fwd %d bwd %d, harness fwd %.3g bwd %.3g" % (
        info["forward_steps"], info["backward_steps"], info["forward_harness"], info["backward_harness"]))

    # FP64 reference gradient through torch.linalg.solve
    At = A.clone().requires_grad_(True)
    Bt = B.clone().requires_grad_(True)
    Xt = torch.linalg.solve(At, Bt)
    (W * Xt).sum().backward()
    rel_x = float((X.detach() - Xt.detach()).abs().max() / Xt.detach().abs().max())
    rel_a = float((Ar.grad - At.grad).abs().max() / At.grad.abs().max())
    rel_b = float((Br.grad - Bt.grad).abs().max() / Bt.grad.abs().max())
    print("vs torch.linalg.solve: x %.3g dA %.3g dB %.3g" % (rel_x, rel_a, rel_b))
    assert rel_x < 1e-9 and rel_a < 1e-9 and rel_b < 1e-9

    # central differences on 32 random entries of A
    rng = np.random.default_rng(7)
    idx = [(int(i), int(j)) for i, j in zip(rng.integers(0, n, 32), rng.integers(0, n, 32))]
    idx[0] = (5, 5)   # one diagonal entry
    h = 1e-2

    def loss_at(Ap):
        Xp = autograd.solve(Ap, B, ctx=ctx, b=512, fmt="f16", tol=1.0, min_steps=1)
        return float((W * Xp).sum())

    fd, an = [], []
    for (i, j) in idx:
        Ap = A.clone()
        Ap[i, j] += h
        lp = loss_at(Ap)
        Ap[i, j] -= 2 * h
        lm = loss_at(Ap)
        fd.append((lp - lm) / (2 * h))
        an.append(float(Ar.grad[i, j]))
    fd, an = np.array(fd), np.array(an)
    rel_fd = np.max(np.abs(fd - an)) / np.max(np.abs(an))
    print("vs central differences (32 entries): %.3g" % rel_fd)
    assert rel_fd < 1e-6
