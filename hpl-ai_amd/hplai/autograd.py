"""torch.autograd binding of the mixed-precision dense solve X = A^-1 B.

Forward: the FP32 copy of A is factored in place by the library (FP32 panel
LU, Schur update in FP16 or e4m3 with FP32 accumulation:
`hplai_factor_f32`), then X is refined in FP64 against the FP64 A until every
column's HPL-AI harness ||r||_inf / ((||A||_inf ||x||_inf + ||b||_inf) n eps/2)
is below `tol` (iterative_refinement.hpp:18-57 applied column by column;
corrections by `hplai_lu_solve_f64`).

Backward (for a loss with dL/dX = W): the same low-precision factors solve
the transposed system A^T Lambda = W (`hplai_lu_solve_f64` with trans = 1,
i.e. U^-T then L^-T), refined in FP64 against A^T; then dA = -Lambda X^T (one
FP64 GEMM) and dB = Lambda.  The factorization is not repeated.

The factorization, the triangular solves and the FP64 products (the
residuals B - A X and W - A^T Lambda and the gradient -Lambda X^T, on the
FP64 matrix cores: `hplai_gemm_f64`) are the library's HIP kernels; torch
does the elementwise work around them (transposed copies, x += d, the norms
of the harness).
"""
import torch

from . import Context, FMTS

DBL_EPS = 2.220446049250313e-16


def harness(A_norm, R, X, B, n):
    """Per-column HPL-AI harness of the residual R = B - op(A) X (columns)."""
    return R.abs().amax(0) / (A_norm * X.abs().amax(0) + B.abs().amax(0)) / (n * DBL_EPS / 2)


def residual(hctx, A, X, B, trans=False):
    """R = B - op(A) X in FP64 with the library's GEMM (op = transpose when
    `trans`).  A (n, n), X and B (n, k): contiguous float64 CUDA tensors.

    A row-major (r, c) tensor is the column-major c x r matrix of its
    transpose, so R^T = B^T - X^T op(A)^T is one column-major GEMM with
    m = k, n = n, K = n: op(A)^T is A's own column-major view (A^T) for
    trans = False and its transpose for trans = True."""
    n, k = X.shape
    R = B.clone()
    hctx.gemm_f64(False, trans, k, n, n, -1.0, X, k, A, n, 1.0, R, k)
    return R


def outer_grad(hctx, Lam, X):
    """dA = -Lam X^T (n x n) with the library's GEMM: column-major
    dA^T = -X Lam^T, i.e. op(A) = X (the transpose of X's column-major view)
    and op(B) = Lam's column-major view, K = k."""
    n, k = X.shape
    dA = torch.empty((n, n), dtype=torch.float64, device=X.device)
    hctx.gemm_f64(True, False, n, n, k, -1.0, X, k, Lam, k, 0.0, dA, n)
    return dA


def refine(hctx, A, F, bs, B, trans=False, tol=16.0, maxit=50, min_steps=0):
    """Solve op(A) X = B (op = transpose when `trans`) with the FP32 factors F
    (column-major, leading dimension n) and FP64 refinement.

    A: (n, n) float64 CUDA tensor; B: (n, k) float64.  Returns (X, steps,
    harness per column)."""
    n, k = B.shape
    A = A.contiguous()
    B = B.contiguous()
    norma = (A.abs().sum(0) if trans else A.abs().sum(1)).amax()
    Xc = B.t().contiguous()                 # column j of X at Xc[j]: leading dimension n
    hctx.lu_solve(n, k, bs, F, n, Xc, n, trans=trans)
    h = None
    steps = 0
    for it in range(maxit):
        X = Xc.t().contiguous()
        R = residual(hctx, A, X, B, trans)
        h = harness(norma, R, X, B, n)
        steps = it + 1
        if it >= min_steps and bool((h < tol).all()):
            return X, steps, h
        Rc = R.t().contiguous()
        hctx.lu_solve(n, k, bs, F, n, Rc, n, trans=trans)
        Xc += Rc
    raise RuntimeError("hplai.autograd: refinement did not reach harness < %g in %d steps (max %.3g)"
                       % (tol, maxit, float(h.max())))


class LUSolve(torch.autograd.Function):
    """X = A^-1 B.  opts: dict(ctx, b, fmt, outer, tol, maxit, min_steps)."""

    @staticmethod
    def forward(fctx, A, B, opts):
        if A.dtype != torch.float64 or B.dtype != torch.float64 or not A.is_cuda or not B.is_cuda:
            raise TypeError("hplai.autograd.solve: A and B must be float64 CUDA tensors")
        n = A.shape[0]
        if A.shape != (n, n) or B.shape[0] != n:
            raise ValueError("hplai.autograd.solve: A must be n x n and B n x k")
        hctx = opts["ctx"]
        F = A.detach().t().contiguous().float()        # FP32 column-major copy, factored in place
        hctx.factor(n, opts["b"], F, n, fmt=opts["fmt"], outer=opts["outer"])
        X, steps, h = refine(hctx, A.detach(), F, opts["b"], B.detach(), False, opts["tol"],
                             opts["maxit"], opts["min_steps"])
        fctx.save_for_backward(A, X)
        fctx.F = F
        fctx.opts = opts
        opts["forward_steps"] = steps
        opts["forward_harness"] = float(h.max())
        return X

    @staticmethod
    def backward(fctx, W):
        A, X = fctx.saved_tensors
        opts = fctx.opts
        Lam, steps, h = refine(opts["ctx"], A, fctx.F, opts["b"], W.contiguous(), True, opts["tol"],
                               opts["maxit"], opts["min_steps"])
        opts["backward_steps"] = steps
        opts["backward_harness"] = float(h.max())
        dA = outer_grad(opts["ctx"], Lam, X) if fctx.needs_input_grad[0] else None
        dB = Lam if fctx.needs_input_grad[1] else None
        return dA, dB, None


def solve(A, B, ctx=None, b=512, fmt="f16", outer=1, tol=16.0, maxit=50, min_steps=0, info=None):
    """Differentiable X = A^-1 B with the library's mixed-precision LU.

    ctx: an hplai.Context (one is created on A's device if None).  info: an
    optional dict that receives the IR step counts and final harness values."""
    if ctx is None:
        ctx = Context(A.device.index or 0)
    opts = info if info is not None else {}
    opts.update(ctx=ctx, b=b, fmt=fmt if isinstance(fmt, str) else
                {v: k for k, v in FMTS.items()}[fmt], outer=outer, tol=tol, maxit=maxit,
                min_steps=min_steps)
    return LUSolve.apply(A, B, opts)
