"""Independent numpy restatement of the blocked mixed-precision LU.

TEST INFRASTRUCTURE ONLY (used by tests/ to pin the C oracle's LU modes and
to show that the parity metric in tests/parity.py rejects plausible
mistakes).  It is written from the reference's algorithm, not from the C
oracle, and uses different arithmetic on purpose (numpy/LAPACK FP32
triangular solves, one FP64 product per trailing update rounded once), so
that agreement with oracle/hplai_oracle.c pins the oracle up to FP32
summation order.

Algorithm (RDMAPanelLU::run, panel_trf.hpp:703-856, on a 1 x 1 grid):
for each diagonal block k of size b:
  getrf_nopiv of the diagonal block (getrf_nopiv.hpp:124-160), FP32;
  L := A U_kk^-1 and U := L_kk^-1 A (update_left/right_panels,
  panel_trf.hpp:143-171), FP32;
  quantize both panels (convert_left/right_panels, panel_trf.hpp:173-233):
  x -> q(x * 2^e) * 2^-e, q = binary16 or OCP e4m3 with round-to-nearest-even;
  e = 0 for plain FP16 (no -d), else the largest e with amax * 2^e <= the
  format's largest finite value, per b x b tile (the per-tile generalisation
  of the reference's scalea, panel_trf.hpp:717-719);
  C -= Lq Uq (schur_updator.hpp:9-47, FP32 C).

The 16-bit trailing storage (the reference's -d double-decker lower deck,
panel.hpp:45-83; lazy initialisation lazy_init.hpp:92-227; back buffer
back_buffer.hpp:81-170) is modelled as the library computes it: the partial
sums P of blocks that have not joined a panel are kept in binary16, scaled
by pscale = sqrt(n sqrt(n)) (hpl_rand.hpp:154), updated once per step as
P := fp16(P - pscale * Lq Uq); a block joining a panel gets
a_ij + (P + 32 Q) / pscale; every 32 steps P folds into the back buffer Q
with a compensated binary16 sum over the blocks that do not join the next
panel.

`mutate` plants a mistake (for the metric's own tests):
  "unscaled"      : tile exponents forced to 0 (FP8 without per-tile scaling)
  "stage"         : in the L panel of step 1, the 64-wide K chunks 0 and 1 of
                    rows 256..511 swap places (one displaced operand stage)
  "tile_exp"      : the L panel's tile exponents shifted by one tile row
  "drop_backbuf"  : the merge at lazy initialisation ignores the back buffer
"""
import numpy as np
import scipy.linalg as sla

FMAX = {"f16": 65504.0, "f16t": 65504.0, "e4m3": 448.0}


def tile_exponent(amax, fmax):
    """Largest integer e with amax * 2^e <= fmax (0 for amax == 0)."""
    amax = float(amax)
    if not amax > 0.0:
        return 0
    e = int(np.floor(np.log2(fmax / amax)))
    while amax * 2.0 ** (e + 1) <= fmax:
        e += 1
    while amax * 2.0 ** e > fmax:
        e -= 1
    return e


def _round(x, fmt):
    """Round FP32 values to the format (RNE) and back to FP32."""
    if fmt == "e4m3":
        import torch
        t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32))
        return t.to(torch.float8_e4m3fn).to(torch.float32).numpy()
    return np.asarray(x, dtype=np.float32).astype(np.float16).astype(np.float32)


def quantize_panel(x, b, fmt, side, mutate=None):
    """Dequantized copy of an L panel (side 0: rows x K) or U panel (side 1:
    K x cols), tiles of b rows (L) or b columns (U)."""
    x = np.asarray(x, dtype=np.float32)
    out = np.empty_like(x)
    ntile = (x.shape[side] + b - 1) // b
    exps = []
    for t in range(ntile):
        sl = np.s_[t * b:(t + 1) * b, :] if side == 0 else np.s_[:, t * b:(t + 1) * b]
        e = 0 if fmt == "f16" else tile_exponent(np.max(np.abs(x[sl])), FMAX[fmt])
        exps.append(e)
    if mutate == "unscaled":
        exps = [0] * ntile
    if mutate == "tile_exp" and side == 0 and ntile > 1:
        exps = exps[1:] + exps[:1]
    for t in range(ntile):
        sl = np.s_[t * b:(t + 1) * b, :] if side == 0 else np.s_[:, t * b:(t + 1) * b]
        e = exps[t]
        s = np.ldexp(x[sl], e).astype(np.float32)
        out[sl] = np.ldexp(_round(s, fmt), -e).astype(np.float32)
    return out


def getrf_nopiv(a):
    """Unblocked right-looking unpivoted LU of a square FP32 block, in place
    (getrf_nopiv_small's arithmetic, getrf_nopiv.hpp:30-46, without its
    16-column blocking)."""
    nb = a.shape[0]
    for k in range(nb - 1):
        inv = np.float32(1.0) / a[k, k]
        a[k + 1:, k] *= inv
        a[k + 1:, k + 1:] -= np.outer(a[k + 1:, k], a[k, k + 1:]).astype(np.float32)
    return a


def lu(A, b, fmt="f16", trail16=0, n=None, mutate=None):
    """Factor the n x n matrix A (numpy row-major, A[i, j] = a_ij, FP32 with
    the diagonal set) and return the LU array in the same layout.

    trail16 = 0: the matrix is factored in place (FP32 C).
    trail16 = 1: the 16-bit trailing storage with lazy initialisation; A
    supplies only the generated a_ij.  trail16 = 2: also the back buffer."""
    A = np.array(A, dtype=np.float32, copy=True)
    n = A.shape[0] if n is None else n
    nb = n // b
    if trail16:
        G = A
        A = np.zeros_like(G)
        pscale = np.float32(np.sqrt(float(n) * np.sqrt(float(n))))
        pdown = np.float32(1.0 / float(pscale))
        P = np.zeros((n, n), dtype=np.float16)
        Q = np.zeros((n, n), dtype=np.float16) if trail16 == 2 else None
    for k in range(nb):
        kk = k * b
        e = kk + b
        if trail16:
            for rs, cs in ((np.s_[kk:n], np.s_[kk:e]), (np.s_[kk:e], np.s_[e:n])):
                x = P[rs, cs].astype(np.float64)
                if Q is not None and mutate != "drop_backbuf":
                    x = (32.0 * Q[rs, cs].astype(np.float64) + x).astype(np.float32).astype(np.float64)
                A[rs, cs] = (float(pdown) * x + G[rs, cs].astype(np.float64)).astype(np.float32)
        getrf_nopiv(A[kk:e, kk:e])
        if e == n:
            break
        U11 = np.triu(A[kk:e, kk:e])
        L11 = np.tril(A[kk:e, kk:e], -1) + np.eye(b, dtype=np.float32)
        # L21 U11 = A21  ->  U11^T L21^T = A21^T
        A[e:, kk:e] = sla.solve_triangular(U11.T, A[e:, kk:e].T, lower=True, check_finite=False).T
        A[kk:e, e:] = sla.solve_triangular(L11, A[kk:e, e:], lower=True, unit_diagonal=True,
                                           check_finite=False)
        mq = mutate if mutate in ("unscaled", "tile_exp") else None
        Lq = quantize_panel(A[e:, kk:e], b, fmt, 0, mq)
        Uq = quantize_panel(A[kk:e, e:], b, fmt, 1, mq if mutate == "unscaled" else None)
        if mutate == "stage" and k == 1 and Lq.shape[0] >= 512:
            blk = Lq[256:512, 0:128].copy()
            Lq[256:512, 0:64], Lq[256:512, 64:128] = blk[:, 64:128], blk[:, 0:64]
        prod = Lq.astype(np.float64) @ Uq.astype(np.float64)
        if trail16:
            P[e:, e:] = (P[e:, e:].astype(np.float64) - float(pscale) * prod).astype(np.float16)
            if Q is not None and (k + 1) % 32 == 0 and e + b < n:
                r1 = e + b
                q0 = Q[r1:, r1:]
                t1 = P[r1:, r1:] * np.float16(1.0 / 32.0)
                t2 = q0 + t1
                t3 = t2 - q0
                t4 = t1 - t3
                Q[r1:, r1:] = t2
                P[r1:, r1:] = t4 * np.float16(32.0)
        else:
            A[e:, e:] = (A[e:, e:].astype(np.float64) - prod).astype(np.float32)
    return A


def to_colmajor(a):
    """numpy row-major [i, j] -> the oracle's / library's [j, i] storage."""
    return np.ascontiguousarray(np.asarray(a).T)


def scaled_matrix(A, b, seed=1):
    """The matrix A ([col, row] storage) with rows and columns scaled by powers
    of two that change from one b-block to the next.  D1 A D2 keeps the
    unpivoted LU stable (L -> D1 L D1^-1, U -> D1 U D2) while the per-tile
    exponents of every panel differ from tile to tile.  Returns [col, row]."""
    n = A.shape[0]
    rng = np.random.default_rng(seed)
    rs = (2.0 ** rng.integers(-3, 4, size=n // b)).repeat(b)
    cs = (2.0 ** rng.integers(-3, 4, size=n // b)).repeat(b)
    Ar = np.asarray(A).T
    return np.ascontiguousarray((Ar * rs[:, None] * cs[None, :]).astype(np.float32).T)
