"""CPU tests of the factor-level parity metric (tests/parity.py) and of the
oracle's whole-LU modes that have no reference run to pin them.

1. The oracle's LU modes (binary16, binary16 with per-tile exponents, e4m3
   with per-tile exponents, and the 16-bit trailing storage with and without
   its back buffer) agree with an independent numpy restatement
   (tests/lu_model.py: LAPACK FP32 triangular solves, one FP64 product per
   trailing update, torch's float8_e4m3fn cast) within the same-format bounds
   the GPU tests use.
2. The 16-bit trailing-storage mode reduces, at the reference's C1
   configuration, to the reference-pinned FP16-accumulate lazy mode (whose
   --checksum matches the reference run) within a stated bound.
3. The metric rejects each plausible mistake, planted in the numpy
   restatement or in the oracle's output, never on the GPU: strict L zeroed,
   e4m3 without per-tile scaling, e4m3 factors passed off as binary16 ones,
   one displaced 64-byte operand stage, per-tile exponents taken from the
   neighbouring tile (on a matrix whose tile exponents differ), a dropped
   back buffer.
"""
import json
import os

import numpy as np
import pytest

import lu_model as M
import parity as P

HERE = os.path.dirname(os.path.abspath(__file__))
SEED = 42
N, B = 1024, 128


@pytest.fixture(scope="module")
def base(oracle):
    A = oracle.gen_matrix(N)          # [col, row]
    return A, A.T.copy()


@pytest.fixture(scope="module")
def oracle_lus(oracle):
    return {f: oracle.lu(m, N, B)[0] for f, m in
            (("f16", oracle.LU_PRECOMP), ("f16t", oracle.LU_F16TILE), ("e4m3", oracle.LU_FP8))}


# ------------------------------------------------------------------ pins
@pytest.mark.parametrize("fmt", ["f16", "f16t", "e4m3"])
def test_model_pins_oracle_lu_modes(base, oracle_lus, fmt):
    """Measured: l, u 2.3e-6 .. 3e-6 (binary16) and 9.8e-7 (e4m3) at n=1024,
    b=128; up to 5.2e-5 for e4m3 at n=8192, b=512."""
    mod = M.to_colmajor(M.lu(base[1], B, fmt))
    err = P.factor_errors(mod, oracle_lus[fmt])
    assert P.within(err, P.bounds(fmt)), P.describe(err)


@pytest.mark.parametrize("fmt,backbuf", [("f16", 0), ("f16", 1), ("e4m3", 0), ("e4m3", 1)])
def test_model_pins_oracle_trail16(oracle, fmt, backbuf):
    """b = 16 gives 64 steps, so the back buffer folds once (after step 32)
    over blocks that later join panels.  Measured l, u <= 4.1e-5."""
    n, b = 1024, 16
    A = oracle.gen_matrix(n)
    ref, _ = oracle.lu_trail16(oracle.LU_PRECOMP if fmt == "f16" else oracle.LU_FP8, backbuf, n, b)
    mod = M.to_colmajor(M.lu(A.T.copy(), b, fmt, trail16=1 + backbuf))
    err = P.factor_errors(mod, ref)
    assert P.within(err, P.bounds(fmt, trail16=True)), P.describe(err)


def test_trail16_reduces_to_reference_lazy_at_c1(oracle):
    """C1 (n=1200, b=60): the reference's -d factors (ORC_LU_LAZY, FP16
    accumulation after every product; its --checksum equals the reference
    run's to 1e-7, test_oracle.py) and the 16-bit trailing storage (FP32 sums,
    one binary16 rounding per update) differ only by where the partial sums
    are rounded.  Measured: l 1.25e-4, u 1.47e-4 (the same as the FP32-C
    factors against the lazy ones, 1.2e-4); checksums within 5.2e-8 of the
    reference run; the refinement converges in the reference's 3 residuals."""
    with open(os.path.join(HERE, "golden", "reference_runs.json")) as f:
        g = json.load(f)["c1_lazy"]
    n, b = g["n"], g["b"]
    lazy, d = oracle.lu(oracle.LU_LAZY, n, b)
    t16, _ = oracle.lu_trail16(oracle.LU_PRECOMP, 0, n, b)
    err = P.factor_errors(t16, lazy)
    assert err["l"] < 3e-4 and err["u"] < 3e-4 and err["d"] < 1e-5, P.describe(err)
    assert np.allclose(oracle.checksum(t16, n), g["check"], rtol=1e-7)
    r = oracle.ir(t16, n, d)
    assert r["steps"] == len(g["ir"])
    assert r["final_harness"] < 16


def test_backbuffer_fold_keeps_the_partial_sums(oracle):
    """The fold is a compensated sum: after it, P + 32 Q equals the partial
    sum before it up to one binary16 rounding of the small remainder, so the
    back buffer changes the factors only at the level of the rounding noise
    (measured 8e-5 at n=8192, b=128), far below a dropped deck."""
    n, b = 1024, 16
    a, _ = oracle.lu_trail16(oracle.LU_PRECOMP, 0, n, b)
    c, _ = oracle.lu_trail16(oracle.LU_PRECOMP, 1, n, b)
    err = P.factor_errors(a, c)
    assert P.within(err, P.bounds("f16", trail16=True)), P.describe(err)
    assert err["l"] > 0  # the fold ran (64 steps > 32)


# ------------------------------------------------------------------ rejection
def test_metric_rejects_zeroed_L(oracle_lus):
    z = oracle_lus["e4m3"].copy()
    z[np.triu_indices(N, 1)] = 0.0
    err = P.factor_errors(z, oracle_lus["e4m3"])
    assert err["l"] == 1.0 and not P.within(err, P.bounds("e4m3"))
    # the max-normalised difference the metric replaces cannot see it
    old = np.max(np.abs(z - oracle_lus["e4m3"])) / np.max(np.abs(oracle_lus["e4m3"]))
    assert old < 1e-4


def test_metric_rejects_fp8_passed_as_f16(oracle_lus):
    err = P.factor_errors(oracle_lus["e4m3"], oracle_lus["f16"])
    assert not P.within(err, P.bounds("e4m3")), P.describe(err)
    assert min(err["l"], err["u"]) > 4 * P.bounds("e4m3")["l"]


@pytest.mark.parametrize("mutation", ["unscaled", "stage"])
def test_metric_rejects_planted_mistakes(base, oracle_lus, mutation):
    mod = M.to_colmajor(M.lu(base[1], B, "e4m3", mutate=mutation))
    err = P.factor_errors(mod, oracle_lus["e4m3"])
    assert not P.within(err, P.bounds("e4m3")), P.describe(err)
    assert max(err["l"], err["u"]) > 20 * P.bounds("e4m3")["l"]


def test_metric_rejects_dropped_backbuffer(oracle):
    n, b = 1024, 16
    ref, _ = oracle.lu_trail16(oracle.LU_FP8, 1, n, b)
    A = oracle.gen_matrix(n)
    mod = M.to_colmajor(M.lu(A.T.copy(), b, "e4m3", trail16=2, mutate="drop_backbuf"))
    err = P.factor_errors(mod, ref)
    assert not P.within(err, P.bounds("e4m3", trail16=True)), P.describe(err)
    assert max(err["l"], err["u"]) > 20 * P.bounds("e4m3", trail16=True)["l"]


def test_tile_exponents_vary_and_are_pinned(oracle):
    """On the LCG matrix every tile of a panel has the same exponent (its
    entries are all about 2/n), so an exponent taken from the wrong tile
    would go unseen; on the scaled matrix the exponents differ, the model
    and the oracle agree (measured 8e-7), and the wrong-tile mistake is
    rejected."""
    As = M.scaled_matrix(oracle.gen_matrix(N), B)
    ref, _ = oracle.lu(oracle.LU_FP8, N, B, A=As.copy(), d=np.zeros(N))
    # the first L panel's tile exponents differ
    panel = ref[:B, B:].T   # rows B..N of block column 0, [row, col]
    exps = {M.tile_exponent(np.max(np.abs(panel[t:t + B])), 448.0) for t in range(0, N - B, B)}
    assert len(exps) > 2
    mod = M.to_colmajor(M.lu(As.T.copy(), B, "e4m3"))
    err = P.factor_errors(mod, ref)
    assert P.within(err, P.bounds("e4m3")), P.describe(err)
    bad = M.to_colmajor(M.lu(As.T.copy(), B, "e4m3", mutate="tile_exp"))
    assert not P.within(P.factor_errors(bad, ref), P.bounds("e4m3"))


def test_chunked_metric_equals_the_whole_array_metric(oracle):
    """parity.factor_errors_chunked (for n too large for whole-array copies)
    gives the numbers of parity.factor_errors, here on an oracle LU with a
    perturbed L and a perturbed U entry."""
    n, b = 1200, 60
    ref, _ = oracle.lu(oracle.LU_PRECOMP, n, b)
    lu = ref.copy()
    lu[5, 100] += 1e-3      # strict L (row 100 > column 5)
    lu[700, 3] *= 1.5       # strict U (row 3 < column 700)
    whole = P.factor_errors(lu, ref)
    for step in (256, 1000, 2048):
        chunked = P.factor_errors_chunked(lu, ref, step=step)
        for k in whole:
            assert abs(whole[k] - chunked[k]) <= 1e-12 * max(1.0, whole[k]), (k, step)
