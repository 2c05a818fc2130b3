"""CPU tests that pin the oracle (oracle/hplai_oracle.c) before it is trusted.

Each oracle function is checked against something independent of its own
formula: the reference's compiled code (fixtures in tests/golden/ref_kernels.npz
written by tests/golden/make_golden.py, and oracle/_ref when built here), the
reference's own run outputs (tests/golden/reference_runs.json), numpy/FP64
brute force, or mathematical invariants.
"""
import json
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
SEED = 42


@pytest.fixture(scope="module")
def fx():
    return np.load(os.path.join(GOLD, "ref_kernels.npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def runs():
    with open(os.path.join(GOLD, "reference_runs.json")) as f:
        return json.load(f)


# ------------------------------------------------------------------ LCG / generator
def test_lcg_pow_matches_reference(oracle, fx):
    for n, a, c in zip(fx["pow_n"], fx["pow_a"], fx["pow_c"]):
        assert oracle.lcg_pow(int(n)) == (int(a), int(c))


def test_lcg_closed_forms(oracle):
    A = 6364136223846793005
    M = 1 << 64
    # pow(1) is the map itself; pow(2) = (a^2, a c + c)
    assert oracle.lcg_pow(1) == (A, 1)
    assert oracle.lcg_pow(2) == ((A * A) % M, (A + 1) % M)
    # stepping one state at a time matches the jump-ahead
    s = (A * SEED + 1) % M
    for idx in range(5):
        assert oracle.lib().orc_lcg_state(SEED, idx) == s
        s = (A * s + 1) % M


def test_generator_blocks_match_reference(oracle, fx):
    for i, (n, i0, j0, m, k) in enumerate(fx["blocks"]):
        want = fx["block%d" % i]
        diag = bool(i0 == j0 and m == k)
        if diag and n > 4096:
            continue
        got = oracle.gen_block(int(n), int(i0), int(j0), int(m), int(k), SEED, with_diag=diag)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (n, i0, j0)


def test_diag_and_rhs_match_reference(oracle, fx):
    d = oracle.gen_diag(1200, SEED)
    assert np.array_equal(d.view(np.uint64), fx["diag1200"].view(np.uint64))
    b = oracle.gen_rhs(1200, SEED)
    assert np.array_equal(b.view(np.uint64), fx["rhs1200"].view(np.uint64))
    for r, v in zip(fx["diag_rows_61440"], fx["diag61440"]):
        assert oracle.gen_diag(61440, SEED, int(r), 1)[0] == v


def test_generator_invariants(oracle):
    n = 600
    A = oracle.gen_matrix(n).T.astype(np.float64)   # row-major view
    off = A - np.diag(np.diag(A))
    assert np.all(np.abs(off) < 0.5)
    d = oracle.gen_diag(n)
    # the FP32 diagonal is the rounded FP64 row sum of the FP32-rounded... no:
    # of the FP64 values; it matches the off-diagonal |a| sum to FP32 accuracy
    assert np.allclose(np.diag(A), d, rtol=1e-6)
    assert np.allclose(np.sum(np.abs(off), axis=1), d, rtol=1e-6)
    assert abs(np.mean(d) - (n - 1) / 4) < 0.02 * n


# ------------------------------------------------------------------ rounding
def test_f16_matches_reference_f16c(oracle, fx):
    got = oracle.f32_to_f16_bits(fx["f16_in"])
    assert np.array_equal(got, fx["f16_out"])


def test_f16_matches_numpy_dense_sweep(oracle):
    # every (sign, exponent, top-13 mantissa bits) with low-bit patterns that
    # hit exact halfway points, just-above/below, and sticky bits
    hi = np.arange(1 << 22, dtype=np.uint32) << 10
    lows = np.array([0, 1, 0x1ff, 0x200, 0x201, 0x3ff], dtype=np.uint32)
    for lo in lows:
        u = hi | lo
        x = u.view(np.float32)
        x = x[np.isfinite(x)]
        want = x.astype(np.float16).view(np.uint16)
        got = oracle.f32_to_f16_bits(x)
        # numpy keeps NaN payloads differently; none here (finite only)
        assert np.array_equal(got, want)


def _e4m3_bruteforce(x):
    codes = np.arange(256, dtype=np.uint8)
    vals = []
    for q in codes:
        q = int(q)
        e, f = (q >> 3) & 0xF, q & 7
        if e == 15 and f == 7:
            vals.append(np.nan)
            continue
        v = f * 2.0 ** -9 if e == 0 else (1 + f / 8) * 2.0 ** (e - 7)
        vals.append(-v if q & 0x80 else v)
    vals = np.array(vals)
    pos = np.array([q for q in range(128) if not np.isnan(vals[q])])
    pv = vals[pos]
    out = np.zeros(x.shape, dtype=np.uint8)
    ax = np.abs(x.astype(np.float64))
    for idx, a in enumerate(ax):
        if a >= 448:
            q = 0x7E
        else:
            dist = np.abs(pv - a)
            best = np.min(dist)
            cand = pos[dist == best]
            q = int(cand[0]) if len(cand) == 1 else int([c for c in cand if (c & 1) == 0][0])
        if x[idx] < 0 or (x[idx] == 0 and np.signbit(x[idx])):
            q |= 0x80
        out[idx] = q
    return out


def test_e4m3_matches_bruteforce(oracle):
    rng = np.random.default_rng(1)
    x = np.concatenate([
        rng.standard_normal(3000).astype(np.float32) * 50,
        (rng.standard_normal(3000) * 1e-2).astype(np.float32),
        (rng.standard_normal(1000) * 1e-3).astype(np.float32),
        np.array([0.0, -0.0, 448.0, 449.0, 464.0, 500.0, 2 ** -9, 2 ** -10, 3 * 2 ** -11,
                  2 ** -6, 1.0625, 1.1875, -1.1875], dtype=np.float32)])
    # every exact halfway point between adjacent positive codes
    grid = np.array(sorted(set(abs(float(oracle.e4m3_to_f32(q))) for q in range(127))))
    mids = ((grid[:-1] + grid[1:]) / 2).astype(np.float32)
    x = np.concatenate([x, mids, -mids])
    got = oracle.f32_to_e4m3_bits(x)
    want = _e4m3_bruteforce(x)
    assert np.array_equal(got, want)


def test_e4m3_matches_torch_float8(oracle):
    """The oracle's FP32 -> e4m3 rounding agrees bit for bit with torch's
    float8_e4m3fn cast (round to nearest even) for every value up to the
    largest finite 448, every halfway point between adjacent codes included
    (beyond 448 torch gives NaN where the quantizer saturates; tile scaling
    keeps the quantizer's inputs at or below 448)."""
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = np.concatenate([
        np.clip(rng.standard_normal(4000) * 60, -448, 448),
        rng.standard_normal(2000) * 1e-2,
        rng.standard_normal(2000) * 3e-3,
        [0.0, -0.0, 448.0, -448.0, 2 ** -9, 2 ** -10, 3 * 2 ** -11, 2 ** -6, 1.0625]]).astype(np.float32)
    grid = np.array(sorted(set(abs(float(oracle.e4m3_to_f32(q))) for q in range(127))))
    mids = ((grid[:-1] + grid[1:]) / 2).astype(np.float32)
    x = np.concatenate([x, mids, -mids])
    got = oracle.f32_to_e4m3_bits(x)
    want = torch.from_numpy(x).to(torch.float8_e4m3fn).view(torch.uint8).numpy()
    assert np.array_equal(got, want)


def test_e4m3_decode_table(oracle):
    assert oracle.e4m3_to_f32(0x7E) == 448.0
    assert oracle.e4m3_to_f32(0x01) == 2.0 ** -9
    assert oracle.e4m3_to_f32(0x38) == 1.0
    assert np.isnan(oracle.e4m3_to_f32(0x7F))


@pytest.mark.parametrize("fc,fmax", [(2, 448.0), (1, 65504.0)])
def test_tile_exponent(oracle, fc, fmax):
    rng = np.random.default_rng(4)
    for a in np.concatenate([10.0 ** rng.uniform(-12, 6, 500), [fmax, fmax * 1.0001, 1.0]]):
        a = np.float32(a)
        e = oracle.tile_exp(fc, a)
        assert np.float64(a) * 2.0 ** e <= fmax
        assert np.float64(a) * 2.0 ** (e + 1) > fmax
    assert oracle.tile_exp(fc, 0.0) == 0


# ------------------------------------------------------------------ GEMM semantics
def test_schur_f16acc_matches_reference_hgemm(oracle, fx):
    m, n, k = (int(v) for v in fx["gemm_shape"])
    c = fx["gemm_c16"].copy()
    a = fx["gemm_a"].copy()   # keep the arrays alive across the C call
    b = fx["gemm_b"].copy()
    oracle.lib().orc_schur_f16acc(m, n, k, oracle._ptr(a), m, oracle._ptr(b), k, oracle._ptr(c), m)
    assert np.array_equal(c, fx["hgemm_out"])


def test_schur_f32acc_matches_reference_shgemm(oracle, fx):
    m, n, k = (int(v) for v in fx["gemm_shape"])
    a = fx["gemm_a"].view(np.float16).astype(np.float32)
    b = fx["gemm_b"].view(np.float16).astype(np.float32)
    c = fx["gemm_c32"].copy()
    oracle.lib().orc_schur_f32acc(m, n, k, oracle._ptr(a), m, oracle._ptr(b), k, oracle._ptr(c), m)
    assert np.array_equal(c.view(np.uint32), fx["shgemm_out"].view(np.uint32))


# ------------------------------------------------------------------ LU blocks
def test_getrf_invariant_and_textbook(oracle):
    n = 96
    A = oracle.gen_block(4 * n, n, n, n, n, SEED)
    LU = A.copy()
    oracle.lib().orc_getrf_nopiv_f32(n, oracle._ptr(LU), n)
    M = LU.T.astype(np.float64)
    L = np.tril(M, -1) + np.eye(n)
    U = np.triu(M)
    Ad = A.T.astype(np.float64)
    assert np.max(np.abs(L @ U - Ad)) / np.max(np.abs(Ad)) < 1e-6
    # textbook Doolittle in FP64 on the same input
    T = Ad.copy()
    for k in range(n):
        T[k + 1:, k] /= T[k, k]
        T[k + 1:, k + 1:] -= np.outer(T[k + 1:, k], T[k, k + 1:])
    assert np.max(np.abs(M - T)) / np.max(np.abs(T)) < 1e-5


def test_trsm_vs_fp64_solve(oracle):
    rng = np.random.default_rng(8)
    b, m = 64, 150
    U = np.triu(rng.standard_normal((b, b))) + 8 * np.eye(b)
    Uc = np.asfortranarray(U, dtype=np.float32).T.copy()   # [col, row]
    P = rng.standard_normal((b, m)).astype(np.float32)      # P[col, row] m x b
    out = P.copy()
    oracle.lib().orc_trsm_runn_f32(m, b, oracle._ptr(Uc), b, oracle._ptr(out), m)
    want = P.T.astype(np.float64) @ np.linalg.inv(Uc.T.astype(np.float64))
    assert np.max(np.abs(out.T - want)) / np.max(np.abs(want)) < 1e-5
    L = np.tril(rng.standard_normal((b, b)), -1) * 0.1 + np.eye(b)
    Lc = L.astype(np.float32).T.copy()
    Q = rng.standard_normal((m, b)).astype(np.float32)      # Q[col, row]: b x m
    outq = Q.copy()
    oracle.lib().orc_trsm_llnu_f32(b, m, oracle._ptr(Lc), b, oracle._ptr(outq), b)
    L1 = np.tril(Lc.T.astype(np.float64), -1) + np.eye(b)
    wantq = np.linalg.solve(L1, Q.T.astype(np.float64))
    assert np.max(np.abs(outq.T - wantq)) / np.max(np.abs(wantq)) < 1e-5


def test_trsv_vs_fp64_like_reference_test(oracle):
    """tests/test_panel_trsv.cpp: tttrsvL/U vs dtrsv, n up to ~3000, rtol 1e-5."""
    rng = np.random.default_rng(12)
    n = 2
    while n < 3000:
        lda = n + 11
        a = rng.uniform(0, 1, (n, lda)).astype(np.float32)   # a[col, row]
        x = rng.uniform(0, 1, n)
        M = a[:, :n].T.astype(np.float64)
        xl = x.copy()
        oracle.lib().orc_trsv_lower_unit(n, oracle._ptr(a), lda, oracle._ptr(xl))
        want = np.linalg.solve(np.tril(M, -1) + np.eye(n), x)
        if n <= 40:   # unit-lower solves with U(0,1) entries grow like 2^n; compare small n
            assert np.allclose(xl, want, rtol=1e-5, atol=1e-8)
        Ud = np.triu(M) + n * np.eye(n)
        a2 = a.copy()
        a2[:, :n] = Ud.T.astype(np.float32)
        xu = x.copy()
        oracle.lib().orc_trsv_upper(n, oracle._ptr(a2), lda, oracle._ptr(xu))
        wantu = np.linalg.solve(a2[:, :n].T.astype(np.float64), x)
        assert np.allclose(xu, wantu, rtol=1e-5, atol=1e-12)
        n = (3 * n + 1) // 2


# ------------------------------------------------------------------ IR and whole solve
def test_otf_gemv_matches_reference_kernel(oracle, fx):
    n, i0, j0, mb, nb = (int(v) for v in fx["otf_block"])
    x = fx["otf_x"]
    # dense FP64 block from the generator, times x
    blk = np.array([[oracle.lib().orc_elem(n, SEED, i0 + i, j0 + j) for j in range(nb)]
                    for i in range(mb)])
    want = fx["otf_y"] - blk @ x
    assert np.max(np.abs(fx["otf_out"] - want)) < 1e-14
    # whole-matrix oracle vs a dense FP64 product at small n
    n = 300
    d = oracle.gen_diag(n)
    Ad = np.array([[oracle.lib().orc_elem(n, SEED, i, j) for j in range(n)] for i in range(n)])
    np.fill_diagonal(Ad, d)
    rng = np.random.default_rng(0)
    xx, yy = rng.standard_normal(n), rng.standard_normal(n)
    got = oracle.otf_gemv(n, d, -1.0, xx, yy)
    scale = np.abs(Ad) @ np.abs(xx) + np.abs(yy)   # FP64 summation error is relative to this
    assert np.max(np.abs(got - (yy - Ad @ xx)) / scale) < n * 2.0 ** -52


def test_step0_known_answer(oracle, runs):
    """README:50: step 0 depends only on A, b and diag(A)."""
    n = 1200
    d = oracle.gen_diag(n)
    A = oracle.gen_matrix(n)
    r = oracle.ir(A, n, d, maxit=1)   # any factors: step 0 is before the first solve
    _, resid0, h0 = runs["readme_c1"]["ir"][0]
    assert abs(r["resid"][0] - resid0) / resid0 < 1e-13
    assert abs(r["harness"][0] - h0) / h0 < 1e-13


@pytest.mark.parametrize("idx", [0, 1])
def test_step0_series(oracle, runs, idx):
    s = runs["step0_series"][idx]
    n = s["n"]
    d = oracle.gen_diag(n)
    rhs = oracle.gen_rhs(n)
    x = rhs / d
    r = oracle.otf_gemv(n, d, -1.0, x, rhs)
    resid0 = np.max(np.abs(r))
    assert abs(resid0 - s["resid0"]) / s["resid0"] < 1e-13


def test_lazy_lu_matches_reference_run(oracle, runs):
    g = runs["c1_lazy"]
    n, b = g["n"], g["b"]
    LU, d = oracle.lu(oracle.LU_LAZY, n, b)
    ck = oracle.checksum(LU, n)
    assert np.allclose(ck, g["check"], rtol=1e-7)
    r = oracle.ir(LU, n, d)
    assert r["steps"] == len(g["ir"])
    step1 = g["ir"][1][1]
    assert abs(r["resid"][1] - step1) / step1 < 0.05
    assert r["final_harness"] < 16


def test_precomp_lu_matches_reference_run(oracle, runs):
    g = runs["c1_precomp"]
    n, b = g["n"], g["b"]
    LU, d = oracle.lu(oracle.LU_PRECOMP, n, b)
    assert np.allclose(oracle.checksum(LU, n), g["check"], rtol=1e-7)
    r = oracle.ir(LU, n, d)
    assert r["steps"] == 3
    step1 = g["ir"][0][1]
    assert abs(r["resid"][1] - step1) / step1 < 0.05
    assert r["final_harness"] < 16


@pytest.mark.parametrize("mode", ["fp8", "f16tile"])
def test_scaled_modes_converge(oracle, mode):
    n, b = 1024, 128
    code = oracle.LU_FP8 if mode == "fp8" else oracle.LU_F16TILE
    LU, d = oracle.lu(code, n, b)
    r = oracle.ir(LU, n, d)
    assert r["final_harness"] < 16
    # the FP8 factors differ from the FP16 ones only by the Schur rounding
    ref, _ = oracle.lu(oracle.LU_PRECOMP, n, b)
    assert np.max(np.abs(LU - ref)) / np.max(np.abs(ref)) < 1e-2


def test_harness_formula(oracle):
    """iterative_refinement.hpp:43: h = ||r|| / (||A|| ||x|| + ||b||) / (n eps/2)."""
    n = 1200
    d = oracle.gen_diag(n)
    rhs = oracle.gen_rhs(n)
    A = oracle.gen_matrix(n)
    r = oracle.ir(A, n, d, rhs, maxit=1)
    x0 = rhs / d
    h = r["resid"][0] / (2 * np.max(d) * np.max(np.abs(x0)) + np.max(np.abs(rhs))) / (n * 2.0 ** -53)
    assert abs(h - r["harness"][0]) / h < 1e-14


# ------------------------------------------------------------------ oracle/_ref cross-check
def test_oracle_vs_ref_shim_when_built(oracle):
    R = oracle.ref()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    for n, i0, j0, m, k in [(1200, 0, 0, 60, 60), (8192, 4096, 4096, 64, 64), (8192, 100, 7000, 50, 33)]:
        A = np.zeros((k, m), dtype=np.float32)
        diag = 1 if (i0 == j0 and m == k) else 0
        R.ref_fill_panel_f32(n, i0, j0, m, k, oracle._ptr(A), m, SEED, diag)
        got = oracle.gen_block(n, i0, j0, m, k, SEED, with_diag=bool(diag))
        assert np.array_equal(got.view(np.uint32), A.view(np.uint32))
