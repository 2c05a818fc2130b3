"""hplai — Python binding of the MI355X HPL-AI solver (libhplai_gfx950.so).

A thin ctypes layer over the C ABI declared in include/hplai.h.  Device
buffers are passed as raw pointers; torch tensors are accepted wherever a
device pointer is expected (torch provides device memory and streams only).

The native library must exist: nothing here falls back to a CPU path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libhplai_gfx950.so")
DRIVER_PATH = os.path.join(os.path.dirname(_HERE), "bin", "driver.out")

OK, EINVAL, EHIP, ENOMEM, ENOCONV, EPIVOT = 0, -1, -2, -3, -4, -5
FMT_F16, FMT_F16_TILE, FMT_E4M3 = 0, 1, 2
FMTS = {"f16": FMT_F16, "f16t": FMT_F16_TILE, "e4m3": FMT_E4M3, "fp8": FMT_E4M3}
MATRIX_HPLAI, MATRIX_HIGHAM, MATRIX_HIGHAM_KAPPA = 0, 1, 2
MATRICES = {"hplai": MATRIX_HPLAI, "higham": MATRIX_HIGHAM, "higham_kappa": MATRIX_HIGHAM_KAPPA}


def higham_coeffs(n, cond=1000.0, rho=0.5):
    """(alpha, beta) of Higham's HPL-AI matrix of order n with cond(A, inf) = cond
    (host computation, no GPU needed)."""
    a, b = _dbl(), _dbl()
    check(lib().hplai_higham_coeffs(n, cond, rho, ctypes.byref(a), ctypes.byref(b)))
    return a.value, b.value

_i64, _u64, _int, _dbl = ctypes.c_int64, ctypes.c_uint64, ctypes.c_int, ctypes.c_double
_vp = ctypes.c_void_p


class HplaiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("hplai error %d: %s" % (code, msg))
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [("n", _i64), ("b", _int), ("seed", _u64), ("fmt", _int), ("outer", _int),
                ("lookahead", _int), ("lazy", _int), ("maxit", _int), ("verbose", _int),
                ("keep_factors", _int), ("profile", _int), ("matrix", _int), ("cond", _dbl),
                ("trail16", _int), ("split", _int)]


class Result(ctypes.Structure):
    _fields_ = [("time_gen", _dbl), ("time_lu", _dbl), ("time_ir", _dbl), ("time_total", _dbl),
                ("gflops", _dbl), ("ir_steps", _int), ("residual", _dbl), ("harness", _dbl),
                ("trace_resid", _dbl * 64), ("trace_harness", _dbl * 64), ("checksum", _dbl * 3),
                ("schur_launches", _i64), ("schur_flops", _dbl), ("schur_ms", _dbl),
                ("schur_large_launches", _i64), ("schur_large_flops", _dbl), ("schur_large_ms", _dbl),
                ("schur_large_bytes", _dbl), ("schur_inner_launches", _i64), ("schur_inner_flops", _dbl),
                ("schur_inner_ms", _dbl), ("schur_class_launches", _i64 * 4), ("schur_class_flops", _dbl * 4),
                ("schur_class_ms", _dbl * 4), ("phase_seconds", _dbl * 21), ("phase_ops", _dbl * 21)]

SCHUR_CLASSES = ["trailing", "rows_below", "in_block", "block_row"]

PHASES = ["DIAG_BCAST", "LCOL_BCAST", "RROW_BCAST", "TEST", "WAIT", "DIAG_LU", "TRSM_L", "TRSM_R", "CONV_L",
          "CONV_R", "GEMM_UPDATE", "GEMM_PROGRESS", "LAZY_INIT", "WRITE_BACK", "IR_GEMV", "IR_GEMV_COMM",
          "IR_TRSV", "IR_TRSV_MV", "IR_TRSV_COMM", "TOTAL", "MISC"]


# (name, restype, argtypes) for every exported entry point of include/hplai.h
class BatchStats(ctypes.Structure):
    _fields_ = [("seconds", _dbl), ("systems_per_s", _dbl), ("p50_us", _dbl), ("p99_us", _dbl),
                ("max_us", _dbl), ("failed", ctypes.c_int32), ("max_steps", ctypes.c_int32),
                ("max_harness", _dbl)]


_SIGS = [
    ("hplai_create", _int, [_int, ctypes.POINTER(_vp)]),
    ("hplai_destroy", _int, [_vp]),
    ("hplai_last_error", ctypes.c_char_p, []),
    ("hplai_version", _int, []),
    ("hplai_gen_matrix_f32", _int, [_vp, _i64, _u64, _i64, _i64, _i64, _i64, _vp, _vp, _i64, _vp]),
    ("hplai_gen_diag_f64", _int, [_vp, _i64, _u64, _i64, _i64, _vp, _vp]),
    ("hplai_gen_rhs_f64", _int, [_vp, _i64, _u64, _i64, _i64, _vp, _vp]),
    ("hplai_getrf_nopiv_f32", _int, [_vp, _int, _vp, _i64, _vp, _vp]),
    ("hplai_trsm_lpanel_f32", _int, [_vp, _i64, _int, _vp, _i64, _vp, _i64, _vp]),
    ("hplai_trsm_upanel_f32", _int, [_vp, _int, _i64, _vp, _i64, _vp, _i64, _vp]),
    ("hplai_quantize", _int, [_vp, _int, _int, _i64, _i64, _int, _vp, _i64, _vp, _i64, _vp, _vp]),
    ("hplai_schur_update", _int, [_vp, _int, _i64, _i64, _int, _int, _vp, _i64, _vp, _i64, _vp, _i64,
                                  _vp, _i64, _int, _vp, _i64, _vp]),
    ("hplai_gemmschur_f16", _int, [_vp, _i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64, _vp]),
    ("hplai_schur_update_f16c", _int, [_vp, _int, _i64, _i64, _int, _int, _vp, _i64, _vp, _i64, _vp, _i64,
                                       _vp, _i64, _int, _vp, _i64, ctypes.c_float, _vp]),
    ("hplai_otf_gemv_f64", _int, [_vp, _i64, _u64, _vp, _dbl, _vp, _vp, _vp]),
    ("hplai_otf_gemv_f64_noalloc", _int, [_vp, _i64, _u64, _vp, _dbl, _vp, _vp, _vp]),
    ("hplai_higham_coeffs", _int, [_i64, _dbl, _dbl, ctypes.POINTER(_dbl), ctypes.POINTER(_dbl)]),
    ("hplai_gen_higham_f32", _int, [_vp, _i64, _dbl, _dbl, _i64, _i64, _i64, _i64, _vp, _i64, _vp]),
    ("hplai_higham_gemv_f64", _int, [_vp, _i64, _dbl, _dbl, _dbl, _vp, _vp, _vp]),
    ("hplai_trsv_lower_f32", _int, [_vp, _i64, _int, _vp, _i64, _vp, _vp]),
    ("hplai_trsv_upper_f32", _int, [_vp, _i64, _int, _vp, _i64, _vp, _vp]),
    ("hplai_norm_inf_f64", _int, [_vp, _i64, _vp, _vp, _vp]),
    ("hplai_factor_f32", _int, [_vp, _i64, _int, _int, _int, _int, _vp, _i64]),
    ("hplai_lu_solve_f64", _int, [_vp, _int, _i64, _int, _int, _vp, _i64, _vp, _i64, _vp]),
    ("hplai_gemm_f64", _int, [_vp, _int, _int, _i64, _i64, _i64, _dbl, _vp, _i64, _vp, _i64, _dbl, _vp, _i64,
                              _vp]),
    ("hplai_batch_create", _int, [_vp, _int, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    ("hplai_batch_destroy", _int, [_vp]),
    ("hplai_batch_set_nb", _int, [_vp, _vp]),
    ("hplai_batch_get_nb", _int, [_vp, _vp]),
    ("hplai_batch_autotune", _int, [_vp, _int, _vp]),
    ("hplai_batch_run", _int, [_vp, _int, _int, ctypes.POINTER(BatchStats)]),
    ("hplai_batch_results", _int, [_vp, _vp, _vp, _vp, _vp]),
    ("hplai_batch_solution", _int, [_vp, _int, _vp]),
    ("hplai_solve", _int, [_vp, ctypes.POINTER(Params), ctypes.POINTER(Result)]),
    ("hplai_solution_x", _int, [_vp, _vp]),
    ("hplai_solution_factors", _int, [_vp, _vp, _i64]),
    ("hplai_solution_block", _int, [_vp, _i64, _i64, _i64, _i64, _vp, _i64]),
    ("hplai_release", _int, [_vp]),
]

_lib = None


def lib():
    """Load libhplai_gfx950.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libhplai_gfx950.so not built (run `make -C hpl-ai_amd` or "
                               "__graft_entry__.build()): %s" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def last_error():
    return lib().hplai_last_error().decode()


def check(rc):
    if rc != OK:
        raise HplaiError(rc, last_error())
    return rc


def ptr(x):
    """Device pointer of a torch tensor, or an int / None passed through."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    return x.data_ptr()


def stream_ptr(stream=None):
    if stream is None:
        try:
            import torch
            return torch.cuda.current_stream().cuda_stream
        except Exception:  # pragma: no cover - torch without a device
            return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


class Context:
    """An hplai_ctx bound to one device (default: torch's current device)."""

    def __init__(self, device=None):
        if device is None:
            try:
                import torch
                device = torch.cuda.current_device()
            except Exception:
                device = 0
        h = _vp()
        check(lib().hplai_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            lib().hplai_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- generator ----
    def gen_matrix(self, n, seed, i0, j0, m, k, d, A, lda, stream=None):
        check(lib().hplai_gen_matrix_f32(self.handle, n, seed, i0, j0, m, k, ptr(d), ptr(A), lda,
                                         stream_ptr(stream)))

    def gen_diag(self, n, seed, i0, cnt, d, stream=None):
        check(lib().hplai_gen_diag_f64(self.handle, n, seed, i0, cnt, ptr(d), stream_ptr(stream)))

    def gen_rhs(self, n, seed, i0, cnt, b, stream=None):
        check(lib().hplai_gen_rhs_f64(self.handle, n, seed, i0, cnt, ptr(b), stream_ptr(stream)))

    # ---- LU blocks ----
    def getrf(self, n, A, lda, info=None, stream=None):
        check(lib().hplai_getrf_nopiv_f32(self.handle, n, ptr(A), lda, ptr(info), stream_ptr(stream)))

    def trsm_lpanel(self, m, b, U, ldu, L, ldl, stream=None):
        check(lib().hplai_trsm_lpanel_f32(self.handle, m, b, ptr(U), ldu, ptr(L), ldl, stream_ptr(stream)))

    def trsm_upanel(self, b, n, Lkk, ldl, U, ldu, stream=None):
        check(lib().hplai_trsm_upanel_f32(self.handle, b, n, ptr(Lkk), ldl, ptr(U), ldu, stream_ptr(stream)))

    def quantize(self, fmt, side, m, k, tb, src, lds, dst, ldd, exps=None, stream=None):
        check(lib().hplai_quantize(self.handle, fmt, side, m, k, tb, ptr(src), lds, ptr(dst), ldd,
                                   ptr(exps), stream_ptr(stream)))

    def schur_update(self, fmt, m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu, eU, ldeu, tb, C,
                     ldc, stream=None):
        check(lib().hplai_schur_update(self.handle, fmt, m, n, nslices, kslice, ptr(Lq), ldl, ptr(eL),
                                       ldel, ptr(Uq), ldu, ptr(eU), ldeu, tb, ptr(C), ldc,
                                       stream_ptr(stream)))

    def gemmschur_f16(self, n, m, k, a, lda, b, ldb, c, ldc, stream=None):
        """The reference's gemmschur<fp16, float>: C(n x m) -= A(n x k) B(k x m),
        column-major binary16 operands, FP32 C."""
        check(lib().hplai_gemmschur_f16(self.handle, n, m, k, ptr(a), lda, ptr(b), ldb, ptr(c), ldc,
                                        stream_ptr(stream)))

    def schur_update_f16c(self, fmt, m, n, nslices, kslice, Lq, ldl, eL, ldel, Uq, ldu, eU, ldeu, tb, C,
                          ldc, cscale, stream=None):
        """C (FP16 bits, m x n column-major) := fp16(C - cscale * L U)."""
        check(lib().hplai_schur_update_f16c(self.handle, fmt, m, n, nslices, kslice, ptr(Lq), ldl, ptr(eL),
                                            ldel, ptr(Uq), ldu, ptr(eU), ldeu, tb, ptr(C), ldc, cscale,
                                            stream_ptr(stream)))

    # ---- IR ----
    # ---- Higham's HPL-AI matrix ----
    def gen_higham(self, n, alpha, beta, i0, j0, m, k, A, lda, stream=None):
        check(lib().hplai_gen_higham_f32(self.handle, n, alpha, beta, i0, j0, m, k, ptr(A), lda,
                                         stream_ptr(stream)))

    def higham_gemv(self, n, alpha, beta, scale, x, y, stream=None):
        check(lib().hplai_higham_gemv_f64(self.handle, n, alpha, beta, scale, ptr(x), ptr(y),
                                          stream_ptr(stream)))

    def otf_gemv(self, n, seed, d, alpha, x, y, stream=None):
        check(lib().hplai_otf_gemv_f64(self.handle, n, seed, ptr(d), alpha, ptr(x), ptr(y),
                                       stream_ptr(stream)))

    def otf_gemv_noalloc(self, n, seed, d, alpha, x, y, stream=None):
        """otf_gemv without a device allocation (partial sums in the library's
        per-device 32 MiB scratch array; n <= 92160)."""
        check(lib().hplai_otf_gemv_f64_noalloc(self.handle, n, seed, ptr(d), alpha, ptr(x), ptr(y),
                                               stream_ptr(stream)))

    def trsv_lower(self, n, bs, LU, lda, x, stream=None):
        check(lib().hplai_trsv_lower_f32(self.handle, n, bs, ptr(LU), lda, ptr(x), stream_ptr(stream)))

    def trsv_upper(self, n, bs, LU, lda, x, stream=None):
        check(lib().hplai_trsv_upper_f32(self.handle, n, bs, ptr(LU), lda, ptr(x), stream_ptr(stream)))

    def norm_inf(self, n, x, out, stream=None):
        check(lib().hplai_norm_inf_f64(self.handle, n, ptr(x), ptr(out), stream_ptr(stream)))

    # ---- dense factor and multi-RHS solves ----
    def factor(self, n, b, A, lda, fmt="f16", outer=1, lookahead=True):
        """LU of the n x n FP32 column-major device matrix A in place."""
        code = FMTS[fmt] if isinstance(fmt, str) else fmt
        check(lib().hplai_factor_f32(self.handle, n, b, code, outer, int(lookahead), ptr(A), lda))

    def gemm_f64(self, transa, transb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, stream=None):
        """C := alpha op(A) op(B) + beta C in FP64 (column-major, op = transpose when trans*)."""
        check(lib().hplai_gemm_f64(self.handle, 1 if transa else 0, 1 if transb else 0, m, n, k, alpha, ptr(A),
                                   lda, ptr(B), ldb, beta, ptr(C), ldc, stream_ptr(stream)))

    def lu_solve(self, n, nrhs, bs, LU, lda, X, ldx, trans=False, stream=None):
        """X := (LU)^-1 X, or (LU)^-T X with trans; X FP64 with nrhs columns."""
        check(lib().hplai_lu_solve_f64(self.handle, 1 if trans else 0, n, nrhs, bs, ptr(LU), lda,
                                       ptr(X), ldx, stream_ptr(stream)))

    # ---- whole solve ----
    def solve(self, n, b, fmt="f16", outer=1, lookahead=True, lazy=False, seed=42, maxit=50,
              verbose=False, keep_factors=False, profile=False, matrix="hplai", cond=0.0, trail16=False,
              split=0):
        """matrix: "hplai" (the reference's LCG matrix), "higham" (--highammat as
        the reference generates it) or "higham_kappa" (Fasi and Higham's signs,
        cond(A, inf) = cond exactly); cond 0 = 1000.  trail16 (with lazy): the
        partial sums of blocks not yet factored kept in FP16 (2: with the
        compensated back buffer).  split (0 < split < outer): each outer panel
        is factored in two parts, the second taking the first split inner
        blocks in one update."""
        p = Params(n=n, b=b, seed=seed, fmt=FMTS[fmt] if isinstance(fmt, str) else fmt, outer=outer,
                   lookahead=int(lookahead), lazy=1 if lazy else 0, maxit=maxit,
                   verbose=1 if verbose else 0, keep_factors=1 if keep_factors else 0,
                   profile=1 if profile else 0, matrix=MATRICES[matrix] if isinstance(matrix, str) else matrix,
                   cond=cond, trail16=int(trail16), split=int(split))
        r = Result()
        rc = lib().hplai_solve(self.handle, ctypes.byref(p), ctypes.byref(r))
        if rc not in (OK, ENOCONV):
            raise HplaiError(rc, last_error())
        k = r.ir_steps
        return dict(converged=(rc == OK), time_gen=r.time_gen, time_lu=r.time_lu, time_ir=r.time_ir,
                    time_total=r.time_total, gflops=r.gflops, ir_steps=k, residual=r.residual,
                    harness=r.harness, trace_resid=list(r.trace_resid[:k]),
                    trace_harness=list(r.trace_harness[:k]), checksum=list(r.checksum),
                    schur_launches=r.schur_launches, schur_flops=r.schur_flops, schur_ms=r.schur_ms,
                    schur_large_launches=r.schur_large_launches, schur_large_flops=r.schur_large_flops,
                    schur_large_ms=r.schur_large_ms, schur_large_bytes=r.schur_large_bytes,
                    schur_inner_launches=r.schur_inner_launches, schur_inner_flops=r.schur_inner_flops,
                    schur_inner_ms=r.schur_inner_ms,
                    schur_classes={name: (r.schur_class_launches[i], r.schur_class_flops[i], r.schur_class_ms[i])
                                   for i, name in enumerate(SCHUR_CLASSES)},
                    phases={PHASES[i]: (r.phase_seconds[i], r.phase_ops[i]) for i in range(21)})

    def solution_x(self, n):
        import numpy as np
        x = np.zeros(n, dtype=np.float64)
        check(lib().hplai_solution_x(self.handle, x.ctypes.data))
        return x

    def solution_factors(self, n):
        import numpy as np
        lu = np.zeros((n, n), dtype=np.float32)   # lu[j, i] = factor (i, j)
        check(lib().hplai_solution_factors(self.handle, lu.ctypes.data, n))
        return lu

    def solution_block(self, i0, j0, m, k):
        """The m x k block of the kept factors at (i0, j0), stored [col, row]."""
        import numpy as np
        blk = np.zeros((k, m), dtype=np.float32)
        check(lib().hplai_solution_block(self.handle, i0, j0, m, k, blk.ctypes.data, m))
        return blk

    def release(self):
        check(lib().hplai_release(self.handle))


def exported_symbols():
    return [name for name, _, _ in _SIGS]


# ---------------------------------------------------------------------------
# batched solver
# ---------------------------------------------------------------------------
BATCH_MODES = {"persistent": 0, "graph": 1, "per_system": 2, "per_system_4streams": 3, "grid_per_system": 4}


def batch_spec(count=4096, seed=2024, nmin=32, nmax=1024, rhs_max=4):
    """Seeded batch description: sizes uniform in [nmin, nmax], 1..rhs_max
    right-hand sides, and one LCG seed per system."""
    import numpy as np
    rng = np.random.default_rng(seed)
    n = rng.integers(nmin, nmax + 1, size=count).astype(np.int32)
    nrhs = rng.integers(1, rhs_max + 1, size=count).astype(np.int32)
    seeds = rng.integers(1, 2 ** 62, size=count, dtype=np.int64).astype(np.uint64)
    return n, nrhs, seeds


class Batch:
    """Many independent HPL-AI systems solved by one persistent launch
    (hplai_batch_* in hplai.h)."""

    def __init__(self, ctx, n, nrhs, seeds):
        import numpy as np
        self.ctx = ctx
        self.n = np.ascontiguousarray(n, dtype=np.int32)
        self.nrhs = np.ascontiguousarray(nrhs, dtype=np.int32)
        self.seeds = np.ascontiguousarray(seeds, dtype=np.uint64)
        self.count = int(self.n.size)
        h = _vp()
        check(lib().hplai_batch_create(ctx.handle, self.count, self.n.ctypes.data, self.nrhs.ctypes.data,
                                       self.seeds.ctypes.data, ctypes.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().hplai_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_nb(self, nb):
        arr = (ctypes.c_int32 * 4)(*nb)
        check(lib().hplai_batch_set_nb(self.handle, arr))

    def get_nb(self):
        arr = (ctypes.c_int32 * 4)()
        check(lib().hplai_batch_get_nb(self.handle, arr))
        return list(arr)

    def autotune(self, maxit=50):
        arr = (ctypes.c_int32 * 4)()
        check(lib().hplai_batch_autotune(self.handle, maxit, arr))
        return list(arr)

    def run(self, mode="persistent", maxit=50):
        st = BatchStats()
        check(lib().hplai_batch_run(self.handle, BATCH_MODES[mode], maxit, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in BatchStats._fields_}

    def results(self):
        import numpy as np
        h = np.zeros((self.count, 4), dtype=np.float64)
        st = np.zeros((self.count, 4), dtype=np.int32)
        info = np.zeros(self.count, dtype=np.int32)
        us = np.zeros(self.count, dtype=np.float64)
        check(lib().hplai_batch_results(self.handle, h.ctypes.data, st.ctypes.data, info.ctypes.data,
                                        us.ctypes.data))
        return h, st, info, us

    def solution(self, s):
        """x of system s as an (n_s, nrhs_s) array."""
        import numpy as np
        x = np.zeros(int(self.n[s]) * int(self.nrhs[s]), dtype=np.float64)
        check(lib().hplai_batch_solution(self.handle, int(s), x.ctypes.data))
        return x.reshape(int(self.nrhs[s]), int(self.n[s])).T

