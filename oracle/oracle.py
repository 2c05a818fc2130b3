"""ctypes wrapper over the CPU oracle (oracle/libhplai_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results against.  The product (hpl-ai_amd/) never
imports it.

Also exposes the reference-source shim (oracle/_ref/libhplai_ref.so) when it
was built here (it needs the reference tree; it is absent on machines where
only this repository exists).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhplai_oracle.so")
REF_PATH = os.path.join(_HERE, "_ref", "libhplai_ref.so")

LU_PRECOMP, LU_LAZY, LU_FP8, LU_F16TILE = 0, 1, 2, 3

_i64, _u64, _int, _dbl, _flt = ctypes.c_int64, ctypes.c_uint64, ctypes.c_int, ctypes.c_double, ctypes.c_float
_vp = ctypes.c_void_p


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_lcg_pow.argtypes = [_u64, ctypes.POINTER(_u64), ctypes.POINTER(_u64)]
        L.orc_lcg_state.argtypes = [_u64, _u64]
        L.orc_lcg_state.restype = _u64
        L.orc_lcg_value.argtypes = [_u64]
        L.orc_lcg_value.restype = _dbl
        L.orc_elem.argtypes = [_i64, _u64, _i64, _i64]
        L.orc_elem.restype = _dbl
        L.orc_gen_block_f32.argtypes = [_i64, _u64, _i64, _i64, _i64, _i64, _vp, _i64, _int]
        L.orc_gen_diag_f64.argtypes = [_i64, _u64, _i64, _i64, _vp]
        L.orc_gen_rhs_f64.argtypes = [_i64, _u64, _i64, _i64, _vp]
        L.orc_f32_to_f16.argtypes = [_flt]
        L.orc_f32_to_f16.restype = ctypes.c_uint16
        L.orc_f16_to_f32.argtypes = [ctypes.c_uint16]
        L.orc_f16_to_f32.restype = _flt
        L.orc_f32_to_e4m3.argtypes = [_flt]
        L.orc_f32_to_e4m3.restype = ctypes.c_uint8
        L.orc_e4m3_to_f32.argtypes = [ctypes.c_uint8]
        L.orc_e4m3_to_f32.restype = _flt
        L.orc_round_f16_array.argtypes = [_i64, _vp, _vp]
        L.orc_f32_to_f16_array.argtypes = [_i64, _vp, _vp]
        L.orc_f32_to_e4m3_array.argtypes = [_i64, _vp, _vp]
        L.orc_tile_exp.argtypes = [_int, _flt]
        L.orc_tile_exp.restype = _int
        L.orc_getrf_nopiv_f32.argtypes = [_int, _vp, _i64]
        L.orc_trsm_runn_f32.argtypes = [_i64, _int, _vp, _i64, _vp, _i64]
        L.orc_trsm_llnu_f32.argtypes = [_int, _i64, _vp, _i64, _vp, _i64]
        L.orc_schur_f32acc.argtypes = [_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64]
        L.orc_schur_f16acc.argtypes = [_i64, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64]
        L.orc_lu.argtypes = [_int, _i64, _int, _u64, _vp, _vp, _i64, _int]
        L.orc_lu.restype = _int
        L.orc_lu_trail16.argtypes = [_int, _int, _i64, _int, _u64, _vp, _vp, _i64, _int]
        L.orc_lu_trail16.restype = _int
        L.orc_otf_gemv.argtypes = [_i64, _u64, _vp, _dbl, _vp, _vp, _int]
        L.orc_otf_gemv_rows.argtypes = [_i64, _u64, _vp, _dbl, _vp, _vp, _i64, _i64, _int]
        L.orc_trsv_lower_unit.argtypes = [_i64, _vp, _i64, _vp]
        L.orc_trsv_upper.argtypes = [_i64, _vp, _i64, _vp]
        L.orc_infnorm.argtypes = [_i64, _vp]
        L.orc_infnorm.restype = _dbl
        L.orc_ir.argtypes = [_i64, _u64, _vp, _i64, _vp, _vp, _vp, _int, _vp, _vp,
                             ctypes.POINTER(_dbl), ctypes.POINTER(_dbl), _int]
        L.orc_ir.restype = _int
        L.orc_checksum.argtypes = [_i64, _vp, _i64, _vp]
        _lib = L
    return _lib


def ref():
    """The reference-source shim, or None when it was not built here."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        R = ctypes.CDLL(REF_PATH)
        R.ref_fill_panel_f32.argtypes = [_int, _int, _int, _int, _int, _vp, ctypes.c_long, _u64, _int]
        R.ref_fill_panel_f64.argtypes = [_int, _int, _int, _int, _int, _vp, ctypes.c_long, _u64, _int]
        R.ref_calc_diag.argtypes = [_int, _int, _u64]
        R.ref_calc_diag.restype = _dbl
        R.ref_mat_elem.argtypes = [_int, _int, _int, _int]
        R.ref_mat_elem.restype = _dbl
        R.ref_coeff_pow.argtypes = [_u64, ctypes.POINTER(_u64), ctypes.POINTER(_u64)]
        R.ref_f32_to_f16.argtypes = [_flt]
        R.ref_f32_to_f16.restype = ctypes.c_uint16
        R.ref_hgemm.argtypes = [_int, _int, _int, _flt, _vp, _int, _vp, _int, _flt, _vp, _int]
        R.ref_shgemm.argtypes = [_int, _int, _int, _flt, _vp, _int, _vp, _int, _flt, _vp, _int]
        R.ref_otf_gemv_kernel.argtypes = [_i64, _int, _int, _dbl, _vp, _vp, _u64]
        _ref = R
    return _ref


# ---------------------------------------------------------------- helpers
SEED = 42  # main.cpp:174


def lcg_pow(n):
    a, c = _u64(), _u64()
    lib().orc_lcg_pow(n, ctypes.byref(a), ctypes.byref(c))
    return a.value, c.value


def gen_matrix(n, seed=SEED, lda=None, with_diag=True):
    lda = lda or n
    A = np.zeros((n, lda), dtype=np.float32)  # row index = column (column-major storage)
    lib().orc_gen_block_f32(n, seed, 0, 0, n, n, _ptr(A), lda, 1 if with_diag else 0)
    return A


def gen_block(n, i0, j0, m, k, seed=SEED, with_diag=True):
    A = np.zeros((k, m), dtype=np.float32)
    lib().orc_gen_block_f32(n, seed, i0, j0, m, k, _ptr(A), m, 1 if with_diag else 0)
    return A


def gen_diag(n, seed=SEED, i0=0, cnt=None):
    cnt = n - i0 if cnt is None else cnt
    d = np.zeros(cnt, dtype=np.float64)
    lib().orc_gen_diag_f64(n, seed, i0, cnt, _ptr(d))
    return d


def gen_rhs(n, seed=SEED, i0=0, cnt=None):
    cnt = n - i0 if cnt is None else cnt
    b = np.zeros(cnt, dtype=np.float64)
    lib().orc_gen_rhs_f64(n, seed, i0, cnt, _ptr(b))
    return b


def lu(mode, n, b, seed=SEED, A=None, d=None, nthreads=0):
    """Factor the generated matrix.  Returns the FP32 LU array (column-major,
    stored as A[col, row]) and the FP64 diagonal."""
    if d is None:
        d = gen_diag(n, seed)
    if A is None:
        A = gen_matrix(n, seed) if mode != LU_LAZY else np.zeros((n, n), dtype=np.float32)
    dhigh = d.astype(np.float32)
    rc = lib().orc_lu(mode, n, b, seed, _ptr(dhigh), _ptr(A), A.shape[1], nthreads)
    if rc != 0:
        raise RuntimeError("orc_lu failed: %d" % rc)
    return A, d


def lu_trail16(qmode, backbuf, n, b, seed=SEED, d=None, nthreads=0):
    """The library's 16-bit trailing storage (lazy initialisation; qmode
    LU_PRECOMP for binary16 panels or LU_FP8), with the back buffer when
    backbuf.  Returns the FP32 LU array [col, row] and the FP64 diagonal."""
    if d is None:
        d = gen_diag(n, seed)
    A = np.zeros((n, n), dtype=np.float32)
    dhigh = d.astype(np.float32)
    rc = lib().orc_lu_trail16(qmode, 1 if backbuf else 0, n, b, seed, _ptr(dhigh), _ptr(A), n, nthreads)
    if rc != 0:
        raise RuntimeError("orc_lu_trail16 failed: %d" % rc)
    return A, d


def ir(LU, n, d, rhs=None, seed=SEED, maxit=50, nthreads=0):
    rhs = gen_rhs(n, seed) if rhs is None else rhs
    x = np.zeros(n, dtype=np.float64)
    tr = np.zeros(maxit, dtype=np.float64)
    th = np.zeros(maxit, dtype=np.float64)
    fr, fh = _dbl(), _dbl()
    it = lib().orc_ir(n, seed, _ptr(LU), LU.shape[1], _ptr(d), _ptr(rhs), _ptr(x), maxit,
                      _ptr(tr), _ptr(th), ctypes.byref(fr), ctypes.byref(fh), nthreads)
    return dict(x=x, steps=it, resid=tr[:it].copy(), harness=th[:it].copy(),
                final_resid=fr.value, final_harness=fh.value)


def checksum(LU, n):
    out = np.zeros(3, dtype=np.float64)
    lib().orc_checksum(n, _ptr(LU), LU.shape[1], _ptr(out))
    return out


def otf_gemv(n, d, alpha, x, y, seed=SEED, nthreads=0):
    y = np.array(y, dtype=np.float64, copy=True)
    lib().orc_otf_gemv(n, seed, _ptr(d), alpha, _ptr(np.ascontiguousarray(x, dtype=np.float64)),
                       _ptr(y), nthreads)
    return y


def otf_gemv_rows(n, d, alpha, x, y, i0, cnt, seed=SEED, nthreads=0):
    """Rows i0 .. i0+cnt-1 of y + alpha A x (A regenerated); d and x full length."""
    y = np.array(y, dtype=np.float64, copy=True)
    lib().orc_otf_gemv_rows(n, seed, _ptr(np.ascontiguousarray(d, dtype=np.float64)), alpha,
                            _ptr(np.ascontiguousarray(x, dtype=np.float64)), _ptr(y), i0, cnt, nthreads)
    return y


def f32_to_f16_bits(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    q = np.zeros(x.shape, dtype=np.uint16)
    lib().orc_f32_to_f16_array(x.size, _ptr(x), _ptr(q))
    return q


def f32_to_e4m3_bits(x):
    x = np.ascontiguousarray(x, dtype=np.float32)
    q = np.zeros(x.shape, dtype=np.uint8)
    lib().orc_f32_to_e4m3_array(x.size, _ptr(x), _ptr(q))
    return q


def tile_exp(fmt_code, amax):
    """fmt_code: 1 = binary16 per-tile, 2 = e4m3 (ORC_LU_FP8)."""
    return lib().orc_tile_exp(fmt_code, float(amax))


_E4M3_TABLE = None


def e4m3_to_f32(q):
    global _E4M3_TABLE
    if _E4M3_TABLE is None:
        _E4M3_TABLE = np.array([lib().orc_e4m3_to_f32(v) for v in range(256)], dtype=np.float32)
    return _E4M3_TABLE[np.asarray(q, dtype=np.uint8)]
