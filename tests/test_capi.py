"""CPU checks of the C ABI boundary (no GPU calls): the library loads, exports
every entry point include/hplai.h declares, the Python binding's struct
layouts match the C compiler's, and driver.out parses its command line."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hplai.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hplai_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared_functions()
    for must in ["hplai_gen_matrix_f32", "hplai_gen_diag_f64", "hplai_gen_rhs_f64",
                 "hplai_getrf_nopiv_f32", "hplai_trsm_lpanel_f32", "hplai_trsm_upanel_f32",
                 "hplai_quantize", "hplai_schur_update", "hplai_otf_gemv_f64",
                 "hplai_trsv_lower_f32", "hplai_trsv_upper_f32", "hplai_norm_inf_f64", "hplai_solve"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    import hplai
    L = hplai.lib()
    for name in declared_functions():
        assert hasattr(L, name), name
    # the binding wraps every declared function
    assert set(declared_functions()) <= set(hplai.exported_symbols())


def test_version_and_error_without_gpu():
    import hplai
    assert hplai.lib().hplai_version() == 100
    # a NULL context is rejected with an error code, not a crash
    rc = hplai.lib().hplai_gen_diag_f64(None, 10, 42, 0, 10, None, None)
    assert rc == hplai.EINVAL
    assert "ctx" in hplai.last_error()


def _c_layout(tmp_path):
    src = tmp_path / "layout.c"
    src.write_text(r'''
#include <stdio.h>
#include <stddef.h>
#include "hplai.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(hplai_params), offsetof(hplai_params, profile),
         sizeof(hplai_result), offsetof(hplai_result, schur_ms),
         offsetof(hplai_result, phase_ops));
  return 0;
}''')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    return [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                           check=True).stdout.split()]


def test_struct_layouts_match_c(tmp_path):
    import hplai
    sp, op, sr, orr, orl = _c_layout(tmp_path)
    assert ctypes.sizeof(hplai.Params) == sp
    assert hplai.Params.profile.offset == op
    assert ctypes.sizeof(hplai.Result) == sr
    assert hplai.Result.schur_ms.offset == orr
    assert hplai.Result.phase_ops.offset == orl


def test_driver_help():
    import hplai
    if not os.path.exists(hplai.DRIVER_PATH):
        pytest.skip("driver.out not built")
    out = subprocess.run([hplai.DRIVER_PATH, "1200", "60", "2", "--help"], capture_output=True,
                         text=True, timeout=60)
    assert "usage: ./driver.out <n> <b> <P>" in out.stdout
