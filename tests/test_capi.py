"""CPU checks of the C ABI boundary (no GPU calls): the library loads, exports
every entry point include/hplai.h declares, the Python binding's struct
layouts match the C compiler's, and driver.out parses its command line."""
import ctypes
import os
import re
import subprocess

import numpy as np
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


def test_e4m3_exponent_clamp(tmp_path, oracle):
    """e4m3 tile exponents become E8M0 block scales 127 - e, so the library
    keeps them in [-127, 127] (device_common.hpp, host code compiled here; no
    GPU call): tiles below ~2^-119 take e = 127 instead of a NaN scale byte,
    and the oracle clamps the same way."""
    src = tmp_path / "clamp.cpp"
    src.write_text(r'''
#include "device_common.hpp"
#include <cstdio>
int main()
{
    const float xs[] = {1.0f, 448.0f, 449.0f, 0x1p-119f, 0x1p-120f, 1e-40f, 0.0f};
    for (float x : xs)
        std::printf("%a %d %d\n", x, hplai::e4m3_exponent(x), hplai::e8m0_scale(hplai::e4m3_exponent(x)));
    std::printf("%d %d\n", hplai::e8m0_scale(200), hplai::e8m0_scale(-200));
    return 0;
}
''')
    exe = tmp_path / "clamp"
    hipcc = "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    subprocess.run([hipcc, "-x", "hip", "--offload-arch=gfx950", "-O1", "-I",
                    os.path.join(ROOT, "hpl-ai_amd", "csrc"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    rows = [ln.split() for ln in out[:7]]
    got = {float.fromhex(r[0]): (int(r[1]), int(r[2])) for r in rows}
    assert got[1.0] == (8, 119) and got[448.0] == (0, 127) and got[449.0] == (-1, 128)
    assert got[2.0 ** -119] == (127, 0) and got[2.0 ** -120] == (127, 0)
    assert got[float(np.float32(1e-40))] == (127, 0) and got[0.0] == (0, 127)
    assert out[7].split() == ["0", "254"]
    for x, (e, _) in got.items():
        assert oracle.tile_exp(2, x) == e, x


def test_seam_example_links_and_reports_without_gpu(tmp_path):
    """bin/seam_example (the reference's gemmschur bound from plain C, built
    with gcc) loads the library; with no GPU its hplai_create call returns an
    error, which it reports instead of crashing."""
    import hplai
    exe = os.path.join(os.path.dirname(hplai.DRIVER_PATH), "seam_example")
    if not os.path.exists(exe):
        pytest.skip("seam_example not built")
    fin = tmp_path / "in.bin"
    np.array([2, 2, 1, 2, 1, 2], dtype=np.int32).tofile(fin)
    with open(fin, "ab") as f:
        np.zeros(2, np.uint16).tofile(f)
        np.zeros(2, np.uint16).tofile(f)
        np.zeros(4, np.float32).tofile(f)
    out = subprocess.run([exe, str(fin), str(tmp_path / "out.bin")], capture_output=True, text=True, timeout=60)
    try:
        import torch
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if has_gpu:
        assert out.returncode == 0, out.stderr
    else:
        assert out.returncode == 1 and "seam_example" in out.stderr
