"""CPU legs of bench.py: the reference's own Schur kernel timing (oracle/_ref)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import bench  # noqa: E402
from oracle import oracle as O  # noqa: E402


def test_reference_schur_kernel_timing():
    if O.ref() is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    r = bench.cpu_reference_schur(2, m=256, k=128, cols=64, reps=1)
    assert r["kind"] == "reference" and r["cores"] == 2 and r["unit"] == "TFLOP/s"
    assert np.isfinite(r["value"]) and r["value"] > 0
    assert "shgemm_opt" in r["sample"]
