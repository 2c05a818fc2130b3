"""C2 (n = 61440, b = 512, plain FP16 Schur update, outer 16) against the
oracle's LU of the same arithmetic at the same n: the GPU's FP32 factors part
by part (strict L, strict U, diagonal, panel_check sums; parity.BOUNDS for
binary16 panels) and the oracle's number of residual evaluations.

The oracle's blocked LU takes about 8 minutes on 16 host threads at this n
(and 15 GB of host memory per copy of the matrix), more than the GPU suite's
other tests together, so this test runs only when HPLAI_C2_ORACLE=1 is set;
test_gpu_fullsize.py checks the same solve on every run through sampled
blocks.  Its measured result is recorded with the design notes.
"""
import os
import time

import numpy as np
import pytest

import parity as P

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(os.environ.get("HPLAI_C2_ORACLE") != "1",
                                 reason="set HPLAI_C2_ORACLE=1: the oracle's LU at n = 61440 takes ~8 min")]
SEED = 42


@pytest.mark.timeout(1500)
def test_c2_factors_match_oracle_at_full_size(ctx, oracle):
    os.environ["HPLAI_ORACLE_PROGRESS"] = "1"   # the oracle reports its LU and IR steps on stderr
    n, b = 61440, 512
    print("GPU solve n=%d b=%d f16 outer 16" % (n, b), flush=True)
    res = ctx.solve(n, b, fmt="f16", outer=16, lookahead=1, keep_factors=True)
    assert res["converged"] and res["harness"] < 16.0, res["trace_harness"]
    lu = ctx.solution_factors(n)
    ctx.release()
    print("  %.3f s, %d residuals; oracle LU (FP16 Schur, FP32 accumulate) ..." % (res["time_total"], res["ir_steps"]),
          flush=True)
    t0 = time.time()
    ref, d = oracle.lu(oracle.LU_PRECOMP, n, b, SEED)
    print("  oracle LU %.0f s" % (time.time() - t0), flush=True)
    err = P.factor_errors_chunked(lu, ref, n)
    print("  factors vs oracle: " + P.describe(err), flush=True)
    assert P.within(err, P.bounds("f16")), P.describe(err)
    del lu
    t0 = time.time()
    orc = oracle.ir(ref, n, d)
    print("  oracle IR %.0f s: %d residuals, harness %s" % (time.time() - t0, orc["steps"], orc.get("final_harness")),
          flush=True)
    assert res["ir_steps"] == orc["steps"], (res["ir_steps"], orc["steps"])
