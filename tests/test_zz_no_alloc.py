"""hplai_otf_gemv_f64_noalloc: the residual GEMV with its partial sums in the
library's per-device scratch array instead of a per-call allocation.

Checked through the C ABI against hplai_otf_gemv_f64 (the same kernels, so
the results must be bit-identical) on two calls in a row, the second with a
smaller n (the array keeps the first call's partial sums), and at the small
size against the oracle's FP64 residual (otf_gemv.cpp:5-131) within the
1e-13 relative bound of the existing GEMV test.  Sizes beyond the array are
rejected on the host before any launch."""
import numpy as np
import pytest

import hplai

pytestmark = pytest.mark.gpu
SEED = 42
N_MAX = 92160     # largest n whose n * ceil(n / 2048) FP64 partial sums fit in 32 MiB


def _call_both(ctx, n, seed):
    import torch
    rng = np.random.default_rng(n)
    d = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.gen_diag(n, seed, 0, n, d)
    x = torch.from_numpy(rng.standard_normal(n)).cuda()
    y0 = torch.from_numpy(rng.standard_normal(n)).cuda()
    y_alloc = y0.clone()
    y_noalloc = y0.clone()
    ctx.otf_gemv(n, seed, d, -1.0, x, y_alloc)
    ctx.otf_gemv_noalloc(n, seed, d, -1.0, x, y_noalloc)
    torch.cuda.synchronize()
    return d.cpu().numpy(), x.cpu().numpy(), y0.cpu().numpy(), y_alloc.cpu().numpy(), y_noalloc.cpu().numpy()


@pytest.mark.parametrize("sizes", [(N_MAX, 3001), (5000, 1200)], ids=["max-then-ragged", "5000-then-1200"])
def test_noalloc_matches_allocating_entry_point(ctx, oracle, sizes):
    for n in sizes:
        print("otf_gemv_noalloc n=%d" % n, flush=True)
        d, x, y0, ya, yn = _call_both(ctx, n, SEED)
        assert np.array_equal(ya.view(np.uint64), yn.view(np.uint64)), n
        if n <= 5000:
            ref = oracle.otf_gemv(n, d, -1.0, x, y0, SEED)
            scale = np.max(np.abs(y0)) + np.max(np.abs(d)) * 2 * np.max(np.abs(x))
            assert np.max(np.abs(yn - ref)) / scale < 1e-13


def test_noalloc_rejects_sizes_beyond_the_array(ctx):
    import torch
    n = N_MAX + 1
    buf = torch.zeros(3, dtype=torch.float64, device="cuda")   # never read: the call is refused first
    rc = hplai.lib().hplai_otf_gemv_f64_noalloc(ctx.handle, n, SEED, hplai.ptr(buf), -1.0, hplai.ptr(buf),
                                                hplai.ptr(buf), None)
    assert rc == hplai.EINVAL
    assert "scratch" in hplai.last_error()
