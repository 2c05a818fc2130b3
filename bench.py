"""Benchmark: HPL-AI time-to-harness<16 and TFLOP/s on one MI355X.

Workload (the metric's configuration): n = 196608, b = 512, the seeded
diagonally dominant LCG matrix (A in FP32 in HBM, ~155 GB), FP32 panel LU and
triangular solves, FP64 iterative refinement until harness < 16.  The primary
form runs the Schur update in FP8 e4m3 with per-512x512-tile power-of-two
scales; the plain FP16 form runs beside it on the same matrix.

A step is one solve: LU + IR, the reference's timed region (main.cpp:379-433);
the matrix generation before it is outside the timed region and reported
separately.  value = (2/3 n^3 + 3/2 n^2) / time (main.cpp:441), in TFLOP/s.

Also reported: the live roofline of the dominant kernel (the Schur update,
every launch timed with HIP events on its own stream inside the library) and
a CPU baseline: the CPU oracle (a port of the reference algorithm) timed on
this host's cores on a bounded sample (a smaller n).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hpl-ai_amd"))
sys.path.insert(0, ROOT)

PEAK_TFLOPS = {"e4m3": 5000.0, "f16": 2500.0, "f16t": 2500.0}   # MI355X dense MFMA
HBM_GBS = 8000.0


def hpl_flops(n):
    return 2.0 / 3.0 * n ** 3 + 3.0 / 2.0 * n ** 2


def _lookahead(args):
    return 0 if args.no_lookahead else args.lookahead


def _storage(args):
    """lazy initialisation / FP16 trailing storage options of the solve"""
    return dict(lazy=bool(args.lazy or args.trail16), trail16=bool(args.trail16))


def _split(args, outer):
    """the two-part panel split, where the form's outer block allows it"""
    return args.split if 0 < args.split < outer else 0


def run_form(ctx, args, fmt, outer):
    res = None
    for _ in range(args.warmup):
        ctx.solve(args.n, args.b, fmt=fmt, outer=outer, lookahead=_lookahead(args), maxit=50,
                  split=_split(args, outer), **_storage(args))
    import torch
    times, results = [], []
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        r = ctx.solve(args.n, args.b, fmt=fmt, outer=outer, lookahead=_lookahead(args),
                      maxit=50, profile=True, split=_split(args, outer), **_storage(args))
        if not r["converged"]:
            raise SystemExit("%s did not reach harness < 16: %s" % (fmt, r))
        times.append(r["time_total"])
        results.append(r)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - w0) / args.steps
    res = results[-1]
    t = sum(times) / len(times)
    tot = {k: sum(r[k] for r in results) for k in
           ("schur_launches", "schur_flops", "schur_ms", "schur_large_launches", "schur_large_flops",
            "schur_large_ms", "schur_large_bytes", "schur_inner_launches", "schur_inner_flops",
            "schur_inner_ms")}
    classes = {}
    for name in results[0]["schur_classes"]:
        nl = sum(r["schur_classes"][name][0] for r in results)
        fl = sum(r["schur_classes"][name][1] for r in results)
        ms = sum(r["schur_classes"][name][2] for r in results)
        classes[name] = {"launches": nl // len(results), "ms_per_solve": ms / len(results),
                         "tflops": fl / (ms * 1e-3) * 1e-12 if ms > 0 else None}
    tot["schur_classes"] = classes
    return dict(fmt=fmt, outer=outer, split=_split(args, outer), time_s=t, wall_s=wall,
                tflops=hpl_flops(args.n) / t * 1e-12,
                time_lu_s=res["time_lu"], time_ir_s=res["time_ir"], time_gen_s=res["time_gen"],
                ir_steps=res["ir_steps"], harness=res["harness"], residual=res["residual"],
                schur_tflops=(tot["schur_flops"] / (tot["schur_ms"] * 1e-3) * 1e-12)
                if tot["schur_ms"] > 0 else None, **tot)


def roofline(p, fmt, n):
    """Roofline entry for the dominant kernel, the streamed 256 x 256 Schur
    update: algorithmic flops per launch (2 m n K) over its mean launch
    duration, both from the library's per-launch HIP events on the launch
    stream.  `traffic` is the measured HBM bytes per launch of that kernel,
    read from the committed rocprofv3 counter summary of this workload when
    one exists (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel);
    `counters` the shader clock, matrix-core occupancy and L2 hit rate the
    committed counter summary measured for it."""
    peak = PEAK_TFLOPS[fmt]
    nl = p["schur_large_launches"]
    if not nl or p["schur_large_ms"] <= 0:
        return {"bound": "mfma", "achieved": None, "peak": peak, "unit": "TFLOP/s", "frac": None,
                "traffic": None}
    flops_per = p["schur_large_flops"] / nl
    ms_per = p["schur_large_ms"] / nl
    ach = flops_per / (ms_per * 1e-3) * 1e-12
    traffic = None
    summ = os.path.join(ROOT, "profiles", "schur_dma_%s_n%d_traffic.txt" % (fmt, n))
    if os.path.exists(summ):
        for line in open(summ):
            if line.startswith("hbm_bytes_per_launch"):
                traffic = float(line.split()[1])
    # clock and matrix-core occupancy of the same kernel from the committed
    # counter summary of this workload (the first kernel block of the file)
    counters = None
    csumm = os.path.join(ROOT, "profiles", "schur_counters_%s_n%d.txt" % (fmt, n))
    if os.path.exists(csumm):
        vals = {}
        for line in open(csumm):
            f = line.split()
            if len(f) == 2 and f[0] in ("sclk_mhz", "mfma_busy_pct", "l2_hit_pct") and f[0] not in vals:
                vals[f[0]] = float(f[1])
        if vals:
            counters = dict(vals, source=os.path.relpath(csumm, ROOT))
            if "sclk_mhz" in vals:
                counters["peak_at_sclk"] = peak * vals["sclk_mhz"] / 2400.0
    return {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
            "traffic": traffic, "counters": counters,
            "kernel": "schur_dma_kernel<%s> (256x256 streamed Schur update)" % fmt,
            "launches": nl, "flops_per_launch": flops_per, "ms_per_launch": ms_per,
            "algorithmic_bytes_per_launch": p["schur_large_bytes"] / nl,
            "all_schur_launches": {"launches": p["schur_launches"], "tflops": p["schur_tflops"],
                                   "ms": p["schur_ms"]},
            "by_class": p["schur_classes"],
            "in_block_launches": {"launches": p["schur_inner_launches"], "ms": p["schur_inner_ms"],
                                   "tflops": (p["schur_inner_flops"] / (p["schur_inner_ms"] * 1e-3) * 1e-12)
                                   if p["schur_inner_ms"] > 0 else None}}


def cpu_baseline(n, b, threads):
    """The CPU oracle (port of the reference's algorithm, FP16 Schur with FP32
    accumulation, FP64 IR) on a bounded sample, timed on this host."""
    from oracle import oracle as O
    O.lib()
    d = O.gen_diag(n)
    rhs = O.gen_rhs(n)
    A = O.gen_matrix(n)
    t0 = time.perf_counter()
    LU, _ = O.lu(O.LU_PRECOMP, n, b, A=A, d=d, nthreads=threads)
    r = O.ir(LU, n, d, rhs, nthreads=threads)
    t = time.perf_counter() - t0
    rate = hpl_flops(n) / t
    return dict(value=rate * 1e-12, unit="TFLOP/s", cores=threads, kind="port",
                sample="oracle LU+IR (FP16 Schur, FP32 accumulate, FP64 IR) at n=%d b=%d: "
                       "%.2f s, harness %.4f" % (n, b, t, r["final_harness"]),
                # not measured: the sample's time scaled by the flop count (about n^3)
                extrapolated_by_n3={"label": "extrapolated from the n=%d sample at its rate, not measured" % n,
                                    "n61440_s": hpl_flops(61440) / rate,
                                    "n196608_s": hpl_flops(196608) / rate})


def cpu_reference_schur(threads, m=4096, k=512, cols=1024, reps=2):
    """The reference's own CPU Schur kernel (shgemm_opt of kernels/avx2.cpp,
    FP16 operands, FP32 C, C -= A B, compiled from the reference's sources
    into oracle/_ref) on this host.  Like the reference's OpenMP Schur update
    (schur_updator.hpp:425-447: each thread takes its own range of columns of
    the trailing matrix) every thread updates its own `cols` columns of an m x
    (threads * cols) C with K = k; ctypes releases the GIL during each call.
    None when oracle/_ref was not built."""
    import threading
    import numpy as np
    from oracle import oracle as O
    R = O.ref()
    if R is None:
        return None
    rng = np.random.default_rng(7)
    a = (rng.random((k, m), dtype=np.float32) - 0.5).astype(np.float16)            # m x k, column-major
    b = (rng.random((threads * cols, k), dtype=np.float32) - 0.5).astype(np.float16)   # k x n, column-major
    c = np.zeros((threads * cols, m), dtype=np.float32)                             # m x n, column-major

    def work(t):
        for _ in range(reps):
            R.ref_shgemm(m, cols, k, -1.0, a.ctypes.data, m, b[t * cols:].ctypes.data, k, 1.0,
                         c[t * cols:].ctypes.data, m)

    pool = [threading.Thread(target=work, args=(t,)) for t in range(threads)]
    t0 = time.perf_counter()
    for th in pool:
        th.start()
    for th in pool:
        th.join()
    t = time.perf_counter() - t0
    flops = 2.0 * m * cols * k * threads * reps
    return dict(value=flops / t * 1e-12, unit="TFLOP/s", cores=threads, kind="reference",
                sample="reference kernels/avx2.cpp shgemm_opt, C(%d x %d) -= A(%d x %d) B, FP16 operands, "
                       "FP32 C, %d threads x %d columns, %d passes: %.2f s" % (m, threads * cols, m, k,
                                                                             threads, cols, reps, t))


def bench_batch(args):
    """Batching form: 4096 seeded systems, sizes uniform in [32, 1024], 1-4
    right-hand sides; autotuned panel widths; the persistent launch replayed
    from a hipGraph against the one-launch-per-system loop."""
    import hplai
    ctx = hplai.Context(0)
    n, nrhs, seeds = hplai.batch_spec(args.batch_count, seed=2024)
    bt = hplai.Batch(ctx, n, nrhs, seeds)
    table = bt.autotune()
    for _ in range(max(args.warmup, 1)):
        bt.run("graph")
    runs = [bt.run("graph") for _ in range(args.steps)]
    _, steps, _, _ = bt.results()
    baselines = {m: bt.run(m) for m in ("per_system", "per_system_4streams", "grid_per_system")}
    base = baselines["per_system"]
    bt.close()
    ctx.close()
    t = sum(r["seconds"] for r in runs) / len(runs)
    last = runs[-1]
    # roofline of batch_kernel: HBM-bound.  Algorithmic bytes per system: the
    # LU's trailing FP32 matrix read and written at every panel step,
    # sum_k 8 (n - k nb)^2 ~ 8 n^3 / (3 nb); the generation's FP64 and FP32
    # copies, 12 n^2; one pass over the FP64 matrix per refinement residual,
    # 8 n^2 per step.  The kernel is the whole timed launch.
    bounds = (128, 256, 512)
    nbytes = 0.0
    for q in range(args.batch_count):
        nq = float(n[q])
        nb = table[sum(1 for x in bounds if n[q] >= x)]
        trail = sum(8.0 * (nq - k * nb) ** 2 for k in range(int(n[q]) // nb + 1) if nq - k * nb > 0)
        nbytes += trail + 12.0 * nq * nq + 8.0 * nq * nq * max(int(v) for v in steps[q][:int(nrhs[q])])
    achieved = nbytes / t * 1e-9
    line = {
        "metric": "batched HPL-AI solves: systems/s, one persistent hipGraph launch, 1 GPU",
        "value": args.batch_count / t, "unit": "systems/s", "n_gpus": 1, "steps": args.steps,
        "warmup": max(args.warmup, 1), "ms_per_step": t * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f16", "data": "synthetic (seeded LCG systems)",
        "config": {"workload": "%d systems, n uniform in [32, 1024], 1-4 FP64 rhs" % args.batch_count,
                   "autotuned_nb_per_bucket": table},
        "p50_us": last["p50_us"], "p99_us": last["p99_us"], "max_us": last["max_us"],
        "failed": last["failed"], "max_harness": last["max_harness"], "max_ir_steps": last["max_steps"],
        "baselines": {m: {"systems_per_s": r["systems_per_s"], "seconds": r["seconds"], "p50_us": r["p50_us"],
                          "p99_us": r["p99_us"], "failed": r["failed"]} for m, r in baselines.items()},
        "speedup_vs_per_system_loop": base["seconds"] / t,
        "speedup_vs_per_system_4streams": baselines["per_system_4streams"]["seconds"] / t,
        "speedup_vs_grid_per_system": baselines["grid_per_system"]["seconds"] / t,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_GBS, "traffic": None, "kernel": "batch_kernel",
                     "algorithmic_bytes": nbytes},
    }
    print(json.dumps(line))


def bench_autograd(args):
    """Differentiation form: X = solve(A, B), n = 8192, B 8192 x 64, loss
    sum(W * X); forward and backward timed."""
    import torch
    import hplai
    from hplai import autograd
    n, k = 8192, 64
    ctx = hplai.Context(0)
    d = torch.empty(n, dtype=torch.float64, device="cuda")
    ctx.gen_diag(n, 42, 0, n, d)
    A32 = torch.empty((n, n), dtype=torch.float32, device="cuda")
    ctx.gen_matrix(n, 42, 0, 0, n, n, None, A32, n)
    A = A32.t().double().contiguous()
    A.diagonal().copy_(d)
    g = torch.Generator(device="cpu").manual_seed(2024)
    B = torch.randn((n, k), generator=g, dtype=torch.float64).cuda()
    W = torch.randn((n, k), generator=g, dtype=torch.float64).cuda()
    fw, bw, info = [], [], {}
    for it in range(max(args.warmup, 1) + args.steps):
        Ar = A.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        X = autograd.solve(Ar, B, ctx=ctx, b=512, fmt="f16", info=info)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        (W * X).sum().backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        if it >= max(args.warmup, 1):
            fw.append(t1 - t0)
            bw.append(t2 - t1)
    At = A.clone().requires_grad_(True)
    (W * torch.linalg.solve(At, B)).sum().backward()
    rel = float((Ar.grad - At.grad).abs().max() / At.grad.abs().max())
    ctx.close()
    tf, tb = sum(fw) / len(fw), sum(bw) / len(bw)
    print(json.dumps({
        "metric": "autograd solve X = A^-1 B, n=8192, 64 rhs: forward + backward time, 1 GPU",
        "value": (tf + tb) * 1e3, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": max(args.warmup, 1),
        "ms_per_step": (tf + tb) * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f16", "data": "synthetic (seeded LCG matrix, N(0,1) B and W)",
        "config": {"workload": "n=8192, B 8192x64 FP64, FP16 Schur LU b=512, FP64 refinement"},
        "forward_ms": tf * 1e3, "backward_ms": tb * 1e3, "forward_ir_steps": info.get("forward_steps"),
        "backward_ir_steps": info.get("backward_steps"), "dA_rel_err_vs_torch_linalg_solve": rel}))


DBL_EPS = 2.220446049250313e-16


def bench_schur(args):
    """Kernel form: the Schur update alone, C -= 2^-(eL+eU) Lq Uq with C in
    FP32 (read and written), m = n = args.schur_m, K = 8192, through the C
    ABI (row-layout operands), beside hipBLASLt's GEMM of the same shape
    through torch (C = A B^T, C not read: FP8 with FP32 and with BF16 output,
    FP16 with FP16 output) as a reference point for the kernel's rate."""
    import torch
    import hplai
    m = n = args.schur_m
    K, tb = 8192, 512
    fl = 2.0 * m * n * K
    ctx = hplai.Context(0)
    g = torch.Generator(device="cuda").manual_seed(1)
    A8 = torch.randn(m, K, device="cuda", generator=g).to(torch.float8_e4m3fn)
    B8 = torch.randn(n, K, device="cuda", generator=g).to(torch.float8_e4m3fn)
    eL = torch.zeros(K // tb, m // tb, dtype=torch.int32, device="cuda")
    eU = torch.zeros(K // tb, n // tb, dtype=torch.int32, device="cuda")
    C = torch.zeros(n, m, dtype=torch.float32, device="cuda")
    one = torch.ones((), device="cuda")
    reps = max(args.steps, 3)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        ts.sort()
        return {"ms_min": ts[0], "ms_median": ts[len(ts) // 2], "tflops": fl / (ts[len(ts) // 2] * 1e-3) / 1e12}

    res = {}
    L8, U8 = A8.view(torch.uint8), B8.view(torch.uint8)
    res["schur_e4m3"] = timed(lambda: ctx.schur_update(2, m, n, K // tb, tb, L8, K, eL, m // tb, U8, K, eU,
                                                       n // tb, tb, C, m))
    res["hipblaslt_fp8_out_f32"] = timed(lambda: torch._scaled_mm(A8, B8.t(), scale_a=one, scale_b=one,
                                                                  out_dtype=torch.float32))
    res["hipblaslt_fp8_out_bf16"] = timed(lambda: torch._scaled_mm(A8, B8.t(), scale_a=one, scale_b=one,
                                                                   out_dtype=torch.bfloat16))
    del A8, B8, L8, U8
    A16 = torch.randn(m, K, device="cuda", generator=g).half()
    B16 = torch.randn(n, K, device="cuda", generator=g).half()
    res["schur_f16"] = timed(lambda: ctx.schur_update(0, m, n, K // tb, tb, A16, K, eL, m // tb, B16, K, eU,
                                                      n // tb, tb, C, m))
    res["hipblaslt_f16_out_f16"] = timed(lambda: torch.mm(A16, B16.t()))
    ctx.close()
    e4 = res["schur_e4m3"]
    print(json.dumps({
        "metric": "Schur update alone, m=n=%d, K=8192: TFLOP/s (median of %d), 1 GPU" % (m, reps),
        "value": e4["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": e4["ms_median"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "e4m3", "data": "synthetic (N(0,1) operands, zero exponents)",
        "config": {"workload": "C(%d x %d, FP32) -= L U, K = 8192" % (m, n)}, "kernels": res}))


def bench_panel(args):
    """Panel form: the kernels of one panel step of the LU, alone (b = 512):
    the diagonal-block LU, and the two panel TRSMs over m = n - b rows /
    columns (each C-ABI TRSM call also inverts the diagonal block's
    triangles).  The tail of the LU is a chain of these, one panel after the
    other; a kernel trace of this form gives each kernel's own duration."""
    import torch
    import hplai
    b, m = args.b, args.n - args.b
    ctx = hplai.Context(0)
    g = torch.Generator(device="cuda").manual_seed(3)
    A0 = (torch.rand(b, b, device="cuda", generator=g) - 0.5) + torch.eye(b, device="cuda") * b
    Ukk = A0.triu().t().contiguous()                      # column-major upper triangle
    L = torch.rand(b, m, device="cuda", generator=g) - 0.5   # column-major m x b
    U = torch.rand(m, b, device="cuda", generator=g) - 0.5   # column-major b x m
    reps = max(args.steps, 5)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1))
        ts.sort()
        return {"ms_min": ts[0], "ms_median": ts[len(ts) // 2]}

    A = A0.clone()   # factored again in place by every call (the factors stay diagonally dominant)
    res = {"getrf_b512": timed(lambda: ctx.getrf(b, A, b)),
           "trsm_l_m%d" % m: timed(lambda: ctx.trsm_lpanel(m, b, Ukk, b, L, m)),
           "trsm_u_n%d" % m: timed(lambda: ctx.trsm_upanel(b, m, Ukk, b, U, b))}
    ctx.close()
    print(json.dumps({
        "metric": "panel-step kernels alone, b=%d, m=%d: ms per call (median of %d), 1 GPU" % (b, m, reps),
        "value": res["getrf_b512"]["ms_median"], "unit": "ms", "n_gpus": 1, "steps": reps, "warmup": 1,
        "ms_per_step": res["getrf_b512"]["ms_median"], "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (uniform operands, dominant diagonal)",
        "config": {"workload": "getrf b=%d; TRSM L m x b and U b x m, m=%d" % (b, m)}, "kernels": res}))


def backward_error(harness, n):
    """Normwise backward error ||b - A x|| / (||A|| ||x|| + ||b||) (inf-norms)
    from the HPL-AI harness, which is that ratio divided by n eps / 2."""
    return harness * n * DBL_EPS / 2 if harness is not None and harness >= 0 else None


def fp64_comparison(ctx, args):
    """The all-FP64 LU (hplai.fp64_lu) at the largest order that fits in HBM
    in FP64 (n = 131072 by default: 137 GB), the mixed forms at the same n
    with the forward error of their x against the FP64 x, and the FP64 time
    extrapolated to args.n at the measured FP64 rate."""
    import torch
    from hplai import fp64_lu
    n = args.fp64_n
    forms = {}
    for fmt, outer in ((args.fmt, args.outer), (args.compare or None, args.compare_outer)):
        if not fmt:
            continue
        r = ctx.solve(n, args.b, fmt=fmt, outer=outer, lookahead=_lookahead(args), keep_factors=True)
        x = torch.from_numpy(ctx.solution_x(n))
        ctx.release()
        forms[fmt] = (r, x)
    torch.cuda.empty_cache()
    f = fp64_lu.fp64_solve(ctx, n)
    x64 = f.pop("x").cpu()
    torch.cuda.empty_cache()
    out = {"n": n, "fp64_time_s": f["time_total"], "fp64_tflops": f["tflops"], "fp64_ir_steps": f["ir_steps"],
           "fp64_backward_error": backward_error(f["harness"], n),
           "fp64_time_extrapolated_to_n%d_s" % args.n: f["time_total"] * (args.n / n) ** 3}
    for fmt, (r, x) in forms.items():
        out[fmt] = {"time_s": r["time_total"], "speedup_vs_fp64": f["time_total"] / r["time_total"],
                    "ir_steps": r["ir_steps"], "backward_error": backward_error(r["harness"], n),
                    "forward_error_vs_fp64": float((x - x64).abs().max() / x64.abs().max())}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--form", default="solve", choices=["solve", "batch", "autograd", "schur", "panel"],
                    help="solve: the metric's workload (default); batch / autograd: the other forms; "
                         "schur: the Schur kernel alone beside hipBLASLt; panel: one panel step's kernels alone")
    ap.add_argument("--schur-m", type=int, default=32768)
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=196608)
    ap.add_argument("--b", type=int, default=512)
    ap.add_argument("--fmt", default="e4m3", choices=["e4m3", "f16", "f16t"])
    ap.add_argument("--outer", type=int, default=16)
    ap.add_argument("--compare", default="f16", help="second form run beside the primary ('' = none)")
    ap.add_argument("--compare-outer", type=int, default=16)
    ap.add_argument("--split", type=int, default=0,
                    help="factor each outer panel in two parts after this many inner blocks (0 = one part)")
    ap.add_argument("--no-lookahead", action="store_true")
    ap.add_argument("--lazy", action="store_true", help="lazy initialisation (-d)")
    ap.add_argument("--trail16", action="store_true",
                    help="FP16 trailing storage of the partial sums (implies --lazy)")
    ap.add_argument("--lookahead", type=int, default=1,
                    help="1: next panel on a second stream; 2: also the next block row")
    ap.add_argument("--cpu-n", type=int, default=16384)
    ap.add_argument("--cpu-b", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-n", type=int, default=131072,
                    help="order of the all-FP64 LU comparison (0 = skip); n = 196608 needs 309 GB in FP64")
    ap.add_argument("--copy-check-bytes", type=int, default=0,
                    help="first run one device-to-device copy of this many bytes (calibrates "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE against a known traffic)")
    args = ap.parse_args()

    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    import hplai

    torch.cuda.set_device(0)
    if args.form == "batch":
        return bench_batch(args)
    if args.form == "autograd":
        return bench_autograd(args)
    if args.form == "schur":
        return bench_schur(args)
    if args.form == "panel":
        return bench_panel(args)
    if args.copy_check_bytes > 0:
        src = torch.ones(args.copy_check_bytes // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        torch.mul(src, 1.0, out=dst)   # an elementwise kernel: reads and writes the bytes once
        torch.cuda.synchronize()
        del src, dst
    ctx = hplai.Context(0)
    # load the code objects once (n = 2048) so the first timed solve does not pay for it
    ctx.solve(2048, 512 if args.b >= 512 else args.b, fmt=args.fmt, outer=1)
    torch.cuda.synchronize()

    primary = run_form(ctx, args, args.fmt, args.outer)
    compare = None
    if args.compare:
        compare = run_form(ctx, args, args.compare, args.compare_outer)
    fp64 = fp64_comparison(ctx, args) if args.fp64_n > 0 else None
    ctx.close()

    roof = roofline(primary, args.fmt, args.n)
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = cpu_baseline(args.cpu_n, args.cpu_b, threads)
        ref_schur = cpu_reference_schur(threads)
        if ref_schur is not None:
            # the dominant GPU kernel against the reference's own CPU kernel for the same update
            ref_schur["gpu_schur_tflops"] = roof["achieved"] if roof else None
            cpu["reference_schur_kernel"] = ref_schur

    line = {
        "metric": "time-to-harness<16 & TFLOP/s at n=%d: FP8-scaled vs FP16 mixed LU, 1 GPU" % args.n,
        "value": primary["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": primary["time_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "e4m3" if args.fmt == "e4m3" else "f16",
        "data": "synthetic (seeded HPL-AI LCG matrix, generated on the GPU)",
        "config": {"workload": "hpl-ai solve n=%d b=%d, Schur %s (outer %d), FP32 panel, FP64 IR%s"
                               % (args.n, args.b, args.fmt, args.outer,
                                  ", lazy init, FP16 trailing storage" if args.trail16
                                  else (", lazy init" if args.lazy else "")),
                   "n": args.n, "b": args.b, "fmt": args.fmt, "outer": args.outer,
                   "lookahead": _lookahead(args), "split": _split(args, args.outer), **_storage(args)},
        "time_to_harness_s": primary["time_s"],
        "ir_steps": primary["ir_steps"],
        "harness": primary["harness"],
        "residual": primary["residual"],
        "time_lu_s": primary["time_lu_s"],
        "time_ir_s": primary["time_ir_s"],
        "time_gen_s": primary["time_gen_s"],
        "wall_s_per_step": primary["wall_s"],
        "backward_error": backward_error(primary["harness"], args.n),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if fp64 is not None:
        line["fp64_comparison"] = fp64
        line["speedup_vs_fp64_extrapolated"] = fp64["fp64_time_extrapolated_to_n%d_s" % args.n] / primary["time_s"]
    if compare is not None:
        line["compare"] = {k: compare[k] for k in ("fmt", "outer", "split", "time_s", "tflops", "ir_steps",
                                                   "harness", "time_lu_s", "time_ir_s",
                                                   "schur_tflops")}
        line["compare"]["backward_error"] = backward_error(compare["harness"], args.n)
        line["speedup_vs_compare"] = compare["time_s"] / primary["time_s"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
