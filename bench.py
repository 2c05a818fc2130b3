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


def run_form(ctx, args, fmt, outer):
    res = None
    for _ in range(args.warmup):
        ctx.solve(args.n, args.b, fmt=fmt, outer=outer, lookahead=not args.no_lookahead, maxit=50)
    import torch
    times, results = [], []
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        r = ctx.solve(args.n, args.b, fmt=fmt, outer=outer, lookahead=not args.no_lookahead,
                      maxit=50, profile=True)
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
            "schur_large_ms", "schur_large_bytes")}
    return dict(fmt=fmt, outer=outer, time_s=t, wall_s=wall, tflops=hpl_flops(args.n) / t * 1e-12,
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
    one exists (FETCH_SIZE + WRITE_SIZE, calibrated against a copy kernel)."""
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
    return {"bound": "mfma", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
            "traffic": traffic,
            "kernel": "schur_dma_kernel<%s> (256x256 streamed Schur update)" % fmt,
            "launches": nl, "flops_per_launch": flops_per, "ms_per_launch": ms_per,
            "algorithmic_bytes_per_launch": p["schur_large_bytes"] / nl,
            "all_schur_launches": {"launches": p["schur_launches"], "tflops": p["schur_tflops"]}}


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
    return dict(value=hpl_flops(n) / t * 1e-12, unit="TFLOP/s", cores=threads, kind="port",
                sample="oracle LU+IR (FP16 Schur, FP32 accumulate, FP64 IR) at n=%d b=%d: "
                       "%.2f s, harness %.4f" % (n, b, t, r["final_harness"]))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--n", type=int, default=196608)
    ap.add_argument("--b", type=int, default=512)
    ap.add_argument("--fmt", default="e4m3", choices=["e4m3", "f16", "f16t"])
    ap.add_argument("--outer", type=int, default=8)
    ap.add_argument("--compare", default="f16", help="second form run beside the primary ('' = none)")
    ap.add_argument("--compare-outer", type=int, default=8)
    ap.add_argument("--no-lookahead", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=16384)
    ap.add_argument("--cpu-b", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
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
    ctx.close()

    roof = roofline(primary, args.fmt, args.n)
    cpu = None
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        cpu = cpu_baseline(args.cpu_n, args.cpu_b, threads)

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
        "config": {"workload": "hpl-ai solve n=%d b=%d, Schur %s (outer %d), FP32 panel, FP64 IR"
                               % (args.n, args.b, args.fmt, args.outer),
                   "n": args.n, "b": args.b, "fmt": args.fmt, "outer": args.outer,
                   "lookahead": not args.no_lookahead},
        "time_to_harness_s": primary["time_s"],
        "ir_steps": primary["ir_steps"],
        "harness": primary["harness"],
        "residual": primary["residual"],
        "time_lu_s": primary["time_lu_s"],
        "time_ir_s": primary["time_ir_s"],
        "time_gen_s": primary["time_gen_s"],
        "wall_s_per_step": primary["wall_s"],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if compare is not None:
        line["compare"] = {k: compare[k] for k in ("fmt", "outer", "time_s", "tflops", "ir_steps",
                                                   "harness", "time_lu_s", "time_ir_s",
                                                   "schur_tflops")}
        line["speedup_vs_compare"] = compare["time_s"] / primary["time_s"]
    print(json.dumps(line))


if __name__ == "__main__":
    main()
