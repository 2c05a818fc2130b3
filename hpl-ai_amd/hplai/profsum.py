"""Summaries of rocprofv3 output for the solver's kernels.

    python -m hplai.profsum stats  KERNEL_STATS_CSV [--top N]
    python -m hplai.profsum traffic --fetch FETCH_CSV --write WRITE_CSV
                                    --kernel NAME_SUBSTRING --copy-bytes BYTES
    python -m hplai.profsum launches KERNEL_TRACE_CSV --kernel NAME_SUBSTRING
                                    [--first N]
    python -m hplai.profsum hbm --fetch FETCH_CSV --write WRITE_CSV
                                --copy-bytes BYTES --kernel SUBSTRING [--kernel ...]
    python -m hplai.profsum counters COUNTER_CSV [COUNTER_CSV ...]
                                --kernel SUBSTRING [--kernel ...] [--simds 1024]

`stats` reads the `*_kernel_stats.csv` written by `rocprofv3 --kernel-trace
--stats --output-format csv` and prints one line per kernel (calls, total and
average duration, share of GPU time).

`launches` reads the `*_kernel_trace.csv` of the same run and prints the
count and mean duration of the first N launches (in dispatch order) of the
kernel whose name contains NAME_SUBSTRING; with N = the launches of one solve
this is the figure to set beside the bench line's live per-launch time when
the run holds more than one solve.

`traffic` reads two `*_counter_collection.csv` files written by
`rocprofv3 --pmc FETCH_SIZE` and `rocprofv3 --pmc WRITE_SIZE` over the same
command.  The command must also run one plain device-to-device copy of
BYTES bytes (bench.py --copy-check-bytes) so that the counters can be
calibrated: the factor is what FETCH_SIZE (WRITE_SIZE) reads for that copy
(summed over its dispatches) divided by BYTES.  The HBM bytes per launch of the kernel whose name contains
NAME_SUBSTRING are then FETCH/factor_f + WRITE/factor_w, averaged over its
launches.

`hbm` applies the same calibration to several kernels and divides each
one's HBM bytes by its summed dispatch durations from the FETCH_SIZE run
(counter collection runs the dispatches one at a time, so these are
standalone durations): launches, HBM bytes per launch, mean duration and
the achieved HBM rate in GB/s.

`counters` reads any `*_counter_collection.csv` files (one rocprofv3 --pmc
pass each) and prints, per kernel, every counter summed over the kernel's
dispatches (each dispatch's value summed over its instances, GRBM_GUI_ACTIVE
summed over the XCDs, so GRBM_GUI_ACTIVE / XCDs is the dispatch's length in
shader clocks) with the mean dispatch duration, then the ratios they support:
the shader clock = GRBM_GUI_ACTIVE / XCDs / duration; MFMA busy =
SQ_VALU_MFMA_BUSY_CYCLES / (GRBM_GUI_ACTIVE / XCDs * SIMDs); F8 MFMA rate = 512 SQ_INSTS_VALU_MFMA_MOPS_F8 /
duration; LDS bank-conflict cycles per LDS-active cycle; L2 hit rate
TCC_HIT / (TCC_HIT + TCC_MISS).
"""
import argparse
import csv
import sys
from collections import defaultdict


def _rows(path):
    with open(path, newline="") as f:
        return list(csv.DictReader(f))


def stats(path, top):
    rows = _rows(path)
    total = sum(float(r["TotalDurationNs"]) for r in rows)
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    print("%-70s %8s %12s %12s %7s" % ("kernel", "calls", "total_ms", "avg_us", "share"))
    for r in rows[:top]:
        name = r["Name"]
        if len(name) > 70:
            name = name[:67] + "..."
        print("%-70s %8d %12.3f %12.3f %6.2f%%" % (
            name, int(r["Calls"]), float(r["TotalDurationNs"]) * 1e-6,
            float(r["AverageNs"]) * 1e-3, 100.0 * float(r["TotalDurationNs"]) / total))


def _per_dispatch(path, counter):
    """{dispatch_id: (kernel_name, value in bytes)} for one counter (KiB in the file)."""
    out = {}
    acc = defaultdict(float)
    names = {}
    for r in _rows(path):
        if r.get("Counter_Name") != counter:
            continue
        d = int(r["Dispatch_Id"])
        acc[d] += float(r["Counter_Value"])
        names[d] = r["Kernel_Name"]
    for d, v in acc.items():
        out[d] = (names[d], v * 1024.0)
    return out


def _copy_factor(per, copy_bytes, name_hint):
    """Counter reading of the calibration copy divided by its bytes.

    The copy of copy_bytes bytes may run as several dispatches (torch splits
    elementwise kernels whose byte offsets exceed 32 bits), so the readings of
    all dispatches of the kernel whose name contains name_hint are summed."""
    got = [v for name, v in per.values() if name_hint in name]
    if not got:
        raise SystemExit("no calibration kernel matching %r found" % name_hint)
    return sum(got) / copy_bytes


def traffic(fetch, write, kernel, copy_bytes, copy_kernel="MulFunctor"):
    pf = _per_dispatch(fetch, "FETCH_SIZE")
    pw = _per_dispatch(write, "WRITE_SIZE")
    ff = _copy_factor(pf, copy_bytes, copy_kernel)
    fw = _copy_factor(pw, copy_bytes, copy_kernel)
    kf = [v for name, v in pf.values() if kernel in name]
    kw = [v for name, v in pw.values() if kernel in name]
    if not kf or not kw:
        raise SystemExit("kernel %r not found" % kernel)
    fetch_per = sum(kf) / len(kf) / ff
    write_per = sum(kw) / len(kw) / fw
    print("kernel %s" % kernel)
    print("launches_fetch_run %d" % len(kf))
    print("launches_write_run %d" % len(kw))
    print("copy_calibration_bytes %d" % copy_bytes)
    print("fetch_factor %.4f" % ff)
    print("write_factor %.4f" % fw)
    print("fetch_bytes_per_launch %.6g" % fetch_per)
    print("write_bytes_per_launch %.6g" % write_per)
    print("hbm_bytes_per_launch %.6g" % (fetch_per + write_per))


def _durations(path, counter):
    """{dispatch_id: duration in ns} from the rows of one counter."""
    out = {}
    for r in _rows(path):
        if r.get("Counter_Name") == counter:
            out[int(r["Dispatch_Id"])] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    return out


def hbm(fetch, write, kernels, copy_bytes, copy_kernel="MulFunctor"):
    pf = _per_dispatch(fetch, "FETCH_SIZE")
    pw = _per_dispatch(write, "WRITE_SIZE")
    ff = _copy_factor(pf, copy_bytes, copy_kernel)
    fw = _copy_factor(pw, copy_bytes, copy_kernel)
    dur = _durations(fetch, "FETCH_SIZE")
    print("fetch_factor %.4f write_factor %.4f (copy of %d bytes)" % (ff, fw, copy_bytes))
    print("%-44s %8s %14s %12s %10s" % ("kernel", "launches", "hbm_B/launch", "mean_us", "GB/s"))
    for k in kernels:
        ids = [d for d, (name, _) in pf.items() if k in name]
        wv = [v for name, v in pw.values() if k in name]
        if not ids or not wv:
            raise SystemExit("kernel %r not found" % k)
        fbytes = sum(pf[d][1] for d in ids) / ff
        wbytes = sum(wv) / fw * len(ids) / len(wv)   # the write run's launches, scaled to the same count
        ns = sum(dur[d] for d in ids)
        per = (fbytes + wbytes) / len(ids)
        print("%-44s %8d %14.6g %12.2f %10.1f" % (k[:44], len(ids), per, ns / len(ids) * 1e-3,
                                                  (fbytes + wbytes) / ns if ns > 0 else 0.0))


def _counter_table(paths):
    """{(dispatch_id, kernel_name): {counter: value}} and durations (ns)."""
    vals = defaultdict(lambda: defaultdict(float))
    dur = {}
    for path in paths:
        for r in _rows(path):
            key = (int(r["Dispatch_Id"]), r["Kernel_Name"], path)
            c, v = r["Counter_Name"], float(r["Counter_Value"])
            vals[key][c] += v
            dur[key] = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
    return vals, dur


def counters(paths, kernels, simds, xcds=8):
    vals, dur = _counter_table(paths)
    for k in kernels:
        keys = [key for key in vals if k in key[1]]
        if not keys:
            raise SystemExit("kernel %r not found" % k)
        tot = defaultdict(float)
        runs = defaultdict(set)
        for key in keys:
            for c, v in vals[key].items():
                tot[c] += v
                runs[c].add(key[2])
        # dispatches per pass (every pass runs the same command)
        per_pass = len(keys) / max(1, len({key[2] for key in keys}))
        ns = sum(dur[key] for key in keys) / max(1, len({key[2] for key in keys}))
        print("kernel %s" % k)
        print("dispatches_per_pass %d" % per_pass)
        print("mean_duration_us %.3f" % (ns / per_pass * 1e-3))
        for c in sorted(tot):
            print("%-28s %.6g per dispatch" % (c, tot[c] / len(runs[c]) / per_pass))
        per = {c: tot[c] / len(runs[c]) for c in tot}   # one pass's sum over the kernel's dispatches
        if "GRBM_GUI_ACTIVE" in per and ns > 0:
            # GRBM_GUI_ACTIVE comes summed over the XCDs: cycles per XCD over the duration is the clock
            print("sclk_mhz %.0f" % (per["GRBM_GUI_ACTIVE"] / xcds / ns * 1e3))
        if "SQ_VALU_MFMA_BUSY_CYCLES" in per and "GRBM_GUI_ACTIVE" in per:
            print("mfma_busy_pct %.1f" % (100.0 * per["SQ_VALU_MFMA_BUSY_CYCLES"] /
                                          (per["GRBM_GUI_ACTIVE"] / xcds * simds)))
        if "SQ_INSTS_VALU_MFMA_MOPS_F8" in per and ns > 0:
            print("f8_mfma_tflops %.6g" % (512.0 * per["SQ_INSTS_VALU_MFMA_MOPS_F8"] / ns * 1e-3))
        if "SQ_INSTS_VALU_MFMA_MOPS_F16" in per and ns > 0:
            print("f16_mfma_tflops %.6g" % (512.0 * per["SQ_INSTS_VALU_MFMA_MOPS_F16"] / ns * 1e-3))
        if "SQ_LDS_BANK_CONFLICT" in per and per.get("SQ_LDS_IDX_ACTIVE", 0) > 0:
            print("lds_bank_conflict_per_active_cycle %.4f" % (per["SQ_LDS_BANK_CONFLICT"] /
                                                              per["SQ_LDS_IDX_ACTIVE"]))
        for hit, miss in (("TCC_HIT", "TCC_MISS"), ("TCC_HIT_sum", "TCC_MISS_sum")):
            if hit in per and miss in per:
                print("l2_hit_pct %.1f" % (100.0 * per[hit] / (per[hit] + per[miss])))
        print()


def launches(path, kernel, first):
    rows = [r for r in _rows(path) if kernel in r["Kernel_Name"]]
    rows.sort(key=lambda r: int(r["Start_Timestamp"]))
    if first > 0:
        rows = rows[:first]
    if not rows:
        raise SystemExit("kernel %r not found" % kernel)
    durs = [int(r["End_Timestamp"]) - int(r["Start_Timestamp"]) for r in rows]
    print("kernel %s" % kernel)
    print("launches %d" % len(durs))
    print("mean_ms %.4f" % (sum(durs) / len(durs) * 1e-6))
    print("total_ms %.3f" % (sum(durs) * 1e-6))


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m hplai.profsum")
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    s.add_argument("--top", type=int, default=25)
    t = sub.add_parser("traffic")
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--kernel", required=True)
    t.add_argument("--copy-bytes", type=int, required=True)
    t.add_argument("--copy-kernel", default="MulFunctor",
                   help="name substring of the calibration copy's kernel (bench.py runs torch.mul)")
    g = sub.add_parser("launches")
    g.add_argument("csv")
    g.add_argument("--kernel", required=True)
    g.add_argument("--first", type=int, default=0)
    h = sub.add_parser("hbm")
    h.add_argument("--fetch", required=True)
    h.add_argument("--write", required=True)
    h.add_argument("--kernel", action="append", required=True)
    h.add_argument("--copy-bytes", type=int, required=True)
    h.add_argument("--copy-kernel", default="MulFunctor")
    c = sub.add_parser("counters")
    c.add_argument("csv", nargs="+")
    c.add_argument("--kernel", action="append", required=True)
    c.add_argument("--simds", type=int, default=1024, help="SIMDs on the device (MI355X: 256 CUs x 4)")
    c.add_argument("--xcds", type=int, default=8, help="XCDs on the device (GRBM_GUI_ACTIVE is summed over them)")
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        stats(a.csv, a.top)
    elif a.cmd == "launches":
        launches(a.csv, a.kernel, a.first)
    elif a.cmd == "counters":
        counters(a.csv, a.kernel, a.simds, a.xcds)
    elif a.cmd == "hbm":
        hbm(a.fetch, a.write, a.kernel, a.copy_bytes, a.copy_kernel)
    else:
        traffic(a.fetch, a.write, a.kernel, a.copy_bytes, a.copy_kernel)


if __name__ == "__main__":
    sys.exit(main())
