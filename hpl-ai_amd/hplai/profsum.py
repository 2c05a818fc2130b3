"""Summaries of rocprofv3 output for the solver's kernels.

    python -m hplai.profsum stats  KERNEL_STATS_CSV [--top N]
    python -m hplai.profsum traffic --fetch FETCH_CSV --write WRITE_CSV
                                    --kernel NAME_SUBSTRING --copy-bytes BYTES
    python -m hplai.profsum launches KERNEL_TRACE_CSV --kernel NAME_SUBSTRING
                                    [--first N]
    python -m hplai.profsum hbm --fetch FETCH_CSV --write WRITE_CSV
                                --copy-bytes BYTES --kernel SUBSTRING [--kernel ...]

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
    a = ap.parse_args(argv)
    if a.cmd == "stats":
        stats(a.csv, a.top)
    elif a.cmd == "launches":
        launches(a.csv, a.kernel, a.first)
    elif a.cmd == "hbm":
        hbm(a.fetch, a.write, a.kernel, a.copy_bytes, a.copy_kernel)
    else:
        traffic(a.fetch, a.write, a.kernel, a.copy_bytes, a.copy_kernel)


if __name__ == "__main__":
    sys.exit(main())
