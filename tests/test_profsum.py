"""CPU tests of the rocprofv3 summary helper (hpl-ai_amd/hplai/profsum.py):
the traffic figure divides the kernel's counter readings by the calibration
copy's reading per byte, summed over all dispatches of the copy."""
import csv
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "hpl-ai_amd"))

from hplai import profsum  # noqa: E402

FIELDS = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for d, name, kib in rows:
            w.writerow({"Dispatch_Id": d, "Kernel_Name": name, "Counter_Name": counter, "Counter_Value": kib})


def test_traffic_sums_split_calibration_copy(tmp_path, capsys):
    gib = 1 << 30
    mul = "vectorized_elementwise_kernel<4, MulFunctor<float>>"
    # a 4 GiB copy split into two dispatches; FETCH_SIZE reads half of what is read,
    # WRITE_SIZE all of it
    fetch = [(1, mul, gib // 1024), (2, mul, gib // 1024),
             (3, "schur_dma_kernel<2, 2, true>", 3 * gib // 1024),
             (4, "schur_dma_kernel<2, 2, true>", 5 * gib // 1024),
             (5, "fill", 7)]
    write = [(1, mul, 2 * gib // 1024), (2, mul, 2 * gib // 1024),
             (3, "schur_dma_kernel<2, 2, true>", gib // 1024),
             (4, "schur_dma_kernel<2, 2, true>", gib // 1024)]
    pf, pw = tmp_path / "f.csv", tmp_path / "w.csv"
    _write(pf, "FETCH_SIZE", fetch)
    _write(pw, "WRITE_SIZE", write)
    profsum.traffic(str(pf), str(pw), "schur_dma_kernel<2, 2, true>", 4 * gib)
    out = dict(line.split(None, 1) for line in capsys.readouterr().out.strip().splitlines())
    assert float(out["fetch_factor"]) == 0.5
    assert float(out["write_factor"]) == 1.0
    # mean fetch reading 4 GiB / 0.5 + mean write reading 1 GiB / 1.0
    assert abs(float(out["hbm_bytes_per_launch"]) / (9 * gib) - 1.0) < 1e-5


def test_hbm_rate_uses_calibrated_bytes_and_fetch_run_durations(tmp_path, capsys):
    gib = 1 << 30
    mul = "vectorized_elementwise_kernel<4, MulFunctor<float>>"
    fields = FIELDS + ["Start_Timestamp", "End_Timestamp"]

    def write(path, counter, rows):
        with open(path, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=fields)
            w.writeheader()
            for d, name, kib, t0, t1 in rows:
                w.writerow({"Dispatch_Id": d, "Kernel_Name": name, "Counter_Name": counter,
                            "Counter_Value": kib, "Start_Timestamp": t0, "End_Timestamp": t1})

    # calibration: 4 GiB read (FETCH reads half) and written (WRITE reads all)
    fetch = [(1, mul, 2 * gib // 1024, 0, 10), (2, "quant_l16_kernel<2, true>", gib // 1024, 100, 100 + 1000),
             (3, "quant_l16_kernel<2, true>", gib // 1024, 2000, 2000 + 3000)]
    wr = [(1, mul, 4 * gib // 1024, 0, 10), (2, "quant_l16_kernel<2, true>", gib // 2048, 0, 1),
          (3, "quant_l16_kernel<2, true>", gib // 2048, 0, 1)]
    pf, pw = tmp_path / "f.csv", tmp_path / "w.csv"
    write(pf, "FETCH_SIZE", fetch)
    write(pw, "WRITE_SIZE", wr)
    profsum.hbm(str(pf), str(pw), ["quant_l16_kernel<2, true>"], 4 * gib)
    line = capsys.readouterr().out.strip().splitlines()[-1].split()
    launches, per, mean_us, rate = int(line[-4]), float(line[-3]), float(line[-2]), float(line[-1])
    # per launch: 1 GiB fetched reading / 0.5 + 0.5 GiB written = 2.5 GiB; 4000 ns over two launches
    assert launches == 2
    assert abs(per / (2.5 * gib) - 1) < 1e-5
    assert abs(mean_us - 2.0) < 1e-9
    assert abs(rate / (2 * 2.5 * gib / 4000) - 1) < 1e-5


def test_counters_ratios(tmp_path, capsys):
    """`counters`: per-kernel sums over dispatches (instances summed, also
    GRBM_GUI_ACTIVE, which counts once per XCD), the clock from GUI-active
    cycles per XCD, MFMA busy over those cycles times SIMDs, F8 rate from the
    MOPS count, L2 hit rate."""
    fields = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "Start_Timestamp", "End_Timestamp"]
    k = "schur_dma_kernel<2, 2, true>"
    p1 = tmp_path / "a.csv"
    with open(p1, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for d in (1, 2):
            for inst in range(2):   # two instances (XCDs): summed
                w.writerow(dict(Dispatch_Id=d, Kernel_Name=k, Counter_Name="SQ_VALU_MFMA_BUSY_CYCLES",
                                Counter_Value=512 * 500, Start_Timestamp=0, End_Timestamp=1000))
                w.writerow(dict(Dispatch_Id=d, Kernel_Name=k, Counter_Name="GRBM_GUI_ACTIVE",
                                Counter_Value=1000, Start_Timestamp=0, End_Timestamp=1000))
                w.writerow(dict(Dispatch_Id=d, Kernel_Name=k, Counter_Name="SQ_INSTS_VALU_MFMA_MOPS_F8",
                                Counter_Value=1000, Start_Timestamp=0, End_Timestamp=1000))
    p2 = tmp_path / "b.csv"
    with open(p2, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=fields)
        w.writeheader()
        for d in (1, 2):
            w.writerow(dict(Dispatch_Id=d, Kernel_Name=k, Counter_Name="TCC_HIT", Counter_Value=3,
                            Start_Timestamp=0, End_Timestamp=1000))
            w.writerow(dict(Dispatch_Id=d, Kernel_Name=k, Counter_Name="TCC_MISS", Counter_Value=1,
                            Start_Timestamp=0, End_Timestamp=1000))
    profsum.main(["counters", str(p1), str(p2), "--kernel", "schur_dma", "--simds", "1024", "--xcds", "2"])
    out = capsys.readouterr().out
    vals = dict(ln.split()[:2] for ln in out.splitlines() if ln and not ln.startswith("kernel"))
    assert vals["dispatches_per_pass"] == "2"
    # 1000 cycles per XCD in 1000 ns
    assert abs(float(vals["sclk_mhz"]) - 1000.0) < 1e-9
    # 2 x 512 x 500 busy cycles per dispatch over 1000 x 1024 SIMD-cycles
    assert abs(float(vals["mfma_busy_pct"]) - 50.0) < 1e-9
    # 2 instances x 1000 MOPS x 512 flops per dispatch over 1000 ns
    assert abs(float(vals["f8_mfma_tflops"]) - 2 * 1000 * 512 / 1000 * 1e-3) < 1e-9
    assert abs(float(vals["l2_hit_pct"]) - 75.0) < 1e-9
