"""The rocprofv3 counter summarizer that fills bench.py's roofline.traffic."""
import csv

from starplat_amd import pmc_summary as P

FIELDS = ["Correlation_Id", "Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"]


def _write(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=FIELDS)
        w.writeheader()
        for i, (name, value) in enumerate(rows, 1):
            w.writerow({"Correlation_Id": i, "Dispatch_Id": i, "Kernel_Name": name, "Counter_Name": counter,
                        "Counter_Value": value})


def test_calibrated_bytes_per_dispatch(tmp_path):
    # A 1 MiB copy that FETCH_SIZE reports as half (512 KiB) and WRITE_SIZE as
    # is; two dispatches of the kernel of interest, one of another kernel.
    f, w = tmp_path / "f.csv", tmp_path / "w.csv"
    _write(f, "FETCH_SIZE", [("__amd_rocclr_fillBufferAligned", 0.0), ("__amd_rocclr_copyBuffer", 512.0),
                             ("void spm::pr_iter<8>(int const*)", 100.0), ("spm::pr_check(float const*)", 7.0),
                             ("void spm::pr_iter<8>(int const*)", 300.0)])
    _write(w, "WRITE_SIZE", [("__amd_rocclr_fillBufferAligned", 1024.0), ("__amd_rocclr_copyBuffer", 1024.0),
                             ("void spm::pr_iter<8>(int const*)", 10.0), ("spm::pr_check(float const*)", 0.0),
                             ("void spm::pr_iter<8>(int const*)", 30.0)])
    s = P.summarize(str(f), str(w), 1.0, ("pr_iter",))
    assert s["dispatches"] == 2
    assert abs(s["fetch_factor"] - 2.0) < 1e-12 and abs(s["write_factor"] - 1.0) < 1e-12
    # (100+300) KiB * 2 + (10+30) KiB * 1, over two dispatches
    assert abs(s["bytes"] - (800 + 40) * 1024 / 2) < 1e-6
    row = P.row("pr-test", ("pr_iter",), s)
    assert P.parse_rows(P.HEADER + "\n" + row) == {("pr-test", ("pr_iter",)): round(s["bytes"])}


def test_kernel_name_and_stats_rows_not_confused():
    assert P.short_name("void spm::tc_count<true>(long, int const*)") == "tc_count"
    assert P.short_name("__amd_rocclr_copyBuffer") == "__amd_rocclr_copyBuffer"
    stats_row = "| sssp_relax | 96 | 13.225 | 137.76 | 3.36 | 662.81 | 88.26 |"
    assert P.parse_rows(stats_row) == {}


def test_report_recomputes_the_fraction_from_the_trace(tmp_path):
    # Two kernels of the leg (400 and 100 dispatches, 2 ms and 0.5 ms in all:
    # 5 us average) and one outside it; bench.py said 4 MB per launch.
    stats = tmp_path / "s.csv"
    with open(stats, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Name", "Calls", "TotalDurationNs", "AverageNs", "Percentage",
                                          "MinNs", "MaxNs"])
        w.writeheader()
        w.writerow({"Name": "void spm::bc_sigma(int)", "Calls": 400, "TotalDurationNs": 2e6, "AverageNs": 5e3,
                    "Percentage": 70, "MinNs": 1e3, "MaxNs": 9e3})
        w.writerow({"Name": "spm::bc_back(int)", "Calls": 100, "TotalDurationNs": 0.5e6, "AverageNs": 5e3,
                    "Percentage": 20, "MinNs": 1e3, "MaxNs": 9e3})
        w.writerow({"Name": "spm::bc_zero(int)", "Calls": 7, "TotalDurationNs": 1e6, "AverageNs": 1e5,
                    "Percentage": 10, "MinNs": 1e3, "MaxNs": 9e3})
    bench = tmp_path / "b.json"
    bench.write_text('[bench] log line\n{"roofline": {"kernels": ["sssp_relax"]}, "per_algo": {"bc": {"roofline": '
                     '{"kernels": ["bc_sigma", "bc_back"], "bytes_per_launch": 4000000, "peak": 8000.0, '
                     '"kernel_ms_per_launch": 0.005, "launches_per_step": 5, "achieved": 800.0, "frac": 0.1, '
                     '"traffic": null}}}}\n')
    text = P.report("t", ["cmd"], str(stats), str(bench), "bc")
    # 4 MB / 5 us = 800 GB/s = 0.1 of 8 TB/s, the same as bench.py's
    assert "500 dispatches, 5.00 us average" in text
    assert "frac 0.1000 (1.000 x bench.py's)" in text
