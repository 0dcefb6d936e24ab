"""HBM traffic per dispatch from two rocprofv3 counter passes.

    rocprofv3 --pmc FETCH_SIZE --output-format csv -d F -o run -- python3 bench.py --algos A --copy-calibration-mib 1024 ...
    rocprofv3 --pmc WRITE_SIZE --output-format csv -d W -o run -- python3 bench.py --algos A --copy-calibration-mib 1024 ...
    python -m starplat_amd.pmc_summary --fetch F/run_counter_collection.csv --write W/run_counter_collection.csv \
        --copy-mib 1024 --workload sssp-rmat22 --kernels sssp_relax

Both counters are reported in KiB per dispatch. Each is scaled by what it
reads for the calibration copy that bench.py runs first (a device-to-device
hipMemcpy of a known size, the first `__amd_rocclr_copyBuffer` dispatch of the
run): fetch factor = copy bytes / FETCH bytes, write factor = copy bytes /
WRITE bytes. The HBM bytes of one dispatch of the named kernels are the scaled
sums over all their dispatches divided by the number of dispatches.

Prints one markdown table row; bench.py reads rows of this form from the
profiles directory to fill `roofline.traffic`.

    python -m starplat_amd.pmc_summary --kernel-stats K/run_kernel_stats.csv [--top 8]

formats the per-kernel summary of a `rocprofv3 --kernel-trace --stats` run.
"""
import argparse
import csv
import re
import sys

_NAME = re.compile(r"spm::(\w+)")
_COPY = "__amd_rocclr_copyBuffer"

HEADER = ("| workload | kernels | dispatches | FETCH_SIZE KiB/dispatch | WRITE_SIZE KiB/dispatch "
          "| fetch factor | write factor | HBM bytes/dispatch |\n|---|---|---|---|---|---|---|---|")


def short_name(kernel_name):
    m = _NAME.search(kernel_name)
    return m.group(1) if m else kernel_name


def read_pass(path, counter):
    """(rows in dispatch order, KiB of the calibration copy) of one counter pass."""
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    copy = next((v for _, k, v in rows if k.startswith(_COPY)), None)
    if copy is None or copy <= 0:
        raise SystemExit(f"{path}: no calibration copy dispatch ({_COPY}) with a non-zero {counter}")
    return rows, copy


def kernel_sum(rows, kernels):
    n, total = 0, 0.0
    for _, k, v in rows:
        if short_name(k) in kernels:
            n += 1
            total += v
    return n, total


def summarize(fetch_csv, write_csv, copy_mib, kernels):
    copy_bytes = float(copy_mib) * (1 << 20)
    frows, fcopy = read_pass(fetch_csv, "FETCH_SIZE")
    wrows, wcopy = read_pass(write_csv, "WRITE_SIZE")
    nf, fk = kernel_sum(frows, kernels)
    nw, wk = kernel_sum(wrows, kernels)
    if nf == 0 or nf != nw:
        raise SystemExit(f"dispatch counts differ or are zero: FETCH pass {nf}, WRITE pass {nw}")
    ff = copy_bytes / (fcopy * 1024.0)
    wf = copy_bytes / (wcopy * 1024.0)
    per = (fk * 1024.0 * ff + wk * 1024.0 * wf) / nf
    return {"dispatches": nf, "fetch_kib": fk / nf, "write_kib": wk / nf, "fetch_factor": ff,
            "write_factor": wf, "bytes": per}


def row(workload, kernels, s):
    return (f"| {workload} | {','.join(kernels)} | {s['dispatches']} | {s['fetch_kib']:.1f} | {s['write_kib']:.1f} "
            f"| {s['fetch_factor']:.4f} | {s['write_factor']:.4f} | {s['bytes']:.0f} |")


def parse_rows(text):
    """{(workload, kernels tuple): bytes per dispatch} from summary tables."""
    out = {}
    for line in text.splitlines():
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) != 8 or not cells[2].isdigit():
            continue
        try:
            out[(cells[0], tuple(cells[1].split(",")))] = float(cells[7])
        except ValueError:
            continue
    return out


def stats_table(path, top=8):
    """Markdown table of the `top` kernels by total time in a rocprofv3
    kernel_stats CSV."""
    with open(path, newline="") as f:
        rows = list(csv.DictReader(f))
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    out = ["| kernel | calls | total ms | average us | min us | max us | % |", "|---|---|---|---|---|---|---|"]
    for r in rows[:top]:
        out.append(f"| {short_name(r['Name'])} | {r['Calls']} | {float(r['TotalDurationNs']) / 1e6:.3f} "
                   f"| {float(r['AverageNs']) / 1e3:.2f} | {float(r['MinNs']) / 1e3:.2f} "
                   f"| {float(r['MaxNs']) / 1e3:.2f} | {float(r['Percentage']):.2f} |")
    return "\n".join(out)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--kernel-stats", help="format a rocprofv3 kernel_stats CSV and exit")
    ap.add_argument("--top", type=int, default=8)
    ap.add_argument("--fetch")
    ap.add_argument("--write")
    ap.add_argument("--copy-mib", type=float)
    ap.add_argument("--workload")
    ap.add_argument("--kernels", help="comma-separated kernel names (without namespace)")
    ap.add_argument("--header", action="store_true")
    a = ap.parse_args(argv)
    if a.kernel_stats:
        print(stats_table(a.kernel_stats, a.top))
        return 0
    if not (a.fetch and a.write and a.copy_mib and a.workload and a.kernels):
        ap.error("--fetch, --write, --copy-mib, --workload and --kernels are required")
    kernels = tuple(a.kernels.split(","))
    s = summarize(a.fetch, a.write, a.copy_mib, kernels)
    if a.header:
        print(HEADER)
    print(row(a.workload, kernels, s))
    return 0


if __name__ == "__main__":
    sys.exit(main())
