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

formats the per-kernel summary of a `rocprofv3 --kernel-trace --stats` run, and

    python -m starplat_amd.pmc_summary --report --kernel-stats K/run_kernel_stats.csv --bench B.json \
        --leg pr_c0 --title "..." --command "..." [--command "..."]

writes the profile summary of one bench leg: the kernel table, bench.py's
live roofline figures for that leg (B.json: the same bench.py command run
without the tracer) and the same fraction recomputed from the trace's
average dispatch time of the leg's kernels.
"""
import argparse
import csv
import json
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


def stats_totals(path, kernels):
    """(dispatches, total ns) of the named kernels in a kernel_stats CSV."""
    n, ns = 0, 0.0
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            if short_name(r["Name"]) in kernels:
                n += int(r["Calls"])
                ns += float(r["TotalDurationNs"])
    return n, ns


def bench_roofline(bench_json, leg):
    """The roofline object of one leg of a bench.py JSON line (the headline
    leg's sits at the top level)."""
    with open(bench_json) as f:
        line = [l for l in f.read().splitlines() if l.strip().startswith("{")][-1]
    out = json.loads(line)
    per = out.get("per_algo", {}).get(leg, {})
    return per.get("roofline") or out["roofline"]


def report(title, commands, stats_csv, bench_json, leg, top=8, traffic_md=None):
    """The markdown summary of one profiled bench.py leg: the kernel table of
    the traced run, bench.py's live roofline for `leg` from an untraced run of
    the same command (bench_json), and the same fraction recomputed from the
    trace's own average dispatch time."""
    rf = bench_roofline(bench_json, leg)
    kernels = tuple(rf["kernels"])
    n, ns = stats_totals(stats_csv, kernels)
    if n == 0:
        raise SystemExit(f"{stats_csv}: no dispatches of {','.join(kernels)}")
    avg_us = ns / n / 1e3
    bpl = float(rf["bytes_per_launch"])
    gbps = bpl / (avg_us * 1e-6) / 1e9
    frac = gbps / float(rf["peak"])
    lines = [f"# {title}", "", "One MI355X, ROCm 7.2. Commands (from the repository root, after "
             "`cd /tmp && export TMPDIR=/tmp` and back):", ""]
    lines += [f"    {c}" for c in commands]
    lines += ["", "The traced run includes the first call on the graph, warm-up and timed steps (graphs "
              "replayed) and bench.py's profiled pass (direct launches), so call counts cover several runs.",
              "", stats_table(stats_csv, top), "",
              "bench.py's live measurement comes from the same bench.py command run without the tracer "
              "(under rocprofv3 the dispatch-stamped events of microsecond kernels read high).", "",
              f"bench.py, untraced ({', '.join(kernels)}): "
              f"{float(rf['kernel_ms_per_launch']) * 1e3:.2f} us per launch, {rf['launches_per_step']} launches "
              f"per run, {bpl:.0f} algorithmic bytes per launch -> {rf['achieved']} GB/s, frac {rf['frac']}.",
              "",
              f"Recomputed from this trace: {n} dispatches, {avg_us:.2f} us average -> "
              f"{bpl:.0f} / {avg_us:.2f} us = {gbps:.1f} GB/s, frac {frac:.4f} "
              f"({frac / float(rf['frac']):.3f} x bench.py's)."]
    traffic = rf.get("traffic")
    if traffic_md:
        with open(traffic_md) as f:
            traffic = parse_rows(f.read()).get((rf["traffic_key"], kernels), traffic)
    if traffic is not None:
        lines += ["", f"HBM traffic per dispatch (`hbm_traffic.md`, key {rf['traffic_key']}): {traffic:.0f} bytes "
                      f"= {float(traffic) / bpl:.2f} x the algorithmic bytes."]
    return "\n".join(lines) + "\n"


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
    ap.add_argument("--report", action="store_true",
                    help="write the profile summary of one bench leg (needs --kernel-stats, --bench, --leg, --title)")
    ap.add_argument("--bench", help="bench.py's JSON output of the same command run without the tracer")
    ap.add_argument("--leg", help="bench leg: sssp, pr, pr_c0, tc, bc")
    ap.add_argument("--title")
    ap.add_argument("--traffic", help="a summary table of counter rows (hbm_traffic.md) to take the leg's traffic from")
    ap.add_argument("--command", action="append", default=[], help="a command line to quote (repeatable)")
    a = ap.parse_args(argv)
    if a.report:
        if not (a.kernel_stats and a.bench and a.leg and a.title):
            ap.error("--report needs --kernel-stats, --bench, --leg and --title")
        sys.stdout.write(report(a.title, a.command, a.kernel_stats, a.bench, a.leg, a.top, a.traffic))
        return 0
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
