"""Command line for the mixed-batch schedule study.

    python -m starplat_amd.sched_sweep study --graphs 2000 --held-out 400
    python -m starplat_amd.sched_sweep sweep --start 0 --stop 500 --rows rows.jsonl

`study` runs the whole study (mixed_batch.run): training sweep, selector
training on the GPU, held-out sweep with output checks and the fresh timed
pass of the fixed, learned and best-autotuned picks; it prints one JSON line.
`sweep` times every schedule class on jobs start..stop-1 of the seeded set
(RMAT scales 12-20 with edge factor 4-32, Erdos-Renyi 2^12-2^20 vertices with
average degree 4-16, grids 64^2 to 1024^2) and appends one JSON line per job;
slices can run in separate processes.
"""
from __future__ import annotations

import argparse
import json
import sys

from . import mixed_batch


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    s = sub.add_parser("sweep")
    s.add_argument("--start", type=int, required=True)
    s.add_argument("--stop", type=int, required=True)
    s.add_argument("--rows", required=True)
    s.add_argument("--seed", type=int, default=5)
    s.add_argument("--reps", type=int, default=2)
    e = sub.add_parser("study")
    e.add_argument("--graphs", type=int, default=2000)
    e.add_argument("--held-out", type=int, default=400)
    e.add_argument("--epochs", type=int, default=400)
    e.add_argument("--seed", type=int, default=5)
    e.add_argument("--batched", action="store_true", help="also time the held-out jobs in batches")
    a = ap.parse_args(argv)
    log = lambda *m: print("[sched]", *m, file=sys.stderr, flush=True)  # noqa: E731
    if a.cmd == "sweep":
        mixed_batch.sweep(a.seed, a.start, a.stop, a.reps, log, rows_path=a.rows)
        return 0
    res = mixed_batch.run(n=a.graphs, held_out=a.held_out, seed=a.seed, epochs=a.epochs, batched=a.batched, log=log)
    print(json.dumps(res))
    return 0


if __name__ == "__main__":
    sys.exit(main())
