"""The mixed-graph batch: autotuner sweep, selector training, evaluation.

A seeded set of uneven graphs -- RMAT (scale, edge factor 4-32), Erdos-Renyi
(vertices, average degree 4-16) and 2-D grids -- each carrying one SSSP job
(int32 weights in [1,100], source 0) or one PageRank job (beta 1e-10, delta
0.85, maxIter 100). Graphs are generated and featurised on a background
thread while the GPU times earlier jobs, one job resident at a time (host and
device memory stay bounded). The first part of the set trains the selector
on (features, fastest class); the held-out part is scored by makespan (sum of
job times on the one GPU) under three policies -- the fixed heuristic (class
0), the learned selector and the best autotuned class per job -- re-timed in a
fresh pass after the selector has picked, with every class's output checked
against class 0's. The held-out jobs can also be run batched -- PageRank jobs
as disjoint unions through one batched launch sequence each (PrBatch), SSSP
jobs likewise (SsspBatch).
"""
from __future__ import annotations

import time

import numpy as np

from .graph import graph
from .schedule import FIXED_CLASS, JOB_PR, JOB_SSSP, N_CLASSES, Job, Selector, features, tie_aware_labels


def _rng(seed, i):
    return np.random.default_rng([seed, i])


RANGES = dict(rmat_scales=(12, 20), er_log2=(12, 20), grid_sides=(64, 1024))


def job_spec(seed, i, rmat_scales=(12, 20), er_log2=(12, 20), grid_sides=(64, 1024)):
    """Job i of the set without building its graph: (kind, params, job type,
    name); kind 0 RMAT (scale, edge factor), 1 Erdos-Renyi (log2 vertices,
    average degree), 2 grid (side). Drawn from a per-index stream, so job i
    depends only on (seed, i)."""
    r = _rng(seed, i)
    kind = int(r.integers(0, 3))
    if kind == 0:
        params = (int(r.integers(rmat_scales[0], rmat_scales[1] + 1)), int(r.integers(4, 33)))
        name = "rmat%d_ef%d" % params
    elif kind == 1:
        params = (int(r.integers(er_log2[0], er_log2[1] + 1)), int(r.integers(4, 17)))
        name = "er%d_d%d" % params
    else:
        params = (int(np.exp(r.uniform(np.log(grid_sides[0]), np.log(grid_sides[1])))),)
        name = "grid%d" % params
    job = JOB_SSSP if r.random() < 0.5 else JOB_PR
    return kind, params, job, name + ("_sssp" if job == JOB_SSSP else "_pr")


def make_job(seed, i, **ranges):
    """Job i of the set: its graph (weights U[1,100]) and job type."""
    kind, params, job, name = job_spec(seed, i, **ranges)
    w = (1, 100)
    gseed = seed * 100003 + i
    if kind == 0:
        g = graph.rmat(params[0], params[1], seed=gseed, weights=w)
    elif kind == 1:
        n = 1 << params[0]
        g = graph.erdos_renyi(n, n * params[1] // 2, seed=gseed, undirected=True, weights=w)
    else:
        g = graph.grid2d(params[0], params[0], seed=gseed, weights=w)
    return Job(g, job, 0, name)


def host_threads():
    """CPU threads this process may use: OMP_NUM_THREADS when set, else the
    CPUs it can run on."""
    import os
    env = os.environ.get("OMP_NUM_THREADS", "")
    n = int(env) if env.isdigit() and int(env) > 0 else len(os.sched_getaffinity(0))
    return max(1, n)


class Prefetch:
    """Iterates (i, job, features) over `indices`, building each job's graph
    and features on a background thread up to `depth` jobs ahead, so graph
    generation on the host overlaps the GPU work on earlier jobs. The builder
    thread uses all but two of the host threads; the caller's thread keeps
    the rest for launching and timing."""

    def __init__(self, seed, indices, depth=2, threads=None, **ranges):
        import queue
        import threading

        self.q = queue.Queue(maxsize=depth)
        self.stop = False
        self.waited = 0.0  # seconds the consumer spent waiting for a job
        n = threads or max(1, host_threads() - 2)

        def work():
            from . import _native as N
            N.host_lib().spmh_set_threads(n)
            try:
                for i in indices:
                    if self.stop:
                        break
                    job = make_job(seed, i, **ranges)
                    self.q.put((i, job, features(job.g, job.kind)))
            except BaseException as e:  # surfaced in the consumer
                self.q.put(e)
            self.q.put(None)

        self.t = threading.Thread(target=work, daemon=True)
        self.t.start()

    def __iter__(self):
        while True:
            t0 = time.perf_counter()
            item = self.q.get()
            self.waited += time.perf_counter() - t0
            if item is None:
                return
            if isinstance(item, BaseException):
                raise item
            yield item

    def close(self):
        self.stop = True
        while self.t.is_alive():
            try:
                self.q.get(timeout=0.1)
            except Exception:
                pass
        self.t.join()


def autotune_cutoff(job, reps=2, cutoff=8.0, outputs=None, stats=None, probe_iters=0):
    """Time every class: its first run (which also builds the class's plan:
    bins, captured graphs) and then the best of `reps` steady runs. A class
    whose first run is already `cutoff` times slower than the best so far
    keeps that first time and is not repeated -- it cannot be the fastest.
    With `probe_iters` > 0, a PageRank class after the first is probed with a
    run of that many iterations (plan build included); if the probe's time
    per iteration times class 0's iteration count already exceeds `cutoff`
    times the best, that estimate is kept and the full run skipped (PageRank's
    cost grows with its iterations, and every class runs the same iterations
    up to round-off).
    With `outputs` (a dict), outputs[k] = (result, iterations) of class k.
    With `stats` (a dict), wall seconds add up under "upload" (the handle's
    creation), "first_runs", "steady_runs" and "probes"."""
    from . import hip
    from .schedule import JOB_PR, apply_class
    out = np.zeros(N_CLASSES)
    best = float("inf")
    st = stats if stats is not None else {}
    t0 = time.perf_counter()
    job.g.device()
    hip.synchronize()
    st["upload"] = st.get("upload", 0.0) + time.perf_counter() - t0
    iters0 = None
    for k in range(N_CLASSES):
        apply_class(job.g.device(), k)
        hip.synchronize()
        if probe_iters > 0 and job.kind == JOB_PR and iters0 and outputs is None:
            t0 = time.perf_counter()
            job.run(max_iter=probe_iters)
            hip.synchronize()
            probe = time.perf_counter() - t0
            st["probes"] = st.get("probes", 0.0) + probe
            est = probe * 1e3 / max(job.iterations, 1) * iters0
            if est > cutoff * best:
                out[k] = est
                continue
        t0 = time.perf_counter()
        job.run()
        hip.synchronize()
        first = (time.perf_counter() - t0) * 1e3
        st["first_runs"] = st.get("first_runs", 0.0) + first * 1e-3
        if k == 0:
            iters0 = job.iterations
        if first > cutoff * best or reps < 1:
            out[k] = first
        else:
            t1 = time.perf_counter()
            t = min(first, time_job_steady(job, reps))
            st["steady_runs"] = st.get("steady_runs", 0.0) + time.perf_counter() - t1
            out[k] = t
            best = min(best, t)
        if outputs is not None:
            outputs[k] = (job.result(), job.iterations)
    return out


def time_job_steady(job, reps):
    from . import hip
    best = float("inf")
    for _ in range(reps):
        hip.synchronize()
        t0 = time.perf_counter()
        job.run()
        hip.synchronize()
        best = min(best, (time.perf_counter() - t0) * 1e3)
    return best


def sweep(seed, start, stop, reps=2, log=None, rows_path=None, **ranges):
    """Features and per-class times of jobs start..stop-1. Job i depends only
    on (seed, i), so disjoint slices can be swept separately; with rows_path
    every job is appended as one JSON line as soon as it is timed."""
    import json

    feats, times, names = [], [], []
    for i in range(start, stop):
        job = make_job(seed, i, **ranges)
        f = features(job.g, job.kind)
        t = autotune_cutoff(job, reps)
        feats.append(f)
        times.append(t)
        names.append(job.name)
        job.release()
        del job
        if rows_path:
            with open(rows_path, "a") as fh:
                fh.write(json.dumps({"i": i, "seed": seed, "name": names[-1], "features": [float(x) for x in f],
                                     "times_ms": [float(x) for x in t]}) + "\n")
        if log and ((i - start) % 25 == 0 or i == stop - 1):
            log(f"autotuned {i + 1 - start}/{stop - start}: {names[-1]} best class {int(np.argmin(t))} "
                f"({t.min():.2f} ms; fixed {t[FIXED_CLASS]:.2f} ms)")
    return np.array(feats, np.float32).reshape(-1, 16), np.array(times).reshape(-1, N_CLASSES), names


def batched_makespan(seed, indices, reps=2, edge_cap=1 << 28, log=None, keep=None, **ranges):
    """Steady-state time (ms) of jobs `indices` run in batches: PageRank jobs
    packed into disjoint unions of at most edge_cap edges, each union one
    spm_pr_multi_device run, and SSSP jobs likewise through
    spm_sssp_multi_device, all with the default schedule. Best of `reps`
    runs per union after an untimed first run (binning, graph capture), as
    the per-job times are taken. With `keep`, only the jobs for which
    keep(job) holds are batched; the others are listed under "left_out"."""
    from . import hip
    from .batch import PrBatch, SsspBatch

    packs = {JOB_PR: ([], 0), JOB_SSSP: ([], 0)}
    out = {"makespan_ms": 0.0, "pr_jobs": 0, "sssp_jobs": 0, "unions": 0, "left_out": []}

    def timed(fn):
        fn()
        best = float("inf")
        for _ in range(reps):
            hip.synchronize()
            t0 = time.perf_counter()
            fn()
            hip.synchronize()
            best = min(best, (time.perf_counter() - t0) * 1e3)
        return best

    def flush(kind):
        graphs, _ = packs[kind]
        if not graphs:
            return
        if kind == JOB_PR:
            b = PrBatch(graphs)
            t = timed(lambda: b.run(1e-10, 0.85, 100))
            out["pr_jobs"] += len(graphs)
        else:
            b = SsspBatch([(g, 0) for g in graphs])
            t = timed(b.run)
            out["sssp_jobs"] += len(graphs)
        b.close()
        out["makespan_ms"] += t
        out["unions"] += 1
        if log:
            log(f"batched {'PageRank' if kind == JOB_PR else 'SSSP'} union of {len(graphs)} jobs: {t:.2f} ms")
        packs[kind] = ([], 0)

    for i in indices:
        job = make_job(seed, i, **ranges)
        if keep is not None and not keep(job):
            out["left_out"].append(int(i))
            job.release()
            continue
        graphs, edges = packs[job.kind]
        if graphs and edges + job.g.E > edge_cap:
            flush(job.kind)
            graphs, edges = packs[job.kind]
        graphs.append(job.g)
        packs[job.kind] = (graphs, edges + job.g.E)
    flush(JOB_PR)
    flush(JOB_SSSP)
    out["makespan_ms"] = round(out["makespan_ms"], 3)
    return out


def compare_outputs(kind, ref, got, tol=1e-5):
    """(ok, rel_linf): SSSP distances must be bit-identical; PageRank ranks
    within `tol` relative L-infinity of the reference class's."""
    if kind == JOB_SSSP:
        return bool(np.array_equal(ref, got)), 0.0
    den = float(np.max(np.abs(ref))) if ref.size else 0.0
    err = float(np.max(np.abs(got.astype(np.float64) - ref))) / den if den > 0 else 0.0
    return bool(err <= tol and np.all(np.isfinite(got))), err


def run(n=2000, held_out=400, seed=5, train_reps=1, eval_reps=2, epochs=400, log=None, on_held_out=None,
        batched=False, **ranges):
    """The mixed-batch study end to end.

    1. Training sweep: jobs 0..n-held_out-1 are built (on a background
       thread), featurised and timed under every schedule class
       (autotune_cutoff with `train_reps` steady runs per class, PageRank
       classes probed with 4 iterations first).
    2. The selector is trained on the GPU on (features, tie-aware labels).
    3. Held-out jobs: each is built, the learned class is picked from its
       features, every class is timed (its output kept and checked against
       class 0's: SSSP bit-exact, PageRank within 1e-5 relative L-infinity),
       and the fastest class of that sweep is the best-autotuned pick. Then a
       fresh pass times the distinct classes the three policies picked
       (fixed = class 0, learned, best), in an order rotated from job to job,
       each run once untimed and then best of `eval_reps`; the makespans are
       the sums of these fresh times. `on_held_out(i, job, outputs, picks)` is
       called with every held-out job's per-class outputs ((result,
       iterations) by class) and the policies' picks before its graph is
       freed.
    """
    from . import hip
    from .schedule import apply_class

    ranges = {**RANGES, **ranges}
    ntr = n - held_out
    feats = np.zeros((n, 16), np.float32)
    times = np.zeros((n, N_CLASSES))
    t0 = time.perf_counter()
    sweep_stats = {}
    pf = Prefetch(seed, range(ntr), **ranges)
    try:
        for c, (i, job, f) in enumerate(pf):
            feats[i] = f
            times[i] = autotune_cutoff(job, train_reps, stats=sweep_stats, probe_iters=4)
            job.release()
            if log and (c % 200 == 0 or c == ntr - 1):
                log(f"trained-set sweep {c + 1}/{ntr}: {job.name} best class {int(np.argmin(times[i]))} "
                    f"({times[i].min():.2f} ms; fixed {times[i][FIXED_CLASS]:.2f} ms)")
    finally:
        pf.close()
    sweep_stats["waiting_for_graphs"] = pf.waited
    t_sweep = time.perf_counter() - t0
    sel = Selector()
    t1 = time.perf_counter()
    loss = sel.fit(feats[:ntr], times[:ntr], epochs=epochs, seed=seed)
    t_train = time.perf_counter() - t1

    policies = ("fixed_heuristic", "learned", "best_autotuned")
    fresh = {p: 0.0 for p in policies}
    picks = {p: np.zeros(held_out, np.int32) for p in policies}
    chk = {"jobs": 0, "class_outputs_compared": 0, "mismatches": [], "pr_max_rel_linf": 0.0,
           "pr_iteration_counts_differ": 0, "tolerance_pr": 1e-5}
    t2 = time.perf_counter()
    pf = Prefetch(seed, range(ntr, n), **ranges)
    try:
        for j, (i, job, f) in enumerate(pf):
            feats[i] = f
            kl = int(sel.choose(f)[0])
            outs = {}
            times[i] = autotune_cutoff(job, eval_reps, outputs=outs)
            kb = int(np.argmin(times[i]))
            pk = {"fixed_heuristic": FIXED_CLASS, "learned": kl, "best_autotuned": kb}
            for p in policies:
                picks[p][j] = pk[p]
            ref, ref_it = outs[FIXED_CLASS]
            chk["jobs"] += 1
            for k in range(N_CLASSES):
                if k == FIXED_CLASS:
                    continue
                ok, err = compare_outputs(job.kind, ref, outs[k][0])
                chk["class_outputs_compared"] += 1
                if job.kind == JOB_PR:
                    chk["pr_max_rel_linf"] = max(chk["pr_max_rel_linf"], err)
                    chk["pr_iteration_counts_differ"] += int(outs[k][1] != ref_it)
                if not ok:
                    chk["mismatches"].append({"job": int(i), "name": job.name, "class": k, "rel_linf": err})
            if on_held_out is not None:
                on_held_out(i, job, outs, pk)
            classes = sorted(set(pk.values()))
            r = j % len(classes)
            ms = {}
            for k in classes[r:] + classes[:r]:
                apply_class(job.g.device(), k)
                job.run()  # plan build for this class, untimed
                hip.synchronize()
                ms[k] = time_job_steady(job, eval_reps)
            for p in policies:
                fresh[p] += ms[pk[p]]
            job.release()
            if log and (j % 100 == 0 or j == held_out - 1):
                log(f"held-out {j + 1}/{held_out}: {job.name} fixed {ms[FIXED_CLASS]:.2f} ms, learned class {kl} "
                    f"{ms[kl]:.2f} ms, best class {kb} {ms[kb]:.2f} ms")
    finally:
        pf.close()
    t_eval = time.perf_counter() - t2
    ho = times[ntr:]
    idx = np.arange(held_out)
    sweep_ms = {"fixed_heuristic": float(ho[:, FIXED_CLASS].sum()), "learned": float(ho[idx, picks["learned"]].sum()),
                "best_autotuned": float(ho.min(axis=1).sum())}
    train_acc = float(np.mean(sel.choose(feats[:ntr]) == tie_aware_labels(times[:ntr])))
    out = {
        "graphs": n, "trained": ntr, "held_out": held_out, "seed": seed,
        "ranges": {k: list(v) for k, v in ranges.items()},
        "makespan_ms": {p: round(fresh[p], 3) for p in policies},
        "learned_vs_fixed_speedup": round(fresh["fixed_heuristic"] / fresh["learned"], 4),
        "learned_over_best": round(fresh["learned"] / fresh["best_autotuned"] - 1.0, 4),
        "sweep_makespan_ms": {p: round(v, 3) for p, v in sweep_ms.items()},
        "outputs_checked": chk["jobs"] == held_out and not chk["mismatches"],
        "output_checks": chk,
        "train_final_loss": round(float(loss), 4), "train_accuracy": round(train_acc, 4),
        "held_out_class_histogram": np.bincount(picks["learned"], minlength=N_CLASSES).tolist(),
        "best_class_histogram": np.bincount(picks["best_autotuned"], minlength=N_CLASSES).tolist(),
        "seconds": {"training_sweep": round(t_sweep, 1), "train_on_gpu": round(t_train, 2),
                    "held_out_sweep_and_fresh_pass": round(t_eval, 1),
                    "training_sweep_parts": {k: round(v, 1) for k, v in sweep_stats.items()}},
        "timing": (f"training jobs: every class timed as its first run (plan build) and best of {train_reps} "
                   "steady runs, a PageRank class after class 0 probed with 4 iterations first and, when that "
                   "extrapolates past 8x the best class so far, kept at the estimate; held-out jobs: every class "
                   f"run in full and timed likewise with best of {eval_reps}, then the "
                   f"policies' distinct picks re-timed in a fresh pass (one untimed run, best of {eval_reps}), "
                   "class order rotated per job; makespan = sum of the fresh-pass times"),
    }
    if batched:
        bat = batched_makespan(seed, range(ntr, n), reps=eval_reps, log=log, **ranges)
        bat["vs_fixed_per_job"] = round(fresh["fixed_heuristic"] / bat["makespan_ms"], 4)
        bat["vs_learned_per_job"] = round(fresh["learned"] / bat["makespan_ms"], 4)
        del bat["left_out"]
        out["batched_held_out"] = bat
    return out
