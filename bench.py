"""Benchmark of the StarPlat hot path on one MI355X.

Headline (the JSON `value`): SSSP Bellman-Ford fixedPoint on a directed RMAT
scale-22 graph (edge factor 16, a,b,c = .57,.19,.19, int32 weights uniform in
[1,100], source 0, seed 2), in GTEPS = m_reach / time, where m_reach is the
number of out-edges of vertices reachable from the source (Graph500
convention). One step = one Compute_SSSP on the graph already resident in HBM.
PageRank, triangle counting, batched BC, the dynamic programs and the
mixed-batch schedule study (2,000 graphs, 400 held out) are measured in the
same run and reported under "per_algo" / "schedule_selection".

Roofline: algorithmic bytes of the dominant kernel's work divided by that
kernel's device time, measured live with start/stop events that each
dispatch stamps itself (hipExtLaunchKernel) on the handle's stream -- the
begin/end times a kernel trace reads, so the figures agree with rocprofv3's.
CPU baseline: the OpenMP restatement of StarPlat's generated code (oracle/),
run after the GPU legs in a child process of this script whose environment
sets OMP_PROC_BIND=close and OMP_PLACES=cores before OpenMP starts; the child
regenerates the same seeded graphs, times the restatement and checks the GPU
outputs handed over in files.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from starplat_amd import _native as N  # noqa: E402
from starplat_amd import graph, hip, sources  # noqa: E402
from starplat_amd import pmc_summary  # noqa: E402

HBM_PEAK_GBPS = 8000.0
METRIC = "GTEPS per algo at RMAT-22 on 1 MI355X, % HBM roofline; learned vs fixed sched"


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def timed_steps(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    hip.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    hip.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def kernel_time(dev, fn, kernels, reps=2):
    """Total device ms and launches of `kernels` over `reps` profiled runs."""
    dev.profile(True)
    dev.profile_reset()
    for _ in range(reps):
        fn()
    out = {k: dev.profile_query(k) for k in kernels}
    dev.profile(False)
    return {k: (ms / reps, n / reps) for k, (ms, n) in out.items()}


def measured_traffic(workload, kernels):
    """HBM bytes per dispatch of `kernels` on `workload` from the rocprofv3
    FETCH_SIZE / WRITE_SIZE summaries in profiles/ (rows written by
    starplat_amd.pmc_summary), or None if that workload was not measured."""
    rows = {}
    pdir = os.path.join(ROOT, "profiles")
    if os.path.isdir(pdir):
        for name in sorted(os.listdir(pdir)):
            if name.endswith(".md"):
                with open(os.path.join(pdir, name)) as f:
                    rows.update(pmc_summary.parse_rows(f.read()))
    return rows.get((workload, tuple(kernels)))


def roofline(bytes_per_run, kms, launches, workload, kernels):
    achieved = bytes_per_run / (kms * 1e-3) / 1e9 if kms > 0 else 0.0
    per_launch = bytes_per_run / max(launches, 1)
    traffic = measured_traffic(workload, kernels)
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None if traffic is None else round(traffic),
           "traffic_source": None if traffic is None else
           "from profiles/ (committed rocprofv3 FETCH_SIZE + WRITE_SIZE passes), not measured in this run",
           "kernel_time_source": "HIP events stamped by each dispatch (hipExtLaunchKernel), graphs off",
           "bytes_per_launch": round(per_launch),
           "kernel_ms_per_launch": round(kms / max(launches, 1), 5), "launches_per_step": launches,
           "kernels": list(kernels), "traffic_key": workload}
    if traffic is not None:
        out["traffic_over_algorithmic"] = round(traffic / per_launch, 3)
    return out


def copy_bandwidth(mib=1024, reps=10):
    """A STREAM-style check of the HBM peak the rooflines are priced against:
    device-to-device copies of `mib` MiB (bytes read + bytes written per copy
    over the time of `reps` copies, events on the null stream)."""
    n = mib << 20
    a = hip.DeviceArray(n // 4, np.int32)
    b = hip.DeviceArray(n // 4, np.int32)
    try:
        hip._chk("hipMemset", hip.lib().hipMemset(a.ptr, 1, n))
        hip._chk("hipMemcpy", hip.lib().hipMemcpy(b.ptr, a.ptr, n, hip.D2D))
        hip.synchronize()
        t0, t1 = hip.Event(), hip.Event()
        t0.record(None)
        for _ in range(reps):
            hip._chk("hipMemcpyAsync", hip.lib().hipMemcpyAsync(b.ptr, a.ptr, n, hip.D2D, None))
        t1.record(None)
        t1.synchronize()
        ms = t0.elapsed_ms(t1)
    finally:
        a.free()
        b.free()
    return {"bytes_per_copy": 2 * n, "copies": reps, "GBps": round(2 * n * reps / (ms * 1e-3) / 1e9, 1),
            "frac_of_peak": round(2 * n * reps / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS, 3)}


def gather_ceiling(n=1 << 22):
    """Random 4-byte loads per second the GPU sustains over an array of n
    floats (the size of RMAT-22's contribution / distance arrays): every
    workgroup over the whole array, and each XCD over its own eighth."""
    out = {"array_floats": n}
    for key, local in (("uniform", 0), ("xcd_local_eighths", 1)):
        r = C.c_double()
        N.check("spm_gather_rate", N.device_lib().spm_gather_rate(n, local, C.byref(r)))
        out[key + "_G_per_s"] = round(r.value / 1e9, 1)
    return out


def host_cores():
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        keep = [l for l in out.splitlines() if l.split(":")[0] in ("Socket(s)", "Core(s) per socket", "Thread(s) per core", "Model name")]
        return "; ".join(" ".join(l.split()) for l in keep)
    except Exception:
        return platform.processor()


# --------------------------------------------------------------------------- graphs
# The seeded inputs of every leg; the CPU child rebuilds them from the same
# arguments (the generators are deterministic).
def graph_for(leg, args):
    if leg in ("sssp", "dyn"):
        return graph.rmat(args.sssp_scale, 16, seed=2, weights=(1, 100))
    if leg in ("pr", "dyn_pr"):
        return graph.rmat(22, 16, seed=1)
    if leg == "pr_c0":
        return graph.rmat(16, 16, seed=1)
    if leg in ("tc", "dyn_tc"):
        return graph.rmat(args.tc_scale, 16, seed=3, undirected=True)
    if leg == "bc":
        return graph.erdos_renyi(args.bc_n, args.bc_n * 4, seed=4, undirected=True)
    raise KeyError(leg)


class Stash:
    """GPU outputs handed to the CPU child: one .npz of arrays and a JSON of
    scalars per leg, in a temporary directory."""

    def __init__(self, path):
        self.path = path

    def put(self, leg, arrays=None, **meta):
        if arrays:
            np.savez(os.path.join(self.path, leg + ".npz"), **arrays)
        with open(os.path.join(self.path, leg + ".json"), "w") as f:
            json.dump(meta, f)

    def legs(self):
        return sorted(n[:-5] for n in os.listdir(self.path) if n.endswith(".json") and n != "cpu.json")

    def get(self, leg):
        with open(os.path.join(self.path, leg + ".json")) as f:
            meta = json.load(f)
        p = os.path.join(self.path, leg + ".npz")
        arrays = dict(np.load(p)) if os.path.exists(p) else {}
        return meta, arrays


def warm_library():
    """Runs every static program once on a tiny graph, so the one-time load
    of the library's code objects (on a process's first launch from each
    source file) is not charged to the first call on the benchmark graphs:
    first_run_ms is the first call on a new graph handle."""
    g = graph.rmat(10, 8, seed=9, weights=(1, 100))
    u = graph.rmat(10, 8, seed=9, undirected=True)
    dev, dev_u = g.device(), u.device()
    a = hip.DeviceArray(g.V, np.int32)
    b = hip.DeviceArray(g.V, np.float32)
    c = hip.DeviceArray(u.V, np.float32)
    dev.sssp_device(0, a)
    dev.set_schedule(relabel=1)
    dev.sssp_device(0, a)
    dev.pr_device(1e-10, 0.85, 10, b)
    dev.set_schedule(pr_mode=0)
    dev.pr_device(1e-10, 0.85, 10, b)
    dev_u.tc()
    dev_u.bc_device(sources(u.V, 4, seed=1), c)
    hip.synchronize()
    for x in (a, b, c):
        x.free()
    g.release_device()
    u.release_device()


# --------------------------------------------------------------------------- SSSP
def bench_sssp(args, stash):
    log("generating RMAT-22 (directed, weights U[1,100], seed 2)")
    g = graph_for("sssp", args)
    r, m, rl = C.c_int64(), C.c_int64(), C.c_int64()
    model = N.host_lib().spmh_model_sssp(g._native(), N.ptr(g.edgeLen, C.c_int32), 0, C.byref(r), C.byref(m), C.byref(rl))
    dev = g.device()
    d_out = hip.DeviceArray(g.V, np.int32)
    hip.synchronize()
    t0 = time.perf_counter()
    rounds = dev.sssp_device(0, d_out)  # includes the one-time layout (relabeled copy, queues)
    first_ms = (time.perf_counter() - t0) * 1e3
    ms = timed_steps(lambda: dev.sssp_device(0, d_out), args.steps, args.warmup)
    kt = kernel_time(dev, lambda: dev.sssp_device(0, d_out), ["sssp_relax"])
    kms, launches = kt["sssp_relax"]
    res = {
        "workload": f"SSSP Bellman-Ford fixedPoint, directed RMAT scale {args.sssp_scale} EF16 (a,b,c)=(.57,.19,.19), "
                    "int32 weights U[1,100], src 0, seed 2",
        "V": g.V, "E": g.E, "m_reach": m.value, "rounds_gpu": rounds, "rounds_sync_trace": r.value,
        "relaxations_sync_trace": rl.value, "ms_per_step": round(ms, 4),
        "first_run_ms": round(first_ms, 3), "relabel": int(dev.schedule().relabel),
        "gteps": m.value / (ms * 1e-3) / 1e9,
        "gteps_first_call": m.value / (first_ms * 1e-3) / 1e9,
        "roofline": roofline(model, kms, launches, f"sssp-rmat{args.sssp_scale}", ["sssp_relax"]),
        "algorithmic_bytes": model,
    }
    stash.put("sssp", {"dist": d_out.to_host()}, m_reach=m.value)
    d_out.free()
    g.release_device()
    return res


# --------------------------------------------------------------------------- PR
def bench_pr(args, stash, leg, label):
    scale = 22 if leg == "pr" else 16
    log(f"generating RMAT-{scale} for PageRank")
    g = graph_for(leg, args)
    model = N.host_lib().spmh_model_pr_bytes(g._native())
    dev = g.device()
    d_out = hip.DeviceArray(g.V, np.float32)
    hip.synchronize()
    t0 = time.perf_counter()
    it = dev.pr_device(1e-10, 0.85, 100, d_out)  # includes the one-time layout (relabeled copy, bins)
    first_ms = (time.perf_counter() - t0) * 1e3
    ms = timed_steps(lambda: dev.pr_device(1e-10, 0.85, 100, d_out), max(1, args.steps // 2), 1)
    # The sweep kernel: pr_iter, pr_iter_lag where the schedule folds the
    # loop test into the sweep (small graphs), or pr_iter_hot where it stages
    # the lowest ids' contributions in LDS (relabeled graphs).
    kt = kernel_time(dev, lambda: dev.pr_device(1e-10, 0.85, 100, d_out), ["pr_iter", "pr_iter_lag", "pr_iter_hot"], reps=1)
    kname = max(kt, key=lambda k: kt[k][1])
    kms, launches = kt[kname]
    res = {"workload": f"{label}: pull PageRank, directed RMAT scale {scale} EF16, beta 1e-10, delta 0.85, maxIter 100, seed 1",
           "V": g.V, "E": g.E, "iterations": it, "ms_per_run": round(ms, 4),
           "first_run_ms": round(first_ms, 3), "relabel": int(dev.schedule().relabel),
           "gteps": g.E * it / (ms * 1e-3) / 1e9,
           "gteps_first_call": g.E * it / (first_ms * 1e-3) / 1e9,
           "pr_mode": int(dev.schedule().pr_mode),
           # in-edge gathers (E per iteration) per second of sweep-kernel time
           "gathers_G_per_s": round(g.E * it / (kms * 1e-3) / 1e9, 1) if kms > 0 else 0.0,
           "roofline": roofline(model * it, kms, launches, f"pr-rmat{scale}", [kname])}
    stash.put(leg, {"ranks": d_out.to_host()}, iterations=it)
    d_out.free()
    g.release_device()
    return res


# --------------------------------------------------------------------------- TC
def bench_tc(args, stash):
    log("generating undirected RMAT-21 for TC")
    g = graph_for("tc", args)
    mu, wk = C.c_int64(), C.c_int64()
    model = N.host_lib().spmh_model_tc_bytes(g._native(), C.byref(mu), C.byref(wk))
    dev = g.device()
    hip.synchronize()
    t0 = time.perf_counter()
    count = dev.tc()  # includes the one-time symmetry test and orientation
    first_ms = (time.perf_counter() - t0) * 1e3
    ms = timed_steps(lambda: dev.tc(), max(1, args.steps // 4), 1)
    # the hashed count is the default; the search runs when the schedule or the graph asks for it
    kt = kernel_time(dev, lambda: dev.tc(), ["tc_count_hash", "tc_count"], reps=1)
    kname = "tc_count_hash" if kt["tc_count_hash"][1] > 0 else "tc_count"
    kms, launches = kt[kname]
    res = {"workload": f"Triangle counting, undirected RMAT scale {args.tc_scale} EF16, symmetrised, sorted, seed 3",
           "V": g.V, "E": g.E, "m_und": mu.value, "triangles": count, "ms_per_run": round(ms, 4),
           "first_run_ms": round(first_ms, 3), "tc_mode": int(dev.schedule().tc_mode), "kernel": kname,
           "gteps": mu.value / (ms * 1e-3) / 1e9,
           "gteps_first_call": mu.value / (first_ms * 1e-3) / 1e9,
           "roofline": roofline(model, kms, launches, f"tc-rmat{args.tc_scale}", [kname])}
    stash.put("tc", triangles=count, m_und=mu.value)
    g.release_device()
    return res


# --------------------------------------------------------------------------- BC
def bc_gather_profile(g, batch):
    """The access pattern of one 64-source batch (host BFS levels): how many
    sources share each per-source row gather of bc_sigma / bc_back and how
    many 128-byte lines those gathers touch."""
    out = np.zeros(7)
    rc = N.host_lib().spmh_bc_gather_profile(g._native(), N.ptr(np.ascontiguousarray(batch), C.c_int32), len(batch),
                                             N.ptr(out, C.c_double))
    if rc != 0:
        return None
    dag, gathers, l64, l32, listed, rows, levels = out
    return {"batch_sources": len(batch), "dag_edges": int(dag), "row_gathers": int(gathers),
            "mean_sources_per_gather": round(dag / gathers, 3),
            "lines_per_gather_fp64_row": round(l64 / gathers, 3), "lines_per_gather_fp32_row": round(l32 / gathers, 3),
            "gather_bytes_per_dag_edge_sigma_fp64": round(l64 * 128 / dag, 2),
            "gather_bytes_per_dag_edge_coef_fp32": round(l32 * 128 / dag, 2),
            "listings_per_vertex": round(listed / g.V, 3), "in_edges_read_per_batch": int(rows), "levels": int(levels)}


def bench_bc(args, stash):
    log("generating Erdos-Renyi 4M / avg degree 8 for BC")
    g = graph_for("bc", args)
    src = sources(g.V, args.bc_sources, seed=4)
    et = C.c_int64()
    log("BC byte model (one BFS per source on the host)")
    model = N.host_lib().spmh_model_bc_bytes(g._native(), N.ptr(src, C.c_int32), len(src), C.byref(et))
    dev = g.device()
    d_out = hip.DeviceArray(g.V, np.float32)
    dev.bc_device(src, d_out)
    ms = timed_steps(lambda: dev.bc_device(src, d_out), max(1, args.steps // 8), 0)
    ks = ["bc_scan", "bc_push", "bc_sigma", "bc_back"]
    kt = kernel_time(dev, lambda: dev.bc_device(src, d_out), ks, reps=1)
    kms = sum(v[0] for v in kt.values())
    launches = sum(v[1] for v in kt.values())
    res = {"workload": f"Batched Brandes BC, Erdos-Renyi V={args.bc_n} avg degree 8, {len(src)} sources (seed 4), FP64 sigma",
           "V": g.V, "E": g.E, "ms_per_run": round(ms, 3),
           "gteps": len(src) * g.E / (ms * 1e-3) / 1e9,
           "kernel_ms": {k: round(v[0], 3) for k, v in kt.items()},
           "roofline": roofline(model, kms, launches, f"bc-er{args.bc_n}-{len(src)}", ks),
           "note": "model bytes count every source's traversal separately; the batched kernels share adjacency reads "
                   "across 64 sources, so achieved model GB/s can exceed the HBM peak",
           "gather_profile": bc_gather_profile(g, src[:min(64, len(src))])}
    # the first full 64-source batch alone, for the CPU child's parity check and timing sample
    k = min(64, len(src))
    dev.bc_device(src[:k], d_out)
    stash.put("bc", {"bc_first_batch": d_out.to_host()}, sources=len(src), batch=k)
    d_out.free()
    g.release_device()
    return res
# --------------------------------------------------------------------------- dynamic SSSP
def dyn_batch(g, src_of, size, rng):
    """One batch of `size` updates: half deletions of edges of the original
    graph, half additions of uniform vertex pairs (weight 1, update.hpp)."""
    nd = size // 2
    e = rng.integers(0, g.E, nd)
    us = np.concatenate([src_of[e], rng.integers(0, g.V, size - nd)]).astype(np.int32)
    ut = np.concatenate([g.edgeList[e], rng.integers(0, g.V, size - nd)]).astype(np.int32)
    return b"d" * nd + b"a" * (size - nd), us, ut


def bench_dyn(args, stash):
    log(f"dynamic SSSP on RMAT-{args.sssp_scale}: {args.dyn_batches} batches of {args.dyn_batch} updates")
    g = graph_for("dyn", args)
    src_of = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(g.indexofNodes))
    rng = np.random.default_rng(5)
    batches = [dyn_batch(g, src_of, args.dyn_batch, rng) for _ in range(args.dyn_batches + 1)]
    dev = g.device()
    hip.synchronize()
    t0 = time.perf_counter()
    dev.dynsssp_init(0)
    t_init = (time.perf_counter() - t0) * 1e3
    dev.dynsssp_batch(*batches[0])  # warm-up (first allocation of the overflow rows)
    times = []
    for b in batches[1:-1]:
        t0 = time.perf_counter()
        dev.dynsssp_batch(*b)
        times.append((time.perf_counter() - t0) * 1e3)
    kt = kernel_time(dev, lambda: dev.dynsssp_batch(*batches[-1]),
                     ["dsp_on_delete", "dyn_delete", "dsp_invalidate", "dsp_pull", "dsp_on_add", "dyn_count",
                      "dyn_difflen", "dyn_diffcopy", "dyn_place", "dsp_push"], reps=1)
    dist, parent = dev.dynsssp_result()
    # the same graph from scratch: fold the updates, static SSSP, parents
    hip.synchronize()
    t0 = time.perf_counter()
    dev.compact()
    t_fold = (time.perf_counter() - t0) * 1e3
    d_out = hip.DeviceArray(g.V, np.int32)
    t0 = time.perf_counter()
    dev.sssp_device(0, d_out)  # rebuilds the layout the fold dropped (relabeled copy, queues)
    t_static_first = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    dev.sssp_device(0, d_out)
    t_static = (time.perf_counter() - t0) * 1e3
    ref = d_out.to_host()
    d_out.free()
    reach = ref != np.iinfo(np.int32).max
    ok = bool(np.array_equal(dist[reach], ref[reach]) and np.all(dist[~reach] >= np.iinfo(np.int32).max // 2))
    ms = statistics.median(times)
    res = {"workload": f"DynSSSP (dynamicBatchSSSP) on directed RMAT scale {args.sssp_scale} EF16 weights U[1,100], "
                       f"src 0; batches of {args.dyn_batch} updates, half deletions of existing edges, half random "
                       "additions of weight 1",
           "V": g.V, "E": g.E, "batch": args.dyn_batch, "batches_timed": len(times),
           "ms_per_batch": round(ms, 4), "updates_per_s": round(args.dyn_batch / (ms * 1e-3)),
           "init_ms": round(t_init, 3), "fold_ms": round(t_fold, 3),
           "static_sssp_first_ms": round(t_static_first, 3), "static_sssp_ms": round(t_static, 3),
           "speedup_vs_fold_plus_static": round((t_fold + t_static) / ms, 2),
           "dist_equals_static_recompute": ok,
           "kernel_ms_one_batch": {k: round(v[0], 4) for k, v in kt.items() if v[1] > 0},
           "roofline": None,
           "note": "per-batch time includes the host's staging copy and one flag read-back per 8 fixed-point "
                   "iterations; the speedup compares with folding the updates plus a steady-state static SSSP "
                   "(static_sssp_first_ms also rebuilds the relabeled layout the fold drops)"}
    stash.put("dyn")
    g.release_device()
    return res


def dyn_tc_batch(g, src_of, size, rng, touched):
    """size directed updates as parseUpdateFile pushes them (each undirected
    update and its reverse): half deletions of existing edges, half additions
    of absent pairs, no pair touched twice over the whole run (so the count
    stays comparable with a recount)."""
    off, adj = g.indexofNodes, g.edgeList
    ts, us, ut = [], [], []
    for i in rng.choice(g.E, size * 2, replace=False).tolist():
        a, c = int(src_of[i]), int(adj[i])
        key = (min(a, c), max(a, c))
        if a == c or key in touched:
            continue
        touched.add(key)
        ts.append("dd"), us.extend([a, c]), ut.extend([c, a])
        if len(us) >= size // 2:
            break
    while len(us) < size:
        a, c = int(rng.integers(0, g.V)), int(rng.integers(0, g.V))
        key = (min(a, c), max(a, c))
        row = adj[off[a]:off[a + 1]]
        j = np.searchsorted(row, c)
        if a == c or key in touched or (j < len(row) and row[j] == c):
            continue
        touched.add(key)
        ts.append("aa"), us.extend([a, c]), ut.extend([c, a])
    return "".join(ts).encode(), np.array(us, np.int32), np.array(ut, np.int32)


def bench_dyn_tc(args, stash):
    log(f"dynamic TC on undirected RMAT-{args.tc_scale}: {args.dyn_batches} batches of {args.dyn_batch} updates")
    g = graph_for("dyn_tc", args)
    src_of = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(g.indexofNodes))
    rng = np.random.default_rng(6)
    touched = set()
    batches = [dyn_tc_batch(g, src_of, args.dyn_batch, rng, touched) for _ in range(args.dyn_batches + 1)]
    dev = g.device()
    hip.synchronize()
    t0 = time.perf_counter()
    dev.dyntc_init()  # staticTC (orientation, hashed count)
    t_init = (time.perf_counter() - t0) * 1e3
    dev.dyntc_batch(*batches[0])  # warm-up (first allocation of the overflow rows and the pair set)
    times = []
    for b in batches[1:-1]:
        hip.synchronize()
        t0 = time.perf_counter()
        dev.dyntc_batch(*b)
        hip.synchronize()
        times.append((time.perf_counter() - t0) * 1e3)
    kt = kernel_time(dev, lambda: dev.dyntc_batch(*batches[-1]),
                     ["dtc_set_insert", "dtc_chunks", "dtc_count", "dtc_apply", "dyn_delete", "dyn_count",
                      "dyn_difflen", "dyn_diffcopy", "dyn_place"], reps=1)
    count = dev.dyntc_count()
    hip.synchronize()
    t0 = time.perf_counter()
    dev.compact()
    t_fold = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    recount = dev.tc()  # re-orients the folded graph
    t_static_first = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    dev.tc()
    t_static = (time.perf_counter() - t0) * 1e3
    ms = statistics.median(times)
    res = {"workload": f"DynTC (dynamicBatchTCV2) on undirected RMAT scale {args.tc_scale} EF16 seed 3; batches of "
                       f"{args.dyn_batch} directed updates (each undirected update and its reverse), half deletions "
                       "of existing edges, half additions of absent pairs",
           "V": g.V, "E": g.E, "batch": args.dyn_batch, "batches_timed": len(times),
           "ms_per_batch": round(ms, 4), "updates_per_s": round(args.dyn_batch / (ms * 1e-3)),
           "init_ms": round(t_init, 3), "fold_ms": round(t_fold, 3), "static_tc_first_ms": round(t_static_first, 3),
           "static_tc_ms": round(t_static, 3), "triangles": count,
           "count_equals_static_recount": bool(count == recount),
           "speedup_vs_fold_plus_static": round((t_fold + t_static) / ms, 2),
           "kernel_ms_one_batch": {k: round(v[0], 4) for k, v in kt.items() if v[1] > 0},
           "roofline": None}
    stash.put("dyn_tc")
    g.release_device()
    return res


def bench_dyn_pr(args):
    log(f"dynamic PageRank on RMAT-{args.sssp_scale}: 3 batches of {args.dyn_batch} updates")
    g = graph_for("dyn_pr", args)
    src_of = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(g.indexofNodes))
    has_out = np.flatnonzero(np.diff(g.indexofNodes) > 0).astype(np.int32)
    rng = np.random.default_rng(7)

    def pr_batch(size):
        half = size // 2
        e = rng.integers(0, g.E, half)
        types = b"d" * half + b"a" * (size - half)
        us = np.concatenate([src_of[e], rng.choice(has_out, size - half)]).astype(np.int32)
        ut = np.concatenate([g.edgeList[e], rng.choice(has_out, size - half)]).astype(np.int32)
        return types, us, ut

    batches = [pr_batch(args.dyn_batch) for _ in range(3)]
    dev = g.device()
    hip.synchronize()
    t0 = time.perf_counter()
    dev.dynpr_init(1e-10, 0.85, 100)
    t_init = (time.perf_counter() - t0) * 1e3
    times, sweeps = [], []
    for b in batches:
        n0 = len(dev.dynpr_iterations())
        t0 = time.perf_counter()
        dev.dynpr_batch(*b)
        times.append((time.perf_counter() - t0) * 1e3)
        sweeps.append(int(dev.dynpr_iterations()[n0:].sum()))
    kt = kernel_time(dev, lambda: dev.dynpr_batch(*pr_batch(args.dyn_batch)),
                     ["dpr_contrib", "dpr_sweep_light", "dpr_sweep_heavy", "dpr_apply", "dpr_diff", "dpr_propagate"],
                     reps=1)
    ksweep = kt["dpr_sweep_light"][0] + kt["dpr_sweep_heavy"][0] + kt["dpr_contrib"][0] + kt["dpr_apply"][0]
    nsweep = kt["dpr_sweep_light"][1]
    res = {"workload": f"DynPR (dynamicBatchPR, as emitted: starting out-degrees, signed diff) on directed RMAT "
                       f"scale {args.sssp_scale} EF16 seed 1, beta 1e-10, delta 0.85, maxIter 100; batches of "
                       f"{args.dyn_batch} updates, half deletions of existing edges, half additions between "
                       "vertices with out-edges",
           "V": g.V, "E": g.E, "batch": args.dyn_batch, "init_ms": round(t_init, 3),
           "ms_per_batch": [round(t, 3) for t in times], "sweeps_per_batch": sweeps,
           "ms_per_sweep": round(ksweep / max(nsweep, 1), 4),  # contributions, both sweep kernels, copy
           "gteps_per_sweep": round(g.E / (ksweep / max(nsweep, 1) * 1e-3) / 1e9, 2),
           "kernel_ms_one_batch": {k: round(v[0], 4) for k, v in kt.items() if v[1] > 0},
           "roofline": None}
    g.release_device()
    return res


# --------------------------------------------------------------------------- schedules
def bench_schedule(args, stash):
    from starplat_amd import mixed_batch

    log(f"mixed batch: {args.sched_graphs} graphs ({args.sched_held_out} held out): sweep, train, held-out pass")
    ranges = dict(rmat_scales=tuple(args.sched_rmat), er_log2=tuple(args.sched_er), grid_sides=tuple(args.sched_grid))
    sample = {}
    ntr = args.sched_graphs - args.sched_held_out

    def keep(i, job, outs, picks):
        # every k-th held-out job: its class-0 and learned-class outputs, for
        # the CPU child's oracle check and timing sample
        if (i - ntr) % args.sched_cpu_every == 0:
            kl = picks["learned"]
            sample[i] = {"kind": job.kind, "learned": kl, "out_fixed": outs[0][0], "it_fixed": outs[0][1],
                         "out_learned": outs[kl][0], "it_learned": outs[kl][1]}

    res = mixed_batch.run(n=args.sched_graphs, held_out=args.sched_held_out, seed=5, epochs=400, log=log,
                          on_held_out=keep, batched=args.sched_batched, **ranges)
    res["workload"] = ("mixed batch: RMAT (scale %d-%d, EF 4-32), Erdos-Renyi (2^%d-2^%d vertices, "
                       "avg degree 4-16) and 2-D grids (%d^2-%d^2), each with one SSSP (weights U[1,100], src 0) or "
                       "PageRank job (beta 1e-10, delta 0.85, maxIter 100), seed 5; 8 schedule classes timed per job; "
                       "MLP 16-64-64-8 trained on the GPU; makespan = sum of held-out job times"
                       % (*ranges["rmat_scales"], *ranges["er_log2"], *ranges["grid_sides"]))
    arrays, meta = {}, {"jobs": [], "ranges": {k: list(v) for k, v in ranges.items()}}
    for i, d in sample.items():
        arrays[f"fixed_{i}"] = d["out_fixed"]
        arrays[f"learned_{i}"] = d["out_learned"]
        meta["jobs"].append({"i": int(i), "kind": int(d["kind"]), "learned": int(d["learned"]),
                             "it_fixed": int(d["it_fixed"]), "it_learned": int(d["it_learned"])})
    meta["held_out"] = args.sched_held_out
    stash.put("sched", arrays, **meta)
    return res


# --------------------------------------------------------------------------- CPU leg
def cpu_median(fn, reps, warm=None):
    """(median seconds of `reps` timed calls after one warm-up call, last result)."""
    (warm or fn)()
    times, out = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
    return statistics.median(times), out


def rel_linf(got, ref):
    den = float(np.max(np.abs(ref)))
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref))) / den if den > 0 else 0.0


def cpu_child(args):
    """The CPU-baseline leg (runs in a child process; see main): the oracle
    timed on this host's cores and checked against the GPU outputs in
    args.cpu_child."""
    nproc = len(os.sched_getaffinity(0))  # before OpenMP binds this thread
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O  # CPU baseline leg only

    st = Stash(args.cpu_child)
    threads = O.num_threads()
    env = {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}
    base = {"cores": threads, "kind": "port", "omp": env}
    out = {"host": {"nproc": nproc, "cpu_count": os.cpu_count(), "lscpu": host_cores(), "omp_threads": threads, **env}}
    reps = args.cpu_reps

    def clog(*a):
        print("[bench-cpu]", *a, file=sys.stderr, flush=True)

    for leg in st.legs():
        meta, arr = st.get(leg)
        clog(leg)
        r = {}
        if leg == "sssp":
            g = graph_for(leg, args)
            t, ref = cpu_median(lambda: O.compute_sssp(g, g.edgeLen, 0), reps)
            r["parity_bit_exact"] = bool(np.array_equal(arr["dist"], ref))
            r["cpu_baseline"] = {**base, "value": round(meta["m_reach"] / t / 1e9, 4), "unit": "GTEPS",
                                 "sample": f"full workload, median of {reps} runs after 1 warm-up", "seconds": round(t, 4)}
        elif leg in ("pr", "pr_c0"):
            g = graph_for(leg, args)
            t, (ref, it_ref) = cpu_median(lambda: O.compute_pr(g, 1e-10, 0.85, 100), reps)
            got, it = arr["ranks"], meta["iterations"]
            fp64, _ = O.pr_fp64(g, 1e-10, 0.85, it)
            fp64_ref, _ = O.pr_fp64(g, 1e-10, 0.85, it_ref)
            # The tolerance is against FP64 at the same iteration count; the FP32
            # restatement's own error is reported beside it (a sequential FP32 sum
            # over a hub's in-edges is itself off by about 2e-5 at RMAT-22).
            r.update({"rel_linf_vs_restatement": rel_linf(got, ref), "iterations_oracle": it_ref,
                      "rel_linf_vs_fp64": rel_linf(got, fp64), "restatement_rel_linf_vs_fp64": rel_linf(ref, fp64_ref),
                      "tolerance_vs_fp64": 1e-5})
            r["cpu_baseline"] = {**base, "value": round(g.E * it_ref / t / 1e9, 4), "unit": "GTEPS",
                                 "sample": f"full run to convergence ({it_ref} iterations), median of {reps} runs "
                                           "after 1 warm-up", "seconds": round(t, 4)}
        elif leg == "tc":
            g = graph_for(leg, args)
            t0 = time.perf_counter()
            ref = O.compute_tc(g)
            t = time.perf_counter() - t0
            r["parity_bit_exact"] = bool(ref == meta["triangles"])
            r["triangles_oracle"] = ref
            r["cpu_baseline"] = {**base, "value": round(meta["m_und"] / t / 1e9, 6), "unit": "GTEPS",
                                 "sample": "full workload (every outer vertex, binary-search is_an_edge), one run",
                                 "seconds": round(t, 3)}
        elif leg == "bc":
            g = graph_for(leg, args)
            src = sources(g.V, meta["sources"], seed=4)[:meta["batch"]]
            t, (ref, _, _) = cpu_median(lambda: O.compute_bc(g, src), reps, warm=lambda: O.compute_bc(g, src[:8]))
            t_all = t * meta["sources"] / len(src)
            r["parity_batch"] = {"sources": len(src), "rel_linf_vs_fp64": rel_linf(arr["bc_first_batch"], ref),
                                 "tolerance": 1e-5}
            r["cpu_baseline"] = {**base, "value": round(meta["sources"] * g.E / t_all / 1e9, 5), "unit": "GTEPS",
                                 "sample": f"the first {len(src)}-source batch (median of {reps} runs after a warm-up "
                                           f"on 8 sources), time scaled to {meta['sources']} sources",
                                 "seconds_estimated": round(t_all, 2)}
        elif leg == "dyn":
            g = graph_for(leg, args)
            src_of = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(g.indexofNodes))
            rng = np.random.default_rng(5)
            b = [dyn_batch(g, src_of, args.dyn_batch, rng) for _ in range(2)][1]
            tm = {}
            O.dyn_sssp(g, g.edgeLen, 0, b[0], b[1], b[2], args.dyn_batch, timing=tm)
            r["cpu_baseline"] = {**base, "cores": 1, "value": round(args.dyn_batch / tm["batch_seconds"]),
                                 "unit": "updates/s", "sample": "one batch after the restatement's own static SSSP "
                                 "(sequential DynSSSP restatement, batch loop timed)",
                                 "seconds": round(tm["batch_seconds"], 4)}
        elif leg == "dyn_tc":
            g = graph_for(leg, args)
            src_of = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(g.indexofNodes))
            rng = np.random.default_rng(6)
            touched = set()
            b = [dyn_tc_batch(g, src_of, args.dyn_batch, rng, touched) for _ in range(2)][1]
            tm = {}
            half, k = args.dyn_batch // 2, 1024  # the batch holds its deletions, then its additions
            sl = np.r_[0:k, half:half + k]
            O.dyn_tc(g, bytes(np.frombuffer(b[0], np.uint8)[sl]), b[1][sl], b[2][sl], 2 * k, timing=tm, initial=0)
            r["cpu_baseline"] = {**base, "cores": 1, "value": round(2 * k / tm["batch_seconds"]), "unit": "updates/s",
                                 "sample": f"the first {k} deletions and first {k} additions of one batch as one "
                                 "batch (sequential DynTC restatement, batch loop timed, static count skipped)",
                                 "seconds": round(tm["batch_seconds"], 4)}
        elif leg == "sched":
            from starplat_amd import mixed_batch
            from starplat_amd.schedule import JOB_SSSP
            ranges = {k: tuple(v) for k, v in meta["ranges"].items()}
            total, checked, bad, worst = 0.0, 0, [], 0.0
            for j in meta["jobs"]:
                job = mixed_batch.make_job(5, j["i"], **ranges)
                g = job.g
                t0 = time.perf_counter()
                if job.kind == JOB_SSSP:
                    ref = O.compute_sssp(g, g.edgeLen, 0)
                else:
                    ref, _ = O.compute_pr(g, 1e-10, 0.85, 100)
                total += time.perf_counter() - t0
                for which in ("fixed", "learned"):
                    got = arr[f"{which}_{j['i']}"]
                    checked += 1
                    if job.kind == JOB_SSSP:
                        ok = bool(np.array_equal(got, ref))
                    else:
                        fp64, _ = O.pr_fp64(g, 1e-10, 0.85, j[f"it_{which}"])
                        e = rel_linf(got, fp64)
                        worst = max(worst, e)
                        ok = e <= 1e-5
                    if not ok:
                        bad.append({"job": j["i"], "name": job.name, "policy": which})
            n = len(meta["jobs"])
            r["oracle_check"] = {"jobs": n, "outputs_checked": checked, "mismatches": bad,
                                 "pr_max_rel_linf_vs_fp64": worst, "tolerance_pr": 1e-5,
                                 "rule": "SSSP bit-exact vs the restatement; PageRank vs the FP64 restatement at the "
                                         "run's own iteration count"}
            est = total * meta["held_out"] / max(n, 1)
            r["cpu_baseline"] = {**base, "value": round(est * 1e3, 1), "unit": "ms (held-out makespan)",
                                 "sample": f"every {meta['held_out'] // max(n, 1)}th held-out job ({n} jobs), each "
                                           f"restated once back to back, time scaled to {meta['held_out']} jobs",
                                 "seconds": round(total, 3)}
        out[leg] = r
    with open(os.path.join(args.cpu_child, "cpu.json"), "w") as f:
        json.dump(out, f)


def run_cpu_leg(args, stash_dir):
    """Runs cpu_child in a child process with OpenMP bound to cores (the
    binding has to be in the environment before the OpenMP runtime starts,
    and must not bind this process's own threads). Returns its results."""
    import subprocess

    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-child", stash_dir, "--cpu-reps", str(args.cpu_reps),
           "--sssp-scale", str(args.sssp_scale), "--tc-scale", str(args.tc_scale), "--bc-n", str(args.bc_n),
           "--dyn-batch", str(args.dyn_batch)]
    log("CPU baseline leg (child process, OMP_PROC_BIND=close OMP_PLACES=cores)")
    rc = subprocess.call(cmd, env=env)
    if rc != 0:
        raise RuntimeError(f"CPU baseline child failed with status {rc}")
    with open(os.path.join(stash_dir, "cpu.json")) as f:
        return json.load(f)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--algos", default="sssp,pr,pr_c0,tc,bc,dyn,dyn_tc,dyn_pr,sched")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--sssp-scale", type=int, default=22)
    ap.add_argument("--tc-scale", type=int, default=21)
    ap.add_argument("--bc-n", type=int, default=4194304)
    ap.add_argument("--bc-sources", type=int, default=256)
    ap.add_argument("--dyn-batch", type=int, default=65536)
    ap.add_argument("--dyn-batches", type=int, default=6)
    ap.add_argument("--sched-graphs", type=int, default=2000)
    ap.add_argument("--sched-held-out", type=int, default=400)
    ap.add_argument("--sched-rmat", type=int, nargs=2, default=[12, 20])
    ap.add_argument("--sched-er", type=int, nargs=2, default=[12, 20])
    ap.add_argument("--sched-grid", type=int, nargs=2, default=[64, 1024])
    ap.add_argument("--sched-cpu-every", type=int, default=20,
                    help="the CPU leg restates every k-th held-out job of the mixed batch")
    ap.add_argument("--sched-batched", action="store_true", help="also time the held-out jobs as batched unions")
    ap.add_argument("--no-warm-up", action="store_true",
                    help="skip the tiny-graph library warm-up (profiled runs: every dispatch of a trace then "
                         "belongs to the measured workloads; first_run_ms then includes code-object loading)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a known-size kernel for "
                         "scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if hip.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration_mib > 0:
        nbytes = args.copy_calibration_mib << 20
        a = hip.DeviceArray(nbytes // 4, np.int32)
        b = hip.DeviceArray(nbytes // 4, np.int32)
        hip.lib().hipMemset(a.ptr, 1, nbytes)
        hip.lib().hipMemcpy(b.ptr, a.ptr, nbytes, hip.D2D)
        hip.synchronize()
        a.free()
        b.free()
        log(f"calibration copy of {args.copy_calibration_mib} MiB done")
    import tempfile

    if not args.no_warm_up:
        warm_library()
    copy_bw = copy_bandwidth() if not args.no_warm_up else None
    gathers = gather_ceiling() if not args.no_warm_up else None
    algos = args.algos.split(",")
    per = {}
    head = None
    sched = None
    with tempfile.TemporaryDirectory(prefix="spm_bench_") as tmp:
        stash = Stash(tmp)
        if "sssp" in algos:
            per["sssp"] = bench_sssp(args, stash)
        if "pr" in algos:
            per["pr"] = bench_pr(args, stash, "pr", "PageRank at RMAT-22")
        if "pr_c0" in algos:
            per["pr_c0"] = bench_pr(args, stash, "pr_c0", "PageRank at RMAT-16 (65,536 vertices)")
        if "tc" in algos:
            per["tc"] = bench_tc(args, stash)
        if "bc" in algos:
            per["bc"] = bench_bc(args, stash)
        if "dyn" in algos:
            per["dyn"] = bench_dyn(args, stash)
        if "dyn_tc" in algos:
            per["dyn_tc"] = bench_dyn_tc(args, stash)
        if "dyn_pr" in algos:
            per["dyn_pr"] = bench_dyn_pr(args)
        if "sched" in algos:
            sched = bench_schedule(args, stash)
        cpu = None if args.no_cpu_baseline else run_cpu_leg(args, tmp)
    if cpu is not None:
        for leg, r in cpu.items():
            if leg == "host":
                continue
            dst = sched if leg == "sched" else per.get(leg)
            if dst is not None:
                dst.update(r)
    if "sssp" in per:
        head = per["sssp"]
        per["sssp"] = {k: v for k, v in head.items() if k not in ("roofline", "cpu_baseline")}
    if head is None and per:
        first = next(iter(per.values()))
        head = {"workload": first["workload"], "gteps": first.get("gteps", float("nan")),
                "ms_per_step": first.get("ms_per_run", first.get("ms_per_batch")),
                "roofline": first["roofline"], "cpu_baseline": first.get("cpu_baseline")}
    if head is None:  # schedule selection only
        head = {"workload": sched["workload"] if sched else "none", "gteps": float("nan"), "ms_per_step": None,
                "roofline": None, "cpu_baseline": None}
    out = {
        "metric": METRIC, "value": round(head["gteps"], 4), "unit": "GTEPS", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int32",
        "data": "synthetic (seeded RMAT / Erdos-Renyi generators)",
        "config": {"workload": head["workload"], "V": head.get("V"), "E": head.get("E"), "m_reach": head.get("m_reach"),
                   "parallelism": "single GPU"},
        "roofline": head["roofline"],
        "cpu_baseline": head.get("cpu_baseline"),
        "host": cpu["host"] if cpu else host_cores(),
        "hbm_copy": copy_bw,
        "gather_ceiling": gathers,
        "per_algo": per,
    }
    if sched is not None:
        out["schedule_selection"] = sched
    print(json.dumps(out))


if __name__ == "__main__":
    main()
