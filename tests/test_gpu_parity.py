"""Parity of the gfx950 kernels (called through the C ABI) with the CPU oracle.

Bar: SSSP distances, TC counts, BFS levels and path counts bit-exact; PageRank
within 1e-5 relative L-inf of the FP64 oracle and of the FP32 restatement (up
to the restatement's own FP64 error, reported); BC within 1e-5 relative L-inf
of the FP64 oracle. Relative L-inf here is max|gpu - ref| / max|ref|.
"""
import json
import os

import numpy as np
import pytest

from starplat_amd import Compute_BC, Compute_PR, Compute_SSSP, Compute_TC, graph, sources
from conftest import GOLDEN

pytestmark = pytest.mark.gpu
EXP = json.load(open(os.path.join(GOLDEN, "sample_graph_expected.json")))
INT_MAX = np.iinfo(np.int32).max
PR_TOL = 1e-5
BC_TOL = 1e-5


def rel_linf(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# ------------------------------------------------------------------ golden fixture
def test_golden_sample_graph(sample_graph):
    g = sample_graph
    assert Compute_TC(g) == EXP["triangles"]
    lev, sig = g.device().bc_paths(0)
    np.testing.assert_array_equal(lev, EXP["bfs_levels_from_0"])
    np.testing.assert_array_equal(sig, EXP["sigma_from_0"])
    bc = np.empty(g.V, np.float32)
    Compute_BC(g, bc, {0})
    np.testing.assert_allclose(bc, EXP["bc_sources_0"], rtol=0, atol=1e-6)
    dist = np.empty(g.V, np.int32)
    Compute_SSSP(g, dist, g.getEdgeLen(), 0)
    np.testing.assert_array_equal(dist, EXP["sssp_from_0_zero_weights"])


# ------------------------------------------------------------------ SSSP
SSSP_CASES = [
    ("rmat", dict(scale=8, edge_factor=8, seed=1)),
    ("rmat", dict(scale=12, edge_factor=16, seed=2)),
    ("rmat", dict(scale=16, edge_factor=16, seed=3)),
    ("er", dict(n=20000, m=80000, seed=4)),
    ("grid", dict(rows=60, cols=70)),
]


def make(kind, kw, weights=(1, 100)):
    if kind == "rmat":
        return graph.rmat(kw["scale"], kw["edge_factor"], seed=kw["seed"], weights=weights)
    if kind == "er":
        return graph.erdos_renyi(kw["n"], kw["m"], seed=kw["seed"], undirected=False, weights=weights)
    return graph.grid2d(kw["rows"], kw["cols"], weights=weights)


@pytest.mark.parametrize("kind,kw", SSSP_CASES)
@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_sssp_parity(oracle, kind, kw, mode):
    g = make(kind, kw)
    dev = g.device()
    dev.set_schedule(sssp_mode=mode)
    for src in (0, g.V // 3, g.V - 1):
        ref = oracle.compute_sssp(g, g.edgeLen, src)
        got = dev.sssp(src)
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("rounds,grid", [(6, 1), (12, 0), (48, 7), (96, 4096)])
def test_sssp_schedules_bit_identical(oracle, rounds, grid):
    g = graph.rmat(14, 16, seed=5, weights=(1, 100))
    ref = oracle.compute_sssp(g, g.edgeLen, 0)
    dev = g.device()
    dev.set_schedule(sssp_rounds=rounds, sssp_grid=grid)
    np.testing.assert_array_equal(dev.sssp(0), ref)
    np.testing.assert_array_equal(dev.sssp(0), ref)  # replay of the captured graph


def test_sssp_rounds_fitted_to_previous_run(oracle):
    # The first replay's length follows the previous run on the handle: a
    # deep run after a shallow one (and after one with no out-edges) must
    # continue with further replays; a shallow run after a deep one stops early.
    n = 400
    src = np.concatenate([np.arange(n - 1), np.arange(0, n - 1, 7)]).astype(np.int32)
    dst = np.concatenate([np.arange(1, n), np.arange(0, n - 1, 7) + 1]).astype(np.int32)
    w = (np.arange(src.size) % 5 + 1).astype(np.int32)
    g = graph.from_edges(src, dst, w)
    dev = g.device()
    for s in (0, n - 1, 0, n - 3, n // 2, 0, 0):
        np.testing.assert_array_equal(dev.sssp(s), oracle.compute_sssp(g, g.edgeLen, s))
    # The same history decides which rounds get compaction kernels: after a
    # run from a vertex without out-edges, the big frontiers from vertex 0 land
    # in rounds captured without them and push through the counter.
    g = graph.rmat(17, 16, seed=9, weights=(1, 100))
    deg = np.diff(g.indexofNodes)
    leaf = int(np.flatnonzero(deg == 0)[0])
    dev = g.device()
    ref0 = oracle.compute_sssp(g, g.edgeLen, 0)
    for s in (leaf, 0, leaf, 0, 0):
        np.testing.assert_array_equal(dev.sssp(s), ref0 if s == 0 else oracle.compute_sssp(g, g.edgeLen, s))


@pytest.mark.parametrize("mode", [0, 2, 3])
def test_sssp_negative_weights_and_many_sources(oracle, mode):
    # Negative weights (no negative cycle: every edge goes up in id) leave the
    # edges unpacked and the first frontier unbanded; the fixed point is the
    # same Bellman-Ford distance.
    rng = np.random.default_rng(11)
    n, m = 3000, 24000
    a = rng.integers(0, n, m)
    b = rng.integers(0, n, m)
    keep = a != b
    s, d = np.minimum(a, b)[keep].astype(np.int32), np.maximum(a, b)[keep].astype(np.int32)
    w = rng.integers(-20, 100, s.size).astype(np.int32)
    g = graph.from_edges(s, d, w)
    dev = g.device()
    dev.set_schedule(sssp_mode=mode)
    for src in (0, 17, 1500):
        np.testing.assert_array_equal(dev.sssp(src), oracle.compute_sssp(g, g.edgeLen, src))
    # Non-negative weights too wide to pack with a 16-bit vertex id (21 + 16
    # bits): unpacked edges, banded first frontier with bands of 2^15.
    g = graph.rmat(16, 16, seed=6, weights=(1, 1 << 20))
    dev = g.device()
    dev.set_schedule(sssp_mode=mode)
    for src in (0, 9):
        np.testing.assert_array_equal(dev.sssp(src), oracle.compute_sssp(g, g.edgeLen, src))
    # Many seeds at once: the seeds' round is itself large, and its frontier
    # is banded over the whole graph.
    from starplat_amd import hip
    g = graph.rmat(16, 16, seed=4, weights=(1, 100))
    srcs = np.arange(0, g.V, 7, dtype=np.int32)
    out = hip.DeviceArray(g.V, np.int32)
    dev = g.device()
    dev.set_schedule(sssp_mode=mode)
    dev.sssp_multi_device(srcs, out)
    # nearest-seed distance = distance from a virtual vertex joined to every
    # seed by a zero-weight edge
    V = g.V
    ss = np.concatenate([np.repeat(np.arange(V, dtype=np.int32), np.diff(g.indexofNodes)), np.full(srcs.size, V, np.int32)])
    dd = np.concatenate([g.edgeList, srcs])
    ww = np.concatenate([g.edgeLen, np.zeros(srcs.size, np.int32)])
    gv = graph.from_edges(ss.astype(np.int32), dd.astype(np.int32), ww.astype(np.int32))
    np.testing.assert_array_equal(out.to_host(), oracle.compute_sssp(gv, gv.edgeLen, V)[:V])


class _RawCsr:
    """A CSR handed to spm_graph_create as is (rows non-decreasing, repeated
    neighbours allowed), without the reverse CSR."""

    def __init__(self, off, adj, w):
        self.indexofNodes = np.ascontiguousarray(off, np.int32)
        self.edgeList = np.ascontiguousarray(adj, np.int32)
        self.edgeLen = np.ascontiguousarray(w, np.int32)
        self.rev_indexofNodes = None
        self.srcList = None
        self.V = len(off) - 1
        self.E = len(adj)

    def num_nodes(self):
        return self.V

    def num_edges(self):
        return self.E


@pytest.mark.parametrize("mode", [0, 2])
def test_sssp_source_row_with_repeated_neighbours(oracle, mode):
    """The banded first frontier of a single source tiles over that source's
    row, skipping a slot that repeats the neighbour before it: rows with
    repeated neighbours (different weights), a self-loop, a row longer than
    V (the vertex tiles then serve instead) and a source without out-edges."""
    from starplat_amd.device import DeviceGraph
    rng = np.random.default_rng(21)
    for V, hub_len in ((3000, 2500), (5000, 9000), (40, 100)):
        rows = []
        for v in range(V):
            n = hub_len if v == 7 else int(rng.integers(0, 12))
            if v == V - 1:
                n = 0
            r = np.sort(rng.integers(0, V, n))
            if v == 7:
                r = np.sort(np.concatenate([r, np.repeat(r[:50], 3), [7, 7]]))
            rows.append(r)
        off = np.zeros(V + 1, np.int64)
        off[1:] = np.cumsum([len(r) for r in rows])
        adj = np.concatenate(rows).astype(np.int32)
        w = rng.integers(1, 100, adj.size).astype(np.int32)
        g = _RawCsr(off, adj, w)
        dev = DeviceGraph(g)
        dev.set_schedule(sssp_mode=mode)
        for src in (7, 0, V - 1):
            np.testing.assert_array_equal(dev.sssp(src), oracle.compute_sssp(g, g.edgeLen, src), err_msg=f"V={V} src={src}")
        dev.close()


def test_sssp_band_local_long_paths(oracle):
    """sssp_mode 3 (each workgroup relaxes its own 4096-vertex band for up to
    sssp_band_rounds rounds per launch, 8 by default): paths far longer than
    one launch's rounds, so
    bands hand queued vertices and cross-band improvements to later launches
    -- a 300 x 300 grid from a corner and from the middle, a 20,000-vertex
    path (with a shortcut at its end) and a ring walked backwards -- plus a
    source without out-edges and weights too wide to pack."""
    n = 20000
    s = np.concatenate([np.arange(n - 1), [n - 10]]).astype(np.int32)
    d = np.concatenate([np.arange(1, n), [5]]).astype(np.int32)
    w = np.concatenate([np.arange(n - 1) % 7 + 1, [1]]).astype(np.int32)
    ring = np.arange(9000, dtype=np.int32)
    cases = [(graph.grid2d(300, 300, weights=(1, 100)), (0, 150 * 300 + 150, 300 * 300 - 1)),
             (graph.from_edges(s, d, w), (0, n // 2, n - 1)),
             (graph.from_edges(ring, (ring - 1) % 9000, np.full(9000, 3, np.int32)), (0, 4500)),
             (graph.grid2d(120, 90, weights=(1, 1 << 20)), (0, 5000))]
    for g, srcs in cases:
        dev = g.device()
        dev.set_schedule(sssp_mode=3)
        for src in srcs:
            np.testing.assert_array_equal(dev.sssp(src), oracle.compute_sssp(g, g.edgeLen, src))
    # Several seeds at once, on the grid and on its degree-relabeled copy
    # (ids no longer local: most edges leave their band), each round cap.
    from starplat_amd import hip
    g = graph.grid2d(200, 150, weights=(1, 100))
    seeds = np.array([0, 7777, 29999], np.int32)
    V = g.V
    ss = np.concatenate([np.repeat(np.arange(V, dtype=np.int32), np.diff(g.indexofNodes)), np.full(seeds.size, V, np.int32)])
    gv = graph.from_edges(ss, np.concatenate([g.edgeList, seeds]).astype(np.int32),
                          np.concatenate([g.edgeLen, np.zeros(seeds.size, np.int32)]))
    ref = oracle.compute_sssp(gv, gv.edgeLen, V)[:V]
    dev = g.device()
    out = hip.DeviceArray(V, np.int32)
    for relabel, k in ((0, 1), (0, 8), (0, 4096), (1, 8)):
        dev.set_schedule(sssp_mode=3, relabel=relabel, sssp_band_rounds=k)
        dev.sssp_multi_device(seeds, out)
        np.testing.assert_array_equal(out.to_host(), ref)
        np.testing.assert_array_equal(dev.sssp(7777), oracle.compute_sssp(g, g.edgeLen, 7777))


def test_sssp_edge_cases(oracle):
    # single vertex, no edges
    g = graph.from_edges(np.array([0], np.int32), np.array([0], np.int32), np.array([5], np.int32))
    np.testing.assert_array_equal(g.device().sssp(0), [0])
    # source with no out-edges; unreachable part; zero and heavy weights; duplicates summed
    s = np.array([1, 1, 2, 3, 3, 4], np.int32)
    d = np.array([2, 3, 4, 4, 4, 0], np.int32)
    w = np.array([0, 7, 2147483, 1, 2, 9], np.int32)
    g = graph.from_edges(s, d, w)
    for src in range(g.V):
        np.testing.assert_array_equal(g.device().sssp(src), oracle.compute_sssp(g, g.edgeLen, src))
    assert g.device().sssp(0)[1] == INT_MAX
    # a caller-supplied weight array replaces the handle's weights
    g = graph.rmat(10, 8, seed=7, weights=(1, 100))
    w2 = (np.arange(g.E) % 13 + 1).astype(np.int32)
    dist = np.empty(g.V, np.int32)
    Compute_SSSP(g, dist, w2, 0)
    np.testing.assert_array_equal(dist, oracle.compute_sssp(g, w2, 0))
    with pytest.raises(Exception):
        g.device().sssp(g.V)  # source out of range -> error status, no launch


def test_graph_create_rejects_reverse_that_is_not_the_transpose():
    """Each CSR on its own is well formed, but the reverse one does not hold
    the forward edges: refused (SPM_EINVAL) by the device-side pair-hash check
    before any algorithm runs."""
    import ctypes

    from starplat_amd import _native as N

    lib = N.device_lib()
    off = np.array([0, 1, 2, 3], np.int32)
    adj = np.array([1, 2, 0], np.int32)         # 0->1, 1->2, 2->0
    for roff, radj in (([0, 1, 2, 3], [1, 2, 0]),  # 1->0, 2->1, 0->2: the opposite cycle
                       ([0, 1, 1, 3], [2, 0, 1])):  # wrong in-degrees
        roff = np.array(roff, np.int32)
        radj = np.array(radj, np.int32)
        h = ctypes.c_void_p()
        rc = lib.spm_graph_create(3, 3, N.ptr(off, ctypes.c_int32), N.ptr(adj, ctypes.c_int32), None,
                                  N.ptr(roff, ctypes.c_int32), N.ptr(radj, ctypes.c_int32), ctypes.byref(h))
        assert rc == -1 and not h.value
        assert b"transpose" in lib.spm_last_error()
    # the true transpose is accepted
    roff = np.array([0, 1, 2, 3], np.int32)
    radj = np.array([2, 0, 1], np.int32)
    h = ctypes.c_void_p()
    rc = lib.spm_graph_create(3, 3, N.ptr(off, ctypes.c_int32), N.ptr(adj, ctypes.c_int32), None,
                              N.ptr(roff, ctypes.c_int32), N.ptr(radj, ctypes.c_int32), ctypes.byref(h))
    assert rc == 0 and h.value
    lib.spm_graph_destroy(h)


# ------------------------------------------------------------------ PR
def exits_agree(oracle, g, it, it_ref, beta=1e-10, maxIter=100):
    """Loop-exit parity. With beta = 1e-10 the exact iteration (FP64) would
    run to maxIter; an FP32 run stops earlier only once round-off makes its
    ranks stationary (a fixed point) -- or never, if they settle into a
    one-ulp 2-cycle. Which of the two happens, and when, depends on the order
    of each in-row sum. So the exits must be equal, unless both lie past the
    iteration k0 where the FP64 diff has already fallen to 1e-6, i.e. where
    the FP32 diff is round-off."""
    if it == it_ref:
        return True
    _, k0 = oracle.pr_fp64(g, 1e-6, 0.85, maxIter)
    return min(it, it_ref) >= k0


def check_pr_values(oracle, g, pr, it, ref, it_ref):
    """Ranks within 1e-5 relative L-inf of the FP64 oracle (same iteration
    count), and of the FP32 restatement up to that restatement's own FP64
    error: a sequential FP32 sum over a hub's 10^5 in-edges is itself ~2e-5
    off at RMAT-22, which no faithful summation can reproduce."""
    fp64, _ = oracle.pr_fp64(g, 1e-10, 0.85, it)
    fp64_ref, _ = oracle.pr_fp64(g, 1e-10, 0.85, it_ref)
    e_gpu, e_ref = rel_linf(pr, fp64), rel_linf(ref, fp64_ref)
    print(f"PR V={g.V}: iters gpu {it} restatement {it_ref}; rel Linf vs FP64: gpu {e_gpu:.2e}, "
          f"restatement {e_ref:.2e}; gpu vs restatement {rel_linf(pr, ref):.2e}")
    assert e_gpu <= PR_TOL
    assert rel_linf(pr, ref) <= PR_TOL + e_ref


@pytest.mark.parametrize("scale,seed", [(8, 1), (12, 2), (16, 1)])
def test_pr_parity(oracle, scale, seed):
    g = graph.rmat(scale, 16, seed=seed)
    ref, it_ref = oracle.compute_pr(g, 1e-10, 0.85, 100)
    pr = np.empty(g.V, np.float32)
    it = Compute_PR(g, 1e-10, 0.85, 100, pr)
    assert exits_agree(oracle, g, it, it_ref), (it, it_ref)
    check_pr_values(oracle, g, pr, it, ref, it_ref)


def test_pr_iteration_count_and_closed_form(oracle):
    n = 100
    s = np.arange(n, dtype=np.int32)
    g = graph.from_edges(s, (s + 1) % n)
    pr = np.empty(n, np.float32)
    assert Compute_PR(g, 1e-10, 0.85, 100, pr) == 1
    np.testing.assert_array_equal(pr, np.float32(1) / np.float32(n))
    g = graph.rmat(12, 8, seed=3)
    for maxIter in (1, 2, 7):
        ref, it_ref = oracle.compute_pr(g, 1e-10, 0.85, maxIter)
        out, it = g.device().pr(1e-10, 0.85, maxIter)
        assert it == it_ref == maxIter
        assert rel_linf(out, ref) <= PR_TOL
    # large beta: stops after the first iteration whose diff <= beta
    ref, it_ref = oracle.compute_pr(g, 1e-3, 0.85, 100)
    out, it = g.device().pr(1e-3, 0.85, 100)
    assert it == it_ref and rel_linf(out, ref) <= PR_TOL


@pytest.mark.parametrize("t1,t2,t3,grp,seg,mode", [(0, 0, 0, 2, 0, 0), (64, 64, 64, 32, 0, 0), (2, 16, 256, 4, 10, 0),
                                                   (8, 128, 100000, 16, 12, 0), (4, 48, 2048, 8, 11, 0),
                                                   (4, 48, 2048, 8, 0, 1), (0, 0, 0, 2, 0, 3),
                                                   (64, 64, 64, 32, 0, 3), (4, 128, 1024, 4, 0, 3)])
def test_pr_schedules(oracle, t1, t2, t3, grp, seg, mode):
    """Lane mapping, source segmentation (8 passes of 1024 sources at seg 10)
    and push vs pull change only the summation order."""
    g = graph.rmat(13, 16, seed=4)
    ref, _ = oracle.compute_pr(g, 1e-10, 0.85, 30)
    dev = g.device()
    dev.set_schedule(pr_t1=t1, pr_t2=t2, pr_t3=t3, pr_group=grp, pr_seg_log2=seg, pr_mode=mode)
    out, it = dev.pr(1e-10, 0.85, 30)
    assert it == 30 and rel_linf(out, ref) <= PR_TOL


@pytest.mark.parametrize("kind", ["rmat16", "rmat17", "odd", "star"])
def test_pr_staged_sweep(oracle, kind):
    """pr_mode 3 (persistent 1024-lane workgroups reading the contributions of
    the lowest ids from LDS): more vertices than the 20,416 staged sources
    (RMAT-16; RMAT-17 on its copy relabeled by out-degree), a
    vertex count that is not a multiple of four (only the first multiple of
    four below it is staged) and a hub row summed by a whole workgroup. Ranks
    within tolerance of the oracle at the restatement's iteration count, and
    equal to pr_mode 0's up to summation order."""
    if kind == "rmat16":
        g = graph.rmat(16, 16, seed=1)
    elif kind == "rmat17":
        g = graph.rmat(17, 16, seed=3)
    elif kind == "odd":
        g = graph.erdos_renyi(30001, 240000, seed=6, undirected=False)
    else:
        n = 25003
        g = graph.from_edges(np.concatenate([np.arange(1, n), np.arange(1, n)]).astype(np.int32),
                             np.concatenate([np.zeros(n - 1), np.arange(2, n + 1) % n]).astype(np.int32))
    dev = g.device()
    relabel = int(kind == "rmat17")
    for beta, mi in ((-1.0, 25), (1e-10, 100)):
        ref, it_ref = oracle.compute_pr(g, beta, 0.85, mi)
        dev.set_schedule(pr_mode=0, pr_t3=1024, relabel=relabel)
        base, it0 = dev.pr(beta, 0.85, mi)
        dev.set_schedule(pr_mode=3, relabel=relabel)
        out, it = dev.pr(beta, 0.85, mi)
        if beta < 0:
            assert it == it0 == it_ref == mi
        else:
            assert exits_agree(oracle, g, it, it_ref), (it, it_ref)
        assert rel_linf(out, base) <= PR_TOL
        check_pr_values(oracle, g, out, it, ref, it_ref)


@pytest.mark.parametrize("kind", ["rmat16", "grid"])
def test_pr_staged_sweep_loop_test(oracle, kind):
    """pr_mode 3 evaluates each sweep's loop test at the start of the next
    sweep, and the last sweep of a captured batch of up to 128 in one check
    kernel. Fixed counts run exactly maxIter sweeps (1, 2 and 130 end or cross
    a batch; from the third sweep on the in-degree-0 rows are skipped) and
    match the FP64 oracle at that count. A converging run stops at the first
    sweep whose diff -- recomputed on the host from the ranks of the runs one
    and two sweeps shorter, which the deterministic sweep reproduces -- is at
    most beta, and not one sweep earlier."""
    g = graph.rmat(16, 16, seed=1) if kind == "rmat16" else graph.grid2d(100, 130)
    dev = g.device()
    dev.set_schedule(pr_mode=3)
    for mi in (1, 2, 37, 130):
        out, it = dev.pr(-1.0, 0.85, mi)
        assert it == mi
        fp64, _ = oracle.pr_fp64(g, -1.0, 0.85, mi)
        assert rel_linf(out, fp64) <= PR_TOL, mi
    # delta 0.95 converges slowly enough to stop in the second batch
    for beta, delta in ((1e-3, 0.85), (1e-6, 0.85), (1e-8, 0.85), (1e-6, 0.95)):
        out, k = dev.pr(beta, delta, 400)
        assert 3 <= k < 400, (beta, k)
        again, k2 = dev.pr(beta, delta, 400)
        assert k2 == k
        np.testing.assert_array_equal(again, out)
        r1, _ = dev.pr(beta, delta, k - 1)
        r2, _ = dev.pr(beta, delta, k - 2)
        d_k = float(np.abs(out.astype(np.float64) - r1).sum())
        d_k1 = float(np.abs(r1.astype(np.float64) - r2).sum())
        print(f"staged loop test {kind} beta {beta} delta {delta}: stop at {k}, diff {d_k:.4e}, "
              f"one sweep earlier {d_k1:.4e}")
        assert d_k <= beta * (1 + 1e-3) and d_k1 > beta * (1 - 1e-3), (beta, k, d_k, d_k1)


# ------------------------------------------------------------------ relabeled layout
@pytest.mark.parametrize("kind", ["rmat12", "rmat16", "grid", "star"])
def test_pr_lagged_loop_test_matches_separate_check(oracle, kind):
    """pr_mode 2 (each sweep evaluates the loop test of the sweep before, the
    check kernel runs once per captured batch) sums the same per-vertex terms
    and the same per-workgroup shares in the same order as pr_mode 0, so ranks
    and iteration counts are identical: fixed counts (beta < 0 runs to
    maxIter; 1 and 130 cross the batch length of 128), convergence inside the
    first batch and after it, and the oracle within tolerance."""
    if kind == "rmat12":
        g = graph.rmat(12, 16, seed=2)
    elif kind == "rmat16":
        g = graph.rmat(16, 16, seed=1)
    elif kind == "grid":
        g = graph.grid2d(100, 130)
    else:  # one hub row long enough for the workgroup-per-vertex bin
        n = 6000
        g = graph.from_edges(np.concatenate([np.arange(1, n), np.arange(1, n)]).astype(np.int32),
                             np.concatenate([np.zeros(n - 1), np.arange(2, n + 1) % n]).astype(np.int32))
    dev = g.device()
    cases = [(-1.0, 37), (-1.0, 1), (-1.0, 2), (-1.0, 130), (1e-10, 100), (1e-10, 300), (1e-6, 300)]
    out = {}
    for mode in (0, 2):
        dev.set_schedule(pr_mode=mode)
        out[mode] = [dev.pr(beta, 0.85, mi) for beta, mi in cases]
    for (beta, mi), a, b in zip(cases, out[0], out[2]):
        assert a[1] == b[1], (beta, mi, a[1], b[1])
        if beta < 0:
            assert a[1] == mi
        np.testing.assert_array_equal(a[0], b[0])
    ref, it_ref = oracle.compute_pr(g, 1e-10, 0.85, 100)
    pr, it = out[2][4]
    assert exits_agree(oracle, g, it, it_ref), (it, it_ref)
    check_pr_values(oracle, g, pr, it, ref, it_ref)


def test_relabel_sssp_bit_identical(oracle):
    """The degree-ordered copy changes vertex ids inside the kernels only:
    distances come back bit-identical in original ids, for both frontier
    modes, several sources and a multi-source seed."""
    g = graph.rmat(14, 16, seed=5, weights=(1, 100))
    dev = g.device()
    for mode in (0, 1):
        dev.set_schedule(relabel=0, sssp_mode=mode)
        base = [dev.sssp(s) for s in (0, 7, g.V - 1)]
        dev.set_schedule(relabel=1, sssp_mode=mode)
        for s, b in zip((0, 7, g.V - 1), base):
            np.testing.assert_array_equal(dev.sssp(s), b)
    np.testing.assert_array_equal(base[1], oracle.compute_sssp(g, g.edgeLen, 7))
    from starplat_amd import hip
    d = hip.DeviceArray(g.V, np.int32)
    srcs = np.array([3, 99, 1000], np.int32)
    dev.set_schedule(relabel=1, sssp_mode=0)
    dev.sssp_multi_device(srcs, d)
    multi = d.to_host()
    d.free()
    np.testing.assert_array_equal(multi, np.min([oracle.compute_sssp(g, g.edgeLen, int(s)) for s in srcs], axis=0))
    # new weights reach the relabeled copy
    w2 = (g.edgeLen * 3 + 1).astype(np.int32)
    dist = np.empty(g.V, np.int32)
    Compute_SSSP(g, dist, w2, 0)
    np.testing.assert_array_equal(dist, oracle.compute_sssp(g, w2, 0))


def test_sssp_packed_and_unpacked_edges(oracle):
    """Edges are read as one (neighbour << wb | weight) word when both fit 32
    bits; new weights that do not fit switch back to the two arrays (and
    weights that fit again repack), with unchanged results."""
    g = graph.rmat(12, 16, seed=9, weights=(1, 100))
    dev = g.device()
    for w in (g.edgeLen,                                           # 7 weight bits: packed
              (g.edgeLen.astype(np.int64) * 1000003 % (1 << 28) + 1).astype(np.int32),  # 29 bits: unpacked
              (g.edgeLen * 2).astype(np.int32)):                   # packed again
        for relabel in (0, 1):
            dev.set_schedule(relabel=relabel)
            dist = np.empty(g.V, np.int32)
            Compute_SSSP(g, dist, w, 3)
            np.testing.assert_array_equal(dist, oracle.compute_sssp(g, w, 3))


def test_relabel_pr_and_updates(oracle):
    g = graph.rmat(14, 16, seed=6)
    dev = g.device()
    dev.set_schedule(relabel=1)
    ref, it_ref = oracle.compute_pr(g, 1e-10, 0.85, 100)
    out, it = dev.pr(1e-10, 0.85, 100)
    assert exits_agree(oracle, g, it, it_ref), (it, it_ref)
    check_pr_values(oracle, g, out, it, ref, it_ref)
    # push mode on the relabeled copy
    dev.set_schedule(relabel=1, pr_mode=1)
    out2, it2 = dev.pr(1e-10, 0.85, 30)
    ref2, _ = oracle.compute_pr(g, 1e-10, 0.85, 30)
    assert it2 == 30 and rel_linf(out2, ref2) <= PR_TOL
    # edge updates drop the copy; the next run sees the updated graph
    dev.set_schedule(relabel=1, pr_mode=0)
    src_of = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(g.indexofNodes))
    dev.delete_edges(src_of[::5], g.edgeList[::5])
    off, adj, _, roff, radj = dev.export()

    class _U:
        V, indexofNodes, edgeList, rev_indexofNodes, srcList = g.V, off, adj, roff, radj
    ref3, it3_ref = oracle.compute_pr(_U, 1e-10, 0.85, 20)
    out3, it3 = dev.pr(1e-10, 0.85, 20)
    assert it3 == 20 and rel_linf(out3, ref3) <= PR_TOL


def test_relabel_default_heuristic():
    small = graph.rmat(12, 16, seed=1)
    assert small.device().default_schedule().relabel == 0
    er = graph.erdos_renyi(1 << 20, 1 << 22, seed=1)
    assert er.device().default_schedule().relabel == 0  # no degree skew


# ------------------------------------------------------------------ TC
@pytest.mark.parametrize("scale,seed", [(8, 1), (12, 2), (16, 3)])
@pytest.mark.parametrize("mode", [0, 1])
def test_tc_parity(oracle, scale, seed, mode):
    g = graph.rmat(scale, 16, seed=seed, undirected=True)
    g.device().set_schedule(tc_mode=mode)
    assert Compute_TC(g) == oracle.compute_tc(g)


def test_tc_long_oriented_rows_fall_back():
    """K_1100: oriented rows reach 1099 entries, past the LDS hash (1024), so
    the count falls back to the edge-centric search; C(1100, 3) triangles."""
    n = 1100
    a, b = np.meshgrid(np.arange(n), np.arange(n))
    m = a != b
    g = graph.from_edges(a[m].astype(np.int32), b[m].astype(np.int32))
    assert Compute_TC(g) == n * (n - 1) * (n - 2) // 6


def test_nearly_empty_graph_all_algorithms(oracle):
    """Ragged input: 1,000 vertices, two edges between the last three, all
    other vertices isolated (no out- or in-edges)."""
    g = graph.from_edges(np.array([997, 998], np.int32), np.array([998, 999], np.int32), np.array([3, 4], np.int32))
    assert g.V == 1000 and g.E == 2
    dev = g.device()
    for s in (0, 500, 997, 999):
        np.testing.assert_array_equal(dev.sssp(s), oracle.compute_sssp(g, g.edgeLen, s))
    assert dev.sssp(997)[999] == 7 and dev.sssp(0)[1] == INT_MAX
    pr = np.empty(g.V, np.float32)
    it = Compute_PR(g, 1e-10, 0.85, 100, pr)
    ref, it_ref = oracle.compute_pr(g)
    assert rel_linf(pr, ref) <= PR_TOL
    assert Compute_TC(g) == 0
    ref_bc, _, _ = oracle.compute_bc(g, [0, 997, 998])
    bc = dev.bc(np.array([0, 997, 998], np.int32))
    np.testing.assert_allclose(bc, ref_bc, rtol=0, atol=1e-6)
    assert bc[998] == 1.0  # on the path 997 -> 998 -> 999
    # the band-local SSSP mode and the staged PageRank sweep on the same
    # ragged input, and on a single vertex
    dev.set_schedule(sssp_mode=3, pr_mode=3)
    for s in (0, 997):
        np.testing.assert_array_equal(dev.sssp(s), oracle.compute_sssp(g, g.edgeLen, s))
    out, it3 = dev.pr(1e-10, 0.85, 100)
    assert rel_linf(out, ref) <= PR_TOL
    one = graph.from_edges(np.array([0], np.int32), np.array([0], np.int32), np.array([5], np.int32))
    d1 = one.device()
    d1.set_schedule(sssp_mode=3, pr_mode=3)
    np.testing.assert_array_equal(d1.sssp(0), [0])
    out1, _ = d1.pr(1e-10, 0.85, 100)
    assert rel_linf(out1, oracle.compute_pr(one)[0]) <= PR_TOL


def test_tc_closed_forms_and_directed(oracle):
    n = 30
    a, b = np.meshgrid(np.arange(n), np.arange(n))
    m = a != b
    assert Compute_TC(graph.from_edges(a[m].astype(np.int32), b[m].astype(np.int32))) == n * (n - 1) * (n - 2) // 6
    assert Compute_TC(graph.grid2d(40, 50)) == 0
    # directed input: the DSL's own sum, not the triangle count
    g = graph.rmat(11, 8, seed=6)
    assert not g.is_symmetric()
    assert Compute_TC(g) == oracle.compute_tc(g, literal=True)
    g = graph.erdos_renyi(3000, 30000, seed=2, undirected=True)
    assert Compute_TC(g) == oracle.compute_tc(g)


# ------------------------------------------------------------------ BC
@pytest.mark.parametrize("kind", ["rmat_dir", "rmat_und", "er", "grid", "path"])
def test_bc_paths_bit_exact(oracle, kind):
    """Levels and path counts of single sources; the grid (up to 178 levels)
    and the path (2,999) run the forward pass over many windows of levels."""
    if kind == "rmat_dir":
        g = graph.rmat(12, 8, seed=1)
    elif kind == "rmat_und":
        g = graph.rmat(12, 8, seed=2, undirected=True)
    elif kind == "grid":
        g = graph.grid2d(80, 100)
    elif kind == "path":
        n = 3000
        g = graph.from_edges(np.arange(n - 1, dtype=np.int32), np.arange(1, n, dtype=np.int32))
    else:
        g = graph.erdos_renyi(10000, 40000, seed=3)
    for s in (0, 1, g.V // 2):
        _, lev_ref, sig_ref = oracle.compute_bc(g, [s])
        lev, sig = g.device().bc_paths(s)
        np.testing.assert_array_equal(lev, lev_ref)
        np.testing.assert_array_equal(sig, sig_ref)


@pytest.mark.parametrize("nsrc", [1, 5, 64, 65, 150])
def test_bc_parity(oracle, nsrc):
    g = graph.erdos_renyi(8000, 32000, seed=4)
    src = sources(g.V, nsrc, seed=4)
    ref, _, _ = oracle.compute_bc(g, src)
    bc = np.empty(g.V, np.float32)
    Compute_BC(g, bc, set(src.tolist()))
    assert rel_linf(bc, ref) <= BC_TOL, rel_linf(bc, ref)


@pytest.mark.parametrize("nsrc", [3, 64, 100])
def test_bc_parity_high_diameter(oracle, nsrc):
    """A 90 x 110 grid (up to 198 levels per batch: the level windows grow
    4, 8, 16, 32 and the levels after the last return at once)."""
    g = graph.grid2d(90, 110)
    src = sources(g.V, nsrc, seed=9)
    ref, _, _ = oracle.compute_bc(g, src)
    bc = np.empty(g.V, np.float32)
    Compute_BC(g, bc, set(src.tolist()))
    assert rel_linf(bc, ref) <= BC_TOL, rel_linf(bc, ref)


def test_bc_hub_frontier(oracle):
    """A hub in a small frontier: the top-down step appends more vertices in
    one workgroup round than it stages in LDS (4096), so the overflow path and
    a level mixing both runs; two hubs share leaves so targets are reached
    from several frontier vertices at once."""
    rng = np.random.default_rng(11)
    n = 20000
    leaves = np.arange(2, n, dtype=np.int32)
    src = [np.zeros(n - 2, np.int32), np.ones(n // 2, np.int32), rng.integers(2, n, 60000).astype(np.int32)]
    dst = [leaves, leaves[: n // 2], rng.integers(2, n, 60000).astype(np.int32)]
    g = graph.from_edges(np.concatenate(src + dst), np.concatenate(dst + src))  # symmetric
    for s in (0, 1, 5):
        _, lev_ref, sig_ref = oracle.compute_bc(g, [s])
        lev, sig = g.device().bc_paths(s)
        np.testing.assert_array_equal(lev, lev_ref)
        np.testing.assert_array_equal(sig, sig_ref)
    srcs = np.array([0, 1] + list(range(100, 162)) + [7], np.int32)
    ref, _, _ = oracle.compute_bc(g, srcs)
    assert rel_linf(g.device().bc(srcs), ref) <= BC_TOL


def test_bc_path_counts_beyond_fp32(oracle):
    """Path counts beyond the FP32 mantissa stay exact (FP64 sigma): a chain
    of 2x2 complete layers doubles them per layer, up to 2^38; a second
    component keeps its counts small, and batches mix both."""
    L = 40
    s, d = [], []
    for i in range(L - 1):
        for a in (2 * i, 2 * i + 1):
            for b in (2 * i + 2, 2 * i + 3):
                s.append(a)
                d.append(b)
    base = 2 * L  # a small undirected ER component after the chain
    rng = np.random.default_rng(3)
    es = rng.integers(0, 500, 3000) + base
    ed = rng.integers(0, 500, 3000) + base
    k = es != ed
    s = np.concatenate([np.array(s), es[k], ed[k]]).astype(np.int32)
    d = np.concatenate([np.array(d), ed[k], es[k]]).astype(np.int32)
    g = graph.from_edges(s, d)
    for src in (0, 1, base + 3):
        _, lev_ref, sig_ref = oracle.compute_bc(g, [src])
        lev, sig = g.device().bc_paths(src)
        np.testing.assert_array_equal(lev, lev_ref)
        np.testing.assert_array_equal(sig, sig_ref)
    assert sig_ref.max() < 2**24 and oracle.compute_bc(g, [0])[2].max() == 2.0 ** (L - 2)
    for srcs in ([0, 1, base + 7], list(range(base, base + 70)), [0] + list(range(base, base + 100))):
        srcs = np.array(srcs, np.int32)
        ref, _, _ = oracle.compute_bc(g, srcs)
        assert rel_linf(g.device().bc(srcs), ref) <= BC_TOL


def test_bc_directed_rmat(oracle):
    g = graph.rmat(12, 16, seed=8)
    src = sources(g.V, 70, seed=1)
    ref, _, _ = oracle.compute_bc(g, src)
    bc = g.device().bc(src)
    assert rel_linf(bc, ref) <= BC_TOL
    with pytest.raises(Exception):
        g.device().bc(np.array([1, 1], np.int32))  # duplicates: not a set


# ------------------------------------------------------------------ full size
@pytest.mark.slow
def test_full_sssp_rmat22(oracle):
    g = graph.rmat(22, 16, seed=2, weights=(1, 100))
    got = g.device().sssp(0)
    ref = oracle.compute_sssp(g, g.edgeLen, 0)
    np.testing.assert_array_equal(got, ref)
    g.device().set_schedule(sssp_mode=1)
    np.testing.assert_array_equal(g.device().sssp(0), ref)


@pytest.mark.slow
def test_full_tc_rmat21(oracle):
    g = graph.rmat(21, 16, seed=3, undirected=True)
    assert Compute_TC(g) == oracle.compute_tc(g)


@pytest.mark.slow
def test_full_pr_rmat22(oracle):
    g = graph.rmat(22, 16, seed=1)
    ref, it_ref = oracle.compute_pr(g, 1e-10, 0.85, 100)
    out, it = g.device().pr(1e-10, 0.85, 100)
    assert exits_agree(oracle, g, it, it_ref)
    check_pr_values(oracle, g, out, it, ref, it_ref)


@pytest.mark.slow
def test_full_bc_er4m_256(oracle):
    """The BC configuration as stated: ER, 4,194,304 vertices, 256 sources in
    four 64-source batches whose dependencies all add into one FP32 BC[v],
    against the FP64 restatement over the whole source set (bc_dslV2:5-20, the
    sourceSet loop); levels and path counts bit-exact for the first source of
    every batch."""
    g = graph.erdos_renyi(4194304, 16777216, seed=4)
    src = sources(g.V, 256, seed=4)
    for s in src[::64]:
        _, lev_ref, sig_ref = oracle.compute_bc(g, [int(s)])
        lev, sig = g.device().bc_paths(int(s))
        np.testing.assert_array_equal(lev, lev_ref)
        np.testing.assert_array_equal(sig, sig_ref)
    ref, _, _ = oracle.compute_bc(g, src)
    got = g.device().bc(src)
    e = rel_linf(got, ref)
    print(f"BC ER-4M 256 sources: rel Linf vs FP64 {e:.2e}")
    assert e <= BC_TOL


# ------------------------------------------------------------------ batched SSSP
def test_sssp_multi_source_and_batch(oracle):
    from starplat_amd import hip
    from starplat_amd.batch import SsspBatch

    gs = [graph.rmat(s, 8, seed=s, weights=(1, 100)) for s in (6, 9, 11)]
    gs += [graph.grid2d(30, 40, weights=(1, 100)), graph.erdos_renyi(3000, 9000, seed=2, undirected=False, weights=(1, 100))]
    srcs = [0, 7, 100, 0, 2999]
    b = SsspBatch(list(zip(gs, srcs)))
    b.run()
    for g, s, got in zip(gs, srcs, b.results()):
        np.testing.assert_array_equal(got, oracle.compute_sssp(g, g.edgeLen, s))
    b.close()
    # several sources in one graph: distance to the nearest source
    g = graph.rmat(12, 8, seed=3, weights=(1, 100))
    src = np.array([0, 5, 5, 4000], np.int32)  # duplicate ignored
    out = hip.DeviceArray(g.V, np.int32)
    g.device().sssp_multi_device(src, out)
    ref = np.min(np.stack([oracle.compute_sssp(g, g.edgeLen, int(s)) for s in set(src.tolist())]), axis=0)
    np.testing.assert_array_equal(out.to_host(), ref)


def test_pr_batch_matches_single_jobs(oracle):
    """Uneven PageRank jobs in one launch sequence: every job's ranks are those
    of its own run (bit-identical: vertices are binned by in-degree with the
    same lane mapping, none of these is large enough to be relabeled) and
    within the tolerance of the restatement; jobs stop independently."""
    from starplat_amd.batch import PrBatch

    gs = [graph.rmat(s, 8, seed=s) for s in (6, 9, 12)]
    gs += [graph.grid2d(30, 40), graph.erdos_renyi(3000, 9000, seed=2, undirected=False), graph.rmat(10, 16, seed=4)]
    b = PrBatch(gs)
    for beta, maxIter in ((1e-10, 100), (1e-4, 37), (1e-10, 7)):
        iters = b.run(beta, 0.85, maxIter)
        for g, got, it in zip(gs, b.results(), iters):
            single, it_single = g.device().pr(beta, 0.85, maxIter)
            assert it == it_single
            np.testing.assert_array_equal(got, single)
            if beta == 1e-10:  # away from the stopping threshold, so both runs stop together
                ref, _ = oracle.compute_pr(g, beta, 0.85, int(it))
                assert rel_linf(got, ref) <= PR_TOL
        if beta == 1e-4:
            assert len(set(iters.tolist())) > 1  # the jobs stopped at different iterations
        if maxIter == 7:
            assert set(iters.tolist()) == {7}
    b.close()


def test_pr_batch_rejects_bad_jobs():
    from starplat_amd import hip

    g = graph.rmat(8, 8, seed=1)
    dev = g.device()
    out = hip.DeviceArray(g.V, np.float32)
    with pytest.raises(Exception):  # an edge joins the two halves
        dev.pr_multi_device([0, g.V // 2, g.V], 1e-10, 0.85, 10, out)
    with pytest.raises(Exception):  # empty job
        dev.pr_multi_device([0, 0, g.V], 1e-10, 0.85, 10, out)
    with pytest.raises(Exception):  # does not cover [0, V)
        dev.pr_multi_device([0, g.V - 1], 1e-10, 0.85, 10, out)
    # one job covering the graph is the ordinary run
    it = dev.pr_multi_device([0, g.V], 1e-10, 0.85, 10, out)
    single, it_single = dev.pr(1e-10, 0.85, 10)
    assert it[0] == it_single
    np.testing.assert_array_equal(out.to_host(), single)
    out.free()
