"""Dynamic-graph path on the GPU through the C ABI: edge updates against the
oracle's update semantics, static algorithms on an updated graph, and
dynamic SSSP against the oracle's DynSSSP restatement (distances and parents
bit-exact) and against Dijkstra on the final graph."""
from collections import Counter

import numpy as np
import pytest

from starplat_amd import Compute_PR, Compute_TC, DynSSSP, graph, parseUpdateFile
from starplat_amd.updates import write_update_file

from test_oracle_dyn import INF2, _G, batch_ops, check_against_dijkstra, random_updates

pytestmark = pytest.mark.gpu


def rows_equal(off, adj, w, roff, radj, ref_off, ref_adj, ref_w, V):
    np.testing.assert_array_equal(off, ref_off)
    for v in range(V):
        a, b = off[v], off[v + 1]
        assert np.all(np.diff(adj[a:b]) >= 0), v
        assert Counter(zip(adj[a:b].tolist(), w[a:b].tolist())) == \
            Counter(zip(ref_adj[a:b].tolist(), ref_w[a:b].tolist())), v
    # the reverse CSR is the transpose of the forward one, rows ascending
    src = np.repeat(np.arange(V), np.diff(off))
    fwd = np.sort(adj.astype(np.int64) * (1 << 32) + src)
    rdst = np.repeat(np.arange(V), np.diff(roff))
    rev = np.sort(rdst.astype(np.int64) * (1 << 32) + radj)
    np.testing.assert_array_equal(fwd, rev)
    for v in range(V):
        assert np.all(np.diff(radj[roff[v]:roff[v + 1]]) >= 0)


@pytest.mark.parametrize("undirected", [False, True])
def test_updates_match_oracle(oracle, undirected):
    g = graph.rmat(11, 8, seed=4, weights=(1, 50), undirected=undirected)
    dev = g.device()
    ops = []
    for k in range(4):
        types, us, ut = random_updates(g, 500, seed=10 + k, symmetric=undirected)
        t = np.frombuffer(types, np.uint8)
        d, a = t == ord("d"), t == ord("a")
        wa = (np.arange(int(a.sum())) % 7 + 1).astype(np.int32)
        dev.delete_edges(us[d], ut[d])
        dev.add_edges(us[a], ut[a], wa)
        ops += [("d", us[d], ut[d]), ("a", us[a], ut[a], wa)]
        if k == 1:  # duplicates of an existing edge, then delete all copies
            s0, t0 = int(np.repeat(np.arange(g.V), np.diff(g.indexofNodes))[5]), int(g.edgeList[5])
            dev.add_edges([s0, s0], [t0, t0], [9, 8])
            ops.append(("a", np.array([s0, s0], np.int32), np.array([t0, t0], np.int32), np.array([9, 8], np.int32)))
            dev.delete_edges([s0, 3], [t0, 3])
            ops.append(("d", np.array([s0, 3], np.int32), np.array([t0, 3], np.int32)))
    ref = oracle.apply_updates(g.V, g.indexofNodes, g.edgeList, g.edgeLen, ops)
    off, adj, w, roff, radj = dev.export()
    assert dev.num_edges() == ref[0][-1]
    rows_equal(off, adj, w, roff, radj, *ref, g.V)
    # a second round of updates on the folded graph
    types, us, ut = random_updates(g, 300, seed=99)
    t = np.frombuffer(types, np.uint8)
    dev.delete_edges(us[t == ord("d")], ut[t == ord("d")])
    dev.add_edges(us[t == ord("a")], ut[t == ord("a")])
    ops2 = [("d", us[t == ord("d")], ut[t == ord("d")]),
            ("a", us[t == ord("a")], ut[t == ord("a")], np.ones(int((t == ord("a")).sum()), np.int32))]
    ref2 = oracle.apply_updates(g.V, *ref, ops2)
    rows_equal(*dev.export(), *ref2, g.V)


def test_update_errors_and_empty_batches(oracle):
    g = graph.rmat(8, 4, seed=1, weights=(1, 5))
    dev = g.device()
    dev.delete_edges([], [])
    dev.add_edges([], [])
    with pytest.raises(Exception):
        dev.delete_edges([0], [g.V])
    with pytest.raises(Exception):
        dev.add_edges([-1], [0])
    off, adj, w, roff, radj = dev.export()
    np.testing.assert_array_equal(off, g.indexofNodes)
    np.testing.assert_array_equal(adj, g.edgeList)
    # delete every edge, then the graph is empty
    src = np.repeat(np.arange(g.V), np.diff(g.indexofNodes)).astype(np.int32)
    dev.delete_edges(src, g.edgeList)
    assert dev.num_edges() == 0
    dev.add_edges([1], [2], [3])
    off, adj, w, roff, radj = dev.export()
    assert off[-1] == 1 and adj.tolist() == [2] and w.tolist() == [3] and radj.tolist() == [1]


def test_static_algorithms_after_updates(oracle):
    g = graph.rmat(11, 8, seed=6, weights=(1, 40))
    dev = g.device()
    types, us, ut = random_updates(g, 800, seed=5)
    t = np.frombuffer(types, np.uint8)
    dev.delete_edges(us[t == ord("d")], ut[t == ord("d")])
    dev.add_edges(us[t == ord("a")], ut[t == ord("a")], np.full(int((t == ord("a")).sum()), 2, np.int32))
    # SSSP on the updated handle sees the updated graph
    d = dev.sssp(0)
    off, adj, w, roff, radj = dev.export()
    fin = _G(g.V, off, adj, w)
    np.testing.assert_array_equal(d, oracle.dijkstra(fin, w, 0))
    # the host mirror follows the device after updateCSR*
    u = graph.rmat(10, 8, seed=8, undirected=True)
    from starplat_amd.updates import update
    src_of = np.repeat(np.arange(u.V), np.diff(u.indexofNodes))
    batch = [update("d", int(src_of[e]), int(u.edgeList[e])) for e in range(0, u.E, 7)]
    batch += [update("d", int(u.edgeList[e]), int(src_of[e])) for e in range(0, u.E, 7)]
    tc0 = oracle.compute_tc(u)
    u.updateCSRDel(batch, 0, len(batch))
    assert u.is_symmetric()
    assert Compute_TC(u) == oracle.compute_tc(u) < tc0
    pr = np.empty(u.V, np.float32)
    Compute_PR(u, 1e-10, 0.85, 20, pr)
    ref, _ = oracle.compute_pr(u, 1e-10, 0.85, 20)
    assert np.max(np.abs(pr - ref)) <= 1e-5 * np.max(np.abs(ref))



def test_fold_refreshes_degree_and_default_schedule(oracle):
    """The fold rebuilds the CSR with the same sorted-pairs builder as the
    relabeled layout and refreshes the handle's max out-degree: a handle on
    the heuristic schedule gets the heuristic of the folded graph (here the
    relabel decision, V >= 2^20 and max out-degree >= 64x the mean, flips
    once a hub appears), a handle whose schedule was set keeps it; SSSP on
    the folded graph matches Dijkstra either way."""
    g = graph.erdos_renyi(1 << 20, 1 << 21, seed=3, weights=(1, 9))
    hub_dst = np.arange(1, 401, dtype=np.int32) * 997
    for explicit in (False, True):
        dev = g.device()
        assert dev.schedule().relabel == 0
        if explicit:
            dev.set_schedule(relabel=0)
        dev.add_edges(np.zeros(len(hub_dst), np.int32), hub_dst, np.full(len(hub_dst), 5, np.int32))
        dev.compact()
        assert dev.schedule().relabel == (0 if explicit else 1)
        d = dev.sssp(0)
        off, adj, w, roff, radj = dev.export()
        assert off[1] - off[0] == g.indexofNodes[1] - g.indexofNodes[0] + len(hub_dst)
        np.testing.assert_array_equal(d, oracle.dijkstra(_G(g.V, off, adj, w), w, 0))
        g.release_device()

@pytest.mark.parametrize("symmetric", [False, True])
@pytest.mark.parametrize("batch", [1, 64, 5000])
def test_dyn_sssp_matches_oracle(oracle, symmetric, batch):
    g = graph.rmat(12, 8, seed=7, weights=(1, 30), undirected=symmetric)
    types, us, ut = random_updates(g, 1500 if batch > 1 else 200, seed=batch, symmetric=symmetric)
    dist_ref, par_ref = oracle.dyn_sssp(g, g.edgeLen, 0, types, us, ut, batch)
    dev = g.device()
    dist, parent = dev.dyn_sssp(0, types, us, ut, batch)
    np.testing.assert_array_equal(dist, dist_ref)
    np.testing.assert_array_equal(parent, par_ref)
    check_against_dijkstra(oracle, g, types, us, ut, batch, dist, parent, 0)


def test_dyn_sssp_stepwise_and_edge_cases(oracle):
    g = graph.from_edges(np.array([0, 1, 2, 5], np.int32), np.array([1, 2, 3, 6], np.int32),
                         np.array([5, 5, 5, 2], np.int32))
    types, us, ut = b"ddaaa", np.array([1, 4, 5, 0, 1], np.int32), np.array([2, 4, 2, 3, 2], np.int32)
    dref, pref = oracle.dyn_sssp(g, g.edgeLen, 0, types, us, ut, 4)
    dev = g.device()
    dev.dynsssp_init(0)
    dev.dynsssp_batch(types[:4], us[:4], ut[:4])
    d1, p1 = dev.dynsssp_result()
    assert d1[2] == INF2 and p1[2] == -1 and d1[3] == 1
    dev.dynsssp_batch(types[4:], us[4:], ut[4:])
    dev.dynsssp_batch(b"", [], [])
    d, p = dev.dynsssp_result()
    np.testing.assert_array_equal(d, dref)
    np.testing.assert_array_equal(p, pref)
    with pytest.raises(Exception):
        dev.dynsssp_batch(b"x", [0], [1])
    # zero weights are refused (parent pointers must strictly shorten)
    z = graph.from_edges(np.array([0, 1], np.int32), np.array([1, 2], np.int32), np.array([0, 1], np.int32))
    with pytest.raises(Exception):
        z.device().dynsssp_init(0)


def test_dynsssp_entry_point_and_update_file(oracle, tmp_path):
    g = graph.rmat(11, 8, seed=9, weights=(1, 25), undirected=True)
    rng = np.random.default_rng(3)
    src_of = np.repeat(np.arange(g.V), np.diff(g.indexofNodes))
    e = rng.integers(0, g.E, 300)
    types = np.where(rng.random(300) < 0.5, ord("d"), ord("a")).astype(np.uint8)
    s = np.where(types == ord("d"), src_of[e], rng.integers(0, g.V, 300))
    t = np.where(types == ord("d"), g.edgeList[e], rng.integers(0, g.V, 300))
    path = tmp_path / "updates.txt"
    write_update_file(str(path), types.tolist(), s, t)
    updateBatch = parseUpdateFile(str(path))
    assert len(updateBatch) == 600
    from starplat_amd.updates import as_arrays
    ty, us, ut, _ = as_arrays(updateBatch)
    dref, pref = oracle.dyn_sssp(g, g.edgeLen, 3, ty, us, ut, 100)
    w0 = g.edgeLen.copy()
    dist = np.empty(g.V, np.int32)
    parent = np.empty(g.V, np.int32)
    DynSSSP(g, dist, parent, w0, updateBatch, 100, 3)
    np.testing.assert_array_equal(dist, dref)
    np.testing.assert_array_equal(parent, pref)
    # g now holds the updated graph
    g0 = graph.rmat(11, 8, seed=9, weights=(1, 25), undirected=True)
    ref = oracle.apply_updates(g0.V, g0.indexofNodes, g0.edgeList, g0.edgeLen, batch_ops(ty, us, ut, 100))
    np.testing.assert_array_equal(g.indexofNodes, ref[0])
    np.testing.assert_array_equal(g.edgeList, ref[1])
    assert g.is_symmetric()


def test_cpp_header_driver(oracle, tmp_path):
    """graph::updateCSRDel / updateCSRAdd and spm_dyn_sssp from C++ through
    starplat_graph.hpp (tests/drivers/dyn_driver.cpp, built by build())."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "build", "bin", "dyn_driver")
    assert os.access(exe, os.X_OK), "run build() first"
    g0 = graph.rmat(10, 8, seed=12, weights=(1, 20), undirected=True)
    gpath, upath = tmp_path / "g.txt", tmp_path / "u.txt"
    g0.write_text(str(gpath))
    g = graph(str(gpath)).parseGraph()
    rng = np.random.default_rng(4)
    src_of = np.repeat(np.arange(g.V), np.diff(g.indexofNodes))
    e = rng.integers(0, g.E, 200)
    types = np.where(rng.random(200) < 0.5, ord("d"), ord("a")).astype(np.uint8)
    s = np.where(types == ord("d"), src_of[e], rng.integers(0, g.V, 200))
    t = np.where(types == ord("d"), g.edgeList[e], rng.integers(0, g.V, 200))
    write_update_file(str(upath), types.tolist(), s, t)
    from starplat_amd.updates import as_arrays
    ty, us, ut, _ = as_arrays(parseUpdateFile(str(upath)))
    out = tmp_path / "o.bin"
    r = subprocess.run([exe, "updates", str(gpath), str(upath), "50", str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.int32)
    off = raw[:g.V + 1]
    E = int(off[-1])
    ref = oracle.apply_updates(g.V, g.indexofNodes, g.edgeList, g.edgeLen, batch_ops(ty, us, ut, 50))
    np.testing.assert_array_equal(off, ref[0])
    for v in range(g.V):
        a, b = off[v], off[v + 1]
        assert Counter(zip(raw[g.V + 1 + a:g.V + 1 + b].tolist(), raw[g.V + 1 + E + a:g.V + 1 + E + b].tolist())) == \
            Counter(zip(ref[1][a:b].tolist(), ref[2][a:b].tolist()))
    r = subprocess.run([exe, "sssp", str(gpath), str(upath), "50", str(out), "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.int32)
    dref, pref = oracle.dyn_sssp(g, g.edgeLen, 2, ty, us, ut, 50)
    np.testing.assert_array_equal(raw[:g.V], dref)
    np.testing.assert_array_equal(raw[g.V:], pref)


@pytest.mark.slow
def test_full_dyn_sssp_rmat22(oracle):
    """RMAT-22 (the SSSP benchmark graph), two batches of 65,536 updates:
    distances and parents equal the restatement's; distances equal a static
    SSSP on the folded graph."""
    g = graph.rmat(22, 16, seed=2, weights=(1, 100))
    rng = np.random.default_rng(5)
    src_of = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(g.indexofNodes))
    parts = []
    for _ in range(2):
        e = rng.integers(0, g.E, 32768)
        parts.append((b"d" * 32768 + b"a" * 32768,
                      np.concatenate([src_of[e], rng.integers(0, g.V, 32768)]).astype(np.int32),
                      np.concatenate([g.edgeList[e], rng.integers(0, g.V, 32768)]).astype(np.int32)))
    types = b"".join(p[0] for p in parts)
    us = np.concatenate([p[1] for p in parts])
    ut = np.concatenate([p[2] for p in parts])
    dref, pref = oracle.dyn_sssp(g, g.edgeLen, 0, types, us, ut, 65536)
    dev = g.device()
    dist, parent = dev.dyn_sssp(0, types, us, ut, 65536)
    np.testing.assert_array_equal(dist, dref)
    np.testing.assert_array_equal(parent, pref)
    static = dev.sssp(0)
    reach = static != np.iinfo(np.int32).max
    np.testing.assert_array_equal(dist[reach], static[reach])
    assert np.all(dist[~reach] >= INF2)
    print(f"DynSSSP V={g.V}: reachable {int(reach.sum())}, parity with restatement and static recompute")


# ---------------------------------------------------------------- DynTC
@pytest.mark.parametrize("undirected", [True, False])
@pytest.mark.parametrize("batch", [1, 64, 5000])
def test_dyn_tc_matches_oracle(oracle, undirected, batch):
    """Per-batch triangle counts equal the restatement's, for random updates
    that include absent deletions, duplicate additions and self loops."""
    g = graph.rmat(11, 8, seed=5, undirected=undirected)
    types, us, ut = random_updates(g, 3000 if batch > 1 else 300, seed=batch, symmetric=undirected)
    ref = oracle.dyn_tc(g, types, us, ut, batch)
    dev = g.device()
    dev.dyntc_init()
    t = np.frombuffer(types, np.uint8)
    for b, k in enumerate(range(0, len(us), batch)):
        dev.dyntc_batch(bytes(t[k:k + batch]), us[k:k + batch], ut[k:k + batch])
        assert dev.dyntc_count() == ref[b], b


def test_dyn_tc_valid_batches_match_recount(oracle):
    """Valid symmetric batches on a simple graph: after every batch the count
    equals a recount of the graph the updates leave (on the GPU: the static
    count of the folded graph), and the restatement's."""
    from test_oracle_dyn import er_edges, triangles, undirected_csr, valid_batches
    V = 2000
    edges = er_edges(V, 24000, seed=1)
    g0 = undirected_csr(V, edges)
    src = np.repeat(np.arange(V, dtype=np.int32), np.diff(g0.indexofNodes))
    g = graph.from_edges(src, g0.edgeList)
    assert g.V == V
    np.testing.assert_array_equal(g.edgeList, g0.edgeList)
    types, us, ut, states = valid_batches(V, edges, 4, 600, seed=5)
    ref = oracle.dyn_tc(g0, types, us, ut, 1200)
    want = [triangles(V, st) for st in states]
    assert ref.tolist() == want
    dev = g.device()
    dev.dyntc_init()
    t = np.frombuffer(types, np.uint8)
    for b, k in enumerate(range(0, len(us), 1200)):
        dev.dyntc_batch(bytes(t[k:k + 1200]), us[k:k + 1200], ut[k:k + 1200])
        assert dev.dyntc_count() == want[b]
    assert dev.tc() == want[-1]  # the static count of the updated graph


def test_dyn_tc_python_mirror(oracle, tmp_path):
    """DynTC(g, parseUpdateFile(...), batchSize): the count, and g's arrays are
    the updated graph (updates applied as the edge multiset the oracle keeps)."""
    from starplat_amd import DynTC
    g = graph.rmat(10, 8, seed=2, undirected=True)
    off0, adj0 = g.indexofNodes.copy(), g.edgeList.copy()
    types, us, ut = random_updates(g, 400, seed=3)
    f = tmp_path / "u.txt"
    write_update_file(f, types, us, ut)
    ups = parseUpdateFile(str(f))
    t = "".join(u.type for u in ups).encode()
    s = np.array([u.source for u in ups], np.int32)
    d = np.array([u.destination for u in ups], np.int32)
    ref = oracle.dyn_tc(_G(g.V, off0, adj0, np.ones(len(adj0), np.int32)), t, s, d, 50)
    assert DynTC(g, ups, 50) == ref[-1]
    want = oracle.apply_updates(g.V, off0, adj0, np.ones(len(adj0), np.int32), batch_ops(t, s, d, 50))
    np.testing.assert_array_equal(g.indexofNodes, want[0])


@pytest.mark.slow
def test_full_dyn_tc_rmat21():
    """The TC configuration graph (undirected RMAT-21) with two batches of
    65,536 directed updates (valid and symmetric): the incremental count equals
    the static count of the updated graph."""
    g = graph.rmat(21, 16, seed=3, undirected=True)
    off, adj = g.indexofNodes, g.edgeList
    rng = np.random.default_rng(9)
    src_of = np.repeat(np.arange(g.V, dtype=np.int32), np.diff(off))

    def is_edge(a, c):
        row = adj[off[a]:off[a + 1]]
        i = np.searchsorted(row, c)
        return i < len(row) and row[i] == c

    dev = g.device()
    dev.dyntc_init()
    assert dev.dyntc_count() == dev.tc()
    touched = set()  # each pair is deleted or added at most once
    for b in range(2):
        ts, us, ut = [], [], []
        for i in rng.choice(g.E, 16384 * 3, replace=False).tolist():
            a, c = int(src_of[i]), int(adj[i])
            key = (min(a, c), max(a, c))
            if a == c or key in touched:
                continue
            touched.add(key)
            ts.append("dd"), us.extend([a, c]), ut.extend([c, a])
            if len(us) >= 32768:
                break
        while len(us) < 65536:
            a, c = int(rng.integers(0, g.V)), int(rng.integers(0, g.V))
            key = (min(a, c), max(a, c))
            if a == c or key in touched or is_edge(a, c):
                continue
            touched.add(key)
            ts.append("aa"), us.extend([a, c]), ut.extend([c, a])
        dev.dyntc_batch("".join(ts).encode(), np.array(us, np.int32), np.array(ut, np.int32))
    assert dev.dyntc_count() == dev.tc()


# ---------------------------------------------------------------- DynPR
def rel_linf(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b)) / np.max(np.abs(b)))


def pr_updates(g, n, seed, symmetric):
    """Deletions of existing edges (and a few absent pairs), additions out of
    vertices that have out-edges: count_outNbrs (the starting out-degree) of
    every source stays positive, so ranks stay finite."""
    rng = np.random.default_rng(seed)
    src_of = np.repeat(np.arange(g.V), np.diff(g.indexofNodes))
    has_out = np.flatnonzero(np.diff(g.indexofNodes) > 0)
    types, us, ut = [], [], []
    for _ in range(n):
        if rng.random() < 0.5:
            if rng.random() < 0.9:
                e = int(rng.integers(0, g.E))
                a, b = int(src_of[e]), int(g.edgeList[e])
            else:
                a, b = int(rng.choice(has_out)), int(rng.integers(0, g.V))
            k = "d"
        else:
            a, b, k = int(rng.choice(has_out)), int(rng.choice(has_out)), "a"
        types.append(k), us.append(a), ut.append(b)
        if symmetric:
            types.append(k), us.append(b), ut.append(a)
    return "".join(types).encode(), np.array(us, np.int32), np.array(ut, np.int32)


def run_dyn_pr(dev, types, us, ut, batch, beta=1e-10, delta=0.85, maxIter=100):
    dev.dynpr_init(beta, delta, maxIter)
    t = np.frombuffer(types, np.uint8)
    for k in range(0, len(us), batch):
        dev.dynpr_batch(bytes(t[k:k + batch]), us[k:k + batch], ut[k:k + batch])
    return dev.dynpr_result(), dev.dynpr_iterations()


# The emitted DynPR divides by the starting out-degree, so additions make mass
# grow through every phase that then runs to maxIter; long streams overflow
# to infinity (in the restatement as on the GPU). These sizes stay finite.
@pytest.mark.parametrize("undirected,n,batch,dels_only", [(True, 200, 100, False), (False, 1200, 200, False),
                                                          (False, 2000, 500, False), (True, 1000, 100, True)])
def test_dyn_pr_matches_oracle(oracle, undirected, n, batch, dels_only):
    """Ranks within 1e-4 relative L-inf of the restatement (sums in another
    order over up to maxIter sweeps per phase); sweep counts equal wherever
    the restatement's phase stopped after one sweep on the sign of its diff.
    (A phase that converges stops where its FP32 diff first reaches beta,
    which depends on the summation order: 97 against 100 sweeps here.)"""
    g = graph.rmat(10, 8, seed=6, undirected=undirected)
    types, us, ut = pr_updates(g, n, seed=batch, symmetric=undirected)
    if dels_only:
        m = np.frombuffer(types, np.uint8) == ord("d")
        types, us, ut = bytes(np.frombuffer(types, np.uint8)[m]), us[m], ut[m]
    ref, its_ref = oracle.dyn_pr(g, 1e-10, 0.85, 100, types, us, ut, batch)
    got, its = run_dyn_pr(g.device(), types, us, ut, batch)
    assert np.all(np.isfinite(ref))
    assert len(its) == len(its_ref)
    sure = its_ref == 1
    np.testing.assert_array_equal(its[sure], its_ref[sure])
    assert rel_linf(got, ref) <= 1e-4, (its.tolist(), its_ref.tolist())


def test_dyn_pr_division_by_zero_as_emitted(oracle):
    """An added edge out of a vertex whose starting out-degree is 0 divides
    by zero in the emitted code (pr / count_outNbrs): the same infinities
    and NaNs appear on the GPU."""
    g = graph.rmat(9, 8, seed=2)
    types, us, ut = random_updates(g, 300, seed=4)
    ref, its_ref = oracle.dyn_pr(g, 1e-10, 0.85, 100, types, us, ut, 100)
    got, its = run_dyn_pr(g.device(), types, us, ut, 100)
    assert not np.all(np.isfinite(ref))
    np.testing.assert_array_equal(np.isnan(got), np.isnan(ref))
    np.testing.assert_array_equal(np.isinf(got), np.isinf(ref))
    np.testing.assert_array_equal(its, its_ref)


def test_dyn_pr_python_mirror(oracle, tmp_path):
    from starplat_amd import DynPR
    g = graph.rmat(9, 8, seed=4)
    off0, adj0 = g.indexofNodes.copy(), g.edgeList.copy()
    types, us, ut = pr_updates(g, 300, seed=5, symmetric=False)
    f = tmp_path / "u.txt"
    write_update_file(f, types, us, ut)
    ups = parseUpdateFile(str(f))
    t = "".join(u.type for u in ups).encode()
    s = np.array([u.source for u in ups], np.int32)
    d = np.array([u.destination for u in ups], np.int32)
    ref, _ = oracle.dyn_pr(_G(g.V, off0, adj0, np.ones(len(adj0), np.int32)), 1e-10, 0.85, 100, t, s, d, 100)
    pr = np.empty(g.V, np.float32)
    DynPR(g, 1e-10, 0.85, 100, pr, ups, 100)
    assert rel_linf(pr, ref) <= 1e-4
    want = oracle.apply_updates(g.V, off0, adj0, np.ones(len(adj0), np.int32), batch_ops(t, s, d, 100))
    np.testing.assert_array_equal(g.indexofNodes, want[0])
