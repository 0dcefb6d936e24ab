"""Pins the oracle's restatements of the dynamic-graph path: the edge-update
semantics (updateCSRDel / updateCSRAdd) against a dictionary model, and the
DynSSSP restatement against textbook Dijkstra on the graph the updates leave,
with planted mistakes that these checks catch."""
from collections import Counter

import numpy as np
import pytest

from starplat_amd import graph
from starplat_amd.updates import as_arrays, getAdditions, getDeletions, parseUpdateFile, update

INF2 = (2**31 - 1) // 2
INT_MAX = 2**31 - 1


def random_updates(g, n, seed, p_del=0.5, symmetric=False):
    """n updates: deletions of existing edges (and a few absent ones),
    additions of random pairs; with symmetric, each followed by its reverse
    (as parseUpdateFile pushes them)."""
    rng = np.random.default_rng(seed)
    src_of = np.repeat(np.arange(g.V), np.diff(g.indexofNodes))
    types, us, ut = [], [], []
    for _ in range(n):
        if rng.random() < p_del:
            if g.E and rng.random() < 0.9:
                e = int(rng.integers(0, g.E))
                s, t = int(src_of[e]), int(g.edgeList[e])
            else:
                s, t = int(rng.integers(0, g.V)), int(rng.integers(0, g.V))
            k = "d"
        else:
            s, t, k = int(rng.integers(0, g.V)), int(rng.integers(0, g.V)), "a"
        types.append(k), us.append(s), ut.append(t)
        if symmetric:
            types.append(k), us.append(t), ut.append(s)
    return "".join(types).encode(), np.array(us, np.int32), np.array(ut, np.int32)


def batch_ops(types, us, ut, batch):
    """The update operations a Dynamic function applies: per batch, all its
    deletions, then all its additions (weight 1)."""
    ops = []
    t = np.frombuffer(types, np.uint8)
    for k in range(0, len(us), batch):
        sl = slice(k, k + batch)
        dm = t[sl] == ord("d")
        ops.append(("d", us[sl][dm], ut[sl][dm]))
        am = ~dm
        ops.append(("a", us[sl][am], ut[sl][am], np.ones(int(am.sum()), np.int32)))
    return ops


class _G:
    def __init__(self, V, off, adj, w):
        self.V, self.indexofNodes, self.edgeList, self.edgeLen = V, off, adj, w


def dict_model(g, ops):
    rows = [Counter() for _ in range(g.V)]
    for v in range(g.V):
        for e in range(g.indexofNodes[v], g.indexofNodes[v + 1]):
            rows[v][(int(g.edgeList[e]), int(g.edgeLen[e]))] += 1
    for op in ops:
        if op[0] == "d":
            for s, t in zip(op[1], op[2]):
                for key in [k for k in rows[s] if k[0] == t]:
                    del rows[s][key]
        else:
            for s, t, w in zip(op[1], op[2], op[3]):
                rows[s][(int(t), int(w))] += 1
    return rows


# ------------------------------------------------------------ update semantics
def test_apply_updates_matches_dictionary_model(oracle):
    g = graph.rmat(8, 6, seed=5, weights=(1, 9))
    types, us, ut = random_updates(g, 300, seed=1)
    ops = batch_ops(types, us, ut, 37)
    # duplicate additions and an add-then-delete of the same pair
    ops.append(("a", np.array([1, 1, 2], np.int32), np.array([2, 2, 3], np.int32), np.array([4, 5, 6], np.int32)))
    ops.append(("d", np.array([2], np.int32), np.array([3], np.int32)))
    off, adj, w = oracle.apply_updates(g.V, g.indexofNodes, g.edgeList, g.edgeLen, ops)
    rows = dict_model(g, ops)
    for v in range(g.V):
        got = Counter(zip(adj[off[v]:off[v + 1]].tolist(), w[off[v]:off[v + 1]].tolist()))
        assert got == rows[v], v
        assert np.all(np.diff(adj[off[v]:off[v + 1]]) >= 0)
    assert rows[1][(2, 4)] == 1 and rows[1][(2, 5)] == 1 and not any(k[0] == 3 for k in rows[2])


def test_apply_updates_invariants_and_plant(oracle):
    g = graph.rmat(7, 8, seed=2, weights=(1, 5))
    src_of = np.repeat(np.arange(g.V), np.diff(g.indexofNodes)).astype(np.int32)
    # deleting every edge leaves no edge; adding them back restores the graph
    off, adj, w = oracle.apply_updates(g.V, g.indexofNodes, g.edgeList, g.edgeLen, [("d", src_of, g.edgeList)])
    assert off[-1] == 0
    off2, adj2, w2 = oracle.apply_updates(g.V, off, adj, w, [("a", src_of, g.edgeList, g.edgeLen)])
    np.testing.assert_array_equal(off2, g.indexofNodes)
    np.testing.assert_array_equal(adj2, g.edgeList)
    np.testing.assert_array_equal(w2, g.edgeLen)
    # planted: a deletion that removes only one copy leaves a duplicate behind
    ops = [("a", np.array([0], np.int32), np.array([1], np.int32), np.array([3], np.int32))] * 2
    ops = ops + [("d", np.array([0], np.int32), np.array([1], np.int32))]
    good = oracle.apply_updates(g.V, g.indexofNodes, g.edgeList, g.edgeLen, ops)
    bad = oracle.apply_updates(g.V, g.indexofNodes, g.edgeList, g.edgeLen, ops, plant=1)
    assert good[0][-1] != bad[0][-1]


def test_update_file_parsing(tmp_path):
    p = tmp_path / "u.txt"
    p.write_text("a 1 2\nd 3 4\n\nx 5 6\na 7\n")
    u = parseUpdateFile(str(p))
    assert [(x.type, x.source, x.destination, x.weight) for x in u] == [
        ("a", 1, 2, 1), ("a", 2, 1, 1), ("d", 3, 4, 1), ("d", 4, 3, 1)]
    assert [x.source for x in getDeletions(0, 3, u)] == [3]
    assert [x.source for x in getAdditions(1, 3, u)] == [2]
    t, s, d, w = as_arrays([update("a", 1, 2, 7), update("d", 3, 4)])
    assert t == b"ad" and s.tolist() == [1, 3] and d.tolist() == [2, 4] and w.tolist() == [7, 1]


# ------------------------------------------------------------ DynSSSP
def check_against_dijkstra(oracle, g, types, us, ut, batch, dist, parent, src):
    off, adj, w = oracle.apply_updates(g.V, g.indexofNodes, g.edgeList, g.edgeLen, batch_ops(types, us, ut, batch))
    fin = _G(g.V, off, adj, w)
    ref = oracle.dijkstra(fin, w, src)
    reach = ref != INT_MAX
    np.testing.assert_array_equal(dist[reach], ref[reach])
    assert np.all(dist[~reach] >= INF2)
    assert parent[src] == -1 and np.all(parent[~reach] == -1)
    for v in np.nonzero(reach)[0]:
        if v == src:
            continue
        p = parent[v]
        row = slice(off[p], off[p + 1])
        assert np.any((adj[row] == v) & (ref[p] + w[row] == ref[v])), (v, p)
    # parent chains reach the source (no cycles)
    for v in np.nonzero(reach)[0][:200]:
        x, steps = v, 0
        while x != src:
            x = parent[x]
            steps += 1
            assert steps <= g.V
    return fin


@pytest.mark.parametrize("symmetric", [False, True])
@pytest.mark.parametrize("batch", [1, 16, 1000])
def test_dyn_sssp_restatement_is_exact(oracle, symmetric, batch):
    g = graph.rmat(9, 8, seed=7, weights=(1, 30), undirected=symmetric)
    types, us, ut = random_updates(g, 300, seed=batch, symmetric=symmetric)
    dist, parent = oracle.dyn_sssp(g, g.edgeLen, 0, types, us, ut, batch)
    check_against_dijkstra(oracle, g, types, us, ut, batch, dist, parent, 0)


def test_dyn_sssp_edge_cases(oracle):
    # path 0->1->2->3, weights 5; delete the middle edge, add a bypass from an
    # unreachable vertex, re-add the middle edge in a later batch
    g = graph.from_edges(np.array([0, 1, 2, 5], np.int32), np.array([1, 2, 3, 6], np.int32),
                         np.array([5, 5, 5, 2], np.int32))
    types = b"ddaa" + b"a"
    us = np.array([1, 4, 5, 0, 1], np.int32)  # 4->4 is absent
    ut = np.array([2, 4, 2, 3, 2], np.int32)
    dist, parent = oracle.dyn_sssp(g, g.edgeLen, 0, types, us, ut, 4)
    # after batch 1: 2 is cut off (INT_MAX/2), 3 reached through 0->3 (weight 1)
    assert dist[1] == 5 and dist[3] == 1 and parent[3] == 0
    # batch 2 re-adds 1->2 with weight 1: dist[2] = 6
    assert dist[2] == 6 and parent[2] == 1
    assert dist[5] == INT_MAX and dist[6] == INT_MAX and parent[5] == -1
    check_against_dijkstra(oracle, g, types, us, ut, 4, dist, parent, 0)


def test_dyn_sssp_plants_are_caught(oracle):
    g = graph.rmat(9, 8, seed=11, weights=(1, 30))
    types, us, ut = random_updates(g, 400, seed=3)
    for plant in (1, 2):
        dist, parent = oracle.dyn_sssp(g, g.edgeLen, 0, types, us, ut, 50, plant=plant)
        with pytest.raises(AssertionError):
            check_against_dijkstra(oracle, g, types, us, ut, 50, dist, parent, 0)


# ---------------------------------------------------------------- DynTC
def undirected_csr(V, edges):
    """Symmetric CSR (sorted rows) of a simple undirected edge set."""
    pairs = sorted({(a, b) for a, b in edges} | {(b, a) for a, b in edges})
    off = np.zeros(V + 1, np.int32)
    for a, _ in pairs:
        off[a + 1] += 1
    off = np.cumsum(off).astype(np.int32)
    adj = np.array([b for _, b in pairs], np.int32)
    return _G(V, off, adj, np.ones(len(adj), np.int32))


def triangles(V, edges):
    A = np.zeros((V, V), np.int64)
    for a, b in edges:
        A[a, b] = A[b, a] = 1
    return int(np.trace(A @ A @ A)) // 6


def valid_batches(V, edges, nbatch, per_batch, seed, whole_triangle=True):
    """Batches of undirected updates, each pushed as two directed updates (as
    parseUpdateFile does): deletions of current edges, additions of absent
    pairs, no pair twice in a batch. The first batch deletes all three edges
    of one triangle and two edges of another (when per_batch allows), so
    every count class occurs. Every batch holds exactly per_batch undirected
    updates. Returns (types, us, ut, edge set after each batch)."""
    rng = np.random.default_rng(seed)
    cur = set(edges)
    types, us, ut, states = [], [], [], []

    def push(k, a, b):
        types.extend([k, k]), us.extend([a, b]), ut.extend([b, a])

    for bi in range(nbatch):
        dels, adds, used = [], [], set()
        if bi == 0 and whole_triangle and per_batch >= 5:
            adjs = {v: set() for v in range(V)}
            for a, b in cur:
                adjs[a].add(b), adjs[b].add(a)
            tris = [(a, b, c) for a, b in sorted(cur) for c in sorted(adjs[a] & adjs[b]) if c > max(a, b)]
            (a, b, c), (x, y, z) = tris[0], tris[-1]
            for e in [(a, b), (a, c), (b, c), (x, y), (x, z)]:
                e = (min(e), max(e))
                if e not in used:
                    dels.append(e), used.add(e)
        pool = sorted(cur)
        while len(dels) + len(adds) < per_batch:
            if rng.random() < 0.5 and pool:
                e = pool[int(rng.integers(0, len(pool)))]
                if e in used:
                    continue
                dels.append(e)
            else:
                a, b = int(rng.integers(0, V)), int(rng.integers(0, V))
                e = (min(a, b), max(a, b))
                if a == b or e in cur or e in used:
                    continue
                adds.append(e)
            used.add(e)
        for a, b in dels:
            push("d", a, b)
        for a, b in adds:
            push("a", a, b)
        cur = (cur - set(dels)) | set(adds)
        states.append(set(cur))
    return "".join(types).encode(), np.array(us, np.int32), np.array(ut, np.int32), states


def er_edges(V, m, seed):
    rng = np.random.default_rng(seed)
    out = set()
    while len(out) < m:
        a, b = int(rng.integers(0, V)), int(rng.integers(0, V))
        if a != b:
            out.add((min(a, b), max(a, b)))
    return out


@pytest.mark.parametrize("per_batch", [1, 4, 32, 400])
def test_dyn_tc_restatement_is_exact(oracle, per_batch):
    """Symmetric, valid batches on a simple graph: after every batch the
    incremental count equals a recount (trace(A^3)/6) of the graph the updates
    leave -- each removed or added triangle is seen from 2, 4 or 6 directed
    updates according to how many of its edges the batch holds."""
    V = 120
    edges = er_edges(V, 900, seed=per_batch)
    g = undirected_csr(V, edges)
    assert oracle.compute_tc(g) == triangles(V, edges)
    nb = max(2, 1200 // max(per_batch, 1) // 10)
    types, us, ut, states = valid_batches(V, edges, nb, per_batch, seed=7)
    counts = oracle.dyn_tc(g, types, us, ut, 2 * per_batch)
    assert [int(c) for c in counts] == [triangles(V, st) for st in states]


def test_dyn_tc_hand_case(oracle):
    """K4 has four triangles. Deleting {0,1} removes 012 and 013; adding it
    back restores them. Deleting the three edges of 012 in one batch removes
    every triangle (each of the others uses one of those edges)."""
    V = 4
    edges = {(a, b) for a in range(4) for b in range(a + 1, 4)}
    g = undirected_csr(V, edges)
    assert oracle.compute_tc(g) == 4
    t = b"dd" + b"aa"
    us = np.array([0, 1, 0, 1], np.int32)
    ut = np.array([1, 0, 1, 0], np.int32)
    # batch 1 deletes {0,1} (triangles 012, 013 go), batch 2 adds it back
    assert oracle.dyn_tc(g, t, us, ut, 2).tolist() == [2, 4]
    # one batch deletes {0,1}, {0,2}, {1,2}
    t = b"dddddd"
    us = np.array([0, 1, 0, 2, 1, 2], np.int32)
    ut = np.array([1, 0, 2, 0, 2, 1], np.int32)
    assert oracle.dyn_tc(g, t, us, ut, 6).tolist() == [triangles(4, edges - {(0, 1), (0, 2), (1, 2)})]
    assert triangles(4, edges - {(0, 1), (0, 2), (1, 2)}) == 0


def test_dyn_tc_plants_are_caught(oracle):
    V = 60
    edges = er_edges(V, 500, seed=3)
    g = undirected_csr(V, edges)
    types, us, ut, states = valid_batches(V, edges, 3, 40, seed=11)
    want = [triangles(V, st) for st in states]
    assert oracle.dyn_tc(g, types, us, ut, 80).tolist() == want
    for plant in (1, 2):
        assert oracle.dyn_tc(g, types, us, ut, 80, plant=plant).tolist() != want


# ---------------------------------------------------------------- DynPR
def _sweep(V, src, dst, deg0, pr, flag, delta):
    """One Jacobi sweep of the DSL's update over the flagged vertices, in
    float64: val = (1 - delta)/N + delta * sum_{u -> v} pr[u] / deg0[u]."""
    contrib = np.zeros(V)
    np.add.at(contrib, dst, pr[src] / deg0[src])
    val = (1 - delta) / V + delta * contrib
    out = pr.copy()
    out[flag] = val[flag]
    return out, float(np.sum(val[flag] - pr[flag]))


def test_dyn_pr_static_closed_forms(oracle):
    # with dangling vertices the signed diff of the first sweep is negative
    # (mass leaves through them), so staticPR stops after one sweep from 1/N
    g = graph.rmat(9, 8, seed=3)
    deg0 = np.diff(g.indexofNodes).astype(np.float64)
    assert (deg0 == 0).any()
    empty = np.zeros(0, np.int32)
    pr, its = oracle.dyn_pr(g, 1e-10, 0.85, 100, b"", empty, empty, 10)
    src = np.repeat(np.arange(g.V), np.diff(g.indexofNodes))
    want, d = _sweep(g.V, src, g.edgeList, np.where(deg0 > 0, deg0, 1), np.full(g.V, 1 / g.V), np.ones(g.V, bool), 0.85)
    assert its.tolist() == [1] and d < 0
    np.testing.assert_allclose(pr, want, rtol=1e-5)
    # a directed cycle is a fixed point: every rank stays 1/N
    n = 50
    cyc = graph.from_edges(np.arange(n, dtype=np.int32), ((np.arange(n) + 1) % n).astype(np.int32))
    pr, its = oracle.dyn_pr(cyc, 1e-10, 0.85, 100, b"", empty, empty, 10)
    np.testing.assert_allclose(pr, 1 / n, rtol=1e-6)


def _dyn_pr_numpy(O, g, types, us, ut, batch, delta, beta, maxIter, its):
    """The per-batch steps recomputed from the updated CSR (apply_updates), in
    float64, running as many sweeps per phase as `its` says and checking that
    the signed-diff rule stops each phase exactly there. O: the oracle module
    (its apply_updates keeps the edge multiset)."""
    V = g.V
    deg0 = np.diff(g.indexofNodes).astype(np.float64)
    off, adj = g.indexofNodes, g.edgeList
    w = np.ones(len(adj), np.int32)
    k_it = [0]

    def phase(pr, off, adj, flag, deg):
        src = np.repeat(np.arange(V), np.diff(off))
        c = int(its[k_it[0]])
        k_it[0] += 1
        # float64 cannot reproduce where a float32 run reaches its fixed point
        # (diff 0), so the rule is checked up to a margin: a phase never stops
        # while its diff is clearly positive, never continues once it is clearly
        # negative or zero
        for i in range(c):
            pr, d = _sweep(V, src, adj, deg, pr, flag, delta)
            if i < c - 1:
                assert d > beta - 1e-7
        assert d <= beta + 1e-7 or c == maxIter
        return pr

    pr = phase(np.full(V, 1 / V), off, adj, np.ones(V, bool), np.where(deg0 > 0, deg0, 1))
    t = np.frombuffer(types, np.uint8)
    for k in range(0, len(us), batch):
        sl = slice(k, k + batch)
        dm, am = t[sl] == ord("d"), t[sl] == ord("a")
        flag = np.zeros(V, bool)
        flag[ut[sl][dm]] = True
        flag[adj] = True  # every destination of the graph before the deletions
        off, adj, w = O.apply_updates(V, off, adj, w, [("d", us[sl][dm], ut[sl][dm])])
        pr = phase(pr, off, adj, flag, deg0)
        flag = np.zeros(V, bool)
        flag[ut[sl][am]] = True
        flag[adj] = True  # after the deletions, before the additions
        off, adj, w = O.apply_updates(V, off, adj, w, [("a", us[sl][am], ut[sl][am], np.ones(int(am.sum()), np.int32))])
        pr = phase(pr, off, adj, flag, deg0)
    return pr


def test_dyn_pr_batches_against_recomputation(oracle):
    """Every vertex the graph keeps some out-edge of keeps a positive base
    degree, so the recomputation never divides by zero; updates delete
    existing edges and add edges between vertices that had out-edges."""
    g = graph.rmat(9, 8, seed=5)
    rng = np.random.default_rng(2)
    src_of = np.repeat(np.arange(g.V), np.diff(g.indexofNodes))
    has_out = np.flatnonzero(np.diff(g.indexofNodes) > 0)
    types, us, ut = [], [], []
    for _ in range(300):
        if rng.random() < 0.5:
            e = int(rng.integers(0, g.E))
            types.append("d"), us.append(int(src_of[e])), ut.append(int(g.edgeList[e]))
        else:
            types.append("a"), us.append(int(rng.choice(has_out))), ut.append(int(rng.integers(0, g.V)))
    types = "".join(types).encode()
    us, ut = np.array(us, np.int32), np.array(ut, np.int32)
    pr, its = oracle.dyn_pr(g, 1e-10, 0.85, 100, types, us, ut, 60)
    # phases stop after one sweep when mass leaves (signed diff < 0) and run
    # to maxIter when it grows (base degrees undercount added out-edges)
    assert its[0] == 1 and len(set(its.tolist())) > 1
    want = _dyn_pr_numpy(oracle, g, types, us, ut, 60, 0.85, 1e-10, 100, its)
    np.testing.assert_allclose(pr, want, rtol=1e-4, atol=1e-9)
    for plant in (1, 2):
        bad, _ = oracle.dyn_pr(g, 1e-10, 0.85, 100, types, us, ut, 60, plant=plant)
        assert not np.allclose(bad, want, rtol=1e-4, atol=1e-9)
