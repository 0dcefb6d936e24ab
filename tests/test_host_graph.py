"""Host data layer: graph.hpp CSR semantics, generators, text round trip."""
import json
import os

import numpy as np
import pytest

from starplat_amd import graph, sources
from conftest import GOLDEN


def brute_csr(src, dst, w):
    """graph.hpp parseEdges/parseEdgesContent restated with dicts (small inputs)."""
    m = {}
    for s, d, x in zip(src, dst, w):
        m[(s, d)] = m.get((s, d), 0) + x  # mpp[p] += weightVal
    V = max([0] + list(src) + list(dst)) + 1
    rows = [[] for _ in range(V)]
    for (s, d), x in sorted(m.items()):
        rows[s].append((d, x))
    off = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    adj = np.array([d for r in rows for d, _ in r], np.int32)
    wts = np.array([x for r in rows for _, x in r], np.int32)
    inc = [[] for _ in range(V)]
    for s in range(V):
        for d, _ in rows[s]:
            inc[d].append(s)
    roff = np.cumsum([0] + [len(r) for r in inc]).astype(np.int32)
    radj = np.array([s for r in inc for s in sorted(r)], np.int32)
    return V, off, adj, wts, roff, radj


def test_sample_graph_parse(sample_graph):
    exp = json.load(open(os.path.join(GOLDEN, "sample_graph_expected.json")))
    g = sample_graph
    assert g.num_nodes() == exp["num_nodes"]
    assert g.num_edges() == exp["num_edges"]
    # no weight column: weights parse as 0
    assert np.all(g.getEdgeLen() == 0)
    for v in range(g.V):
        row = g.neighbors(v)
        assert np.all(np.diff(row) > 0)
    assert g.is_symmetric()


def test_text_semantics_duplicates_selfloops(tmp_path):
    p = tmp_path / "g.txt"
    p.write_text("# comment\n3 1 5\n0 2 1\n3 1 7\n% skipped\n2 2 4\n0 1\n\n5 0 2\n")
    g = graph(str(p)).parseGraph()
    V, off, adj, w, roff, radj = brute_csr([3, 0, 3, 2, 0, 5], [1, 2, 1, 2, 1, 0], [5, 1, 7, 4, 0, 2])
    assert g.V == V == 6
    np.testing.assert_array_equal(g.indexofNodes, off)
    np.testing.assert_array_equal(g.edgeList, adj)
    np.testing.assert_array_equal(g.edgeLen, w)
    np.testing.assert_array_equal(g.rev_indexofNodes, roff)
    np.testing.assert_array_equal(g.srcList, radj)
    assert g.getEdge(3, 1) == (12, int(g.indexofNodes[3]))  # duplicate weights summed
    assert g.check_if_nbr(2, 2) and not g.check_if_nbr(1, 3)


def test_from_edges_matches_brute():
    rng = np.random.default_rng(7)
    s = rng.integers(0, 40, 500).astype(np.int32)
    d = rng.integers(0, 45, 500).astype(np.int32)
    w = rng.integers(-5, 50, 500).astype(np.int32)
    g = graph.from_edges(s, d, w)
    V, off, adj, wt, roff, radj = brute_csr(s.tolist(), d.tolist(), w.tolist())
    assert g.V == V
    np.testing.assert_array_equal(g.indexofNodes, off)
    np.testing.assert_array_equal(g.edgeList, adj)
    np.testing.assert_array_equal(g.edgeLen, wt)
    np.testing.assert_array_equal(g.rev_indexofNodes, roff)
    np.testing.assert_array_equal(g.srcList, radj)


def test_text_round_trip(tmp_path):
    g = graph.rmat(10, 8, seed=11, weights=(1, 100))
    p = tmp_path / "rmat.txt"
    g.write_text(p, weights=True)
    h = graph(str(p)).parseGraph()
    for a in ("indexofNodes", "edgeList", "edgeLen", "rev_indexofNodes", "srcList"):
        np.testing.assert_array_equal(getattr(g, a), getattr(h, a))


def check_csr(g):
    off, adj = g.indexofNodes, g.edgeList
    assert off[0] == 0 and off[-1] == g.E and np.all(np.diff(off) >= 0)
    src = np.repeat(np.arange(g.V), np.diff(off))
    assert not np.any(src == adj), "self-loop survived"
    key = src.astype(np.int64) * g.V + adj
    assert np.all(np.diff(key) > 0), "rows not strictly ascending (duplicates?)"
    # reverse CSR is the transpose with ascending rows
    rsrc = np.repeat(np.arange(g.V), np.diff(g.rev_indexofNodes))
    rkey = rsrc.astype(np.int64) * g.V + g.srcList
    assert np.all(np.diff(rkey) > 0)
    np.testing.assert_array_equal(np.sort(adj.astype(np.int64) * g.V + src), rkey)


@pytest.mark.parametrize("undirected", [False, True])
def test_rmat_properties(undirected):
    g = graph.rmat(12, 16, seed=3, undirected=undirected, weights=(1, 100))
    check_csr(g)
    assert g.V <= 1 << 12
    deg = np.diff(g.indexofNodes)
    assert deg[0] == deg.max(), "vertex 0 is the RMAT hub (no relabelling)"
    assert g.edgeLen.min() >= 1 and g.edgeLen.max() <= 100
    if undirected:
        assert g.is_symmetric()
    # determinism: same seed, same graph; other seed, other graph
    h = graph.rmat(12, 16, seed=3, undirected=undirected, weights=(1, 100))
    np.testing.assert_array_equal(g.edgeList, h.edgeList)
    np.testing.assert_array_equal(g.edgeLen, h.edgeLen)
    k = graph.rmat(12, 16, seed=4, undirected=undirected)
    assert k.E != g.E or not np.array_equal(k.edgeList, g.edgeList)


def test_rmat_quadrant_frequencies():
    # With b = c and scale 1 the two endpoint bits follow (a,b,c,d) directly.
    g = graph.rmat(8, 64, a=0.57, b=0.19, c=0.19, seed=5)
    # Top-left quadrant (both ids < 128) must hold the most edges.
    src = np.repeat(np.arange(g.V), np.diff(g.indexofNodes))
    q = [np.sum((src < 128) & (g.edgeList < 128)), np.sum((src < 128) & (g.edgeList >= 128)),
         np.sum((src >= 128) & (g.edgeList < 128)), np.sum((src >= 128) & (g.edgeList >= 128))]
    assert q[0] > q[1] > q[3] and q[0] > q[2] > q[3]


def test_er_and_grid():
    g = graph.erdos_renyi(5000, 20000, seed=4, undirected=True)
    check_csr(g)
    assert g.is_symmetric()
    assert abs(g.E - 40000) < 400  # few duplicates / loops at this density
    gr = graph.grid2d(7, 9)
    check_csr(gr)
    assert gr.V == 63 and gr.E == 2 * (7 * 8 + 6 * 9)


def test_sources_distinct_sorted():
    s = sources(1000, 256, seed=4)
    assert len(set(s.tolist())) == 256 and np.all(np.diff(s) > 0)
    np.testing.assert_array_equal(s, sources(1000, 256, seed=4))
    with pytest.raises(ValueError):
        sources(10, 11, seed=1)


def test_union_is_block_diagonal():
    gs = [graph.rmat(8, 8, seed=i, weights=(1, 100)) for i in range(3)] + [graph.grid2d(5, 7, weights=(1, 9))]
    u, bases = graph.union(gs)
    check_csr(u)
    assert list(np.diff(bases)) == [g.V for g in gs] and u.E == sum(g.E for g in gs)
    for i, g in enumerate(gs):
        lo, hi = bases[i], bases[i + 1]
        rows = u.indexofNodes[lo:hi + 1] - u.indexofNodes[lo]
        np.testing.assert_array_equal(rows, g.indexofNodes)
        seg = slice(u.indexofNodes[lo], u.indexofNodes[hi])
        np.testing.assert_array_equal(u.edgeList[seg] - lo, g.edgeList)
        np.testing.assert_array_equal(u.edgeLen[seg], g.edgeLen)


def test_binary_csr_round_trip_and_corruption(tmp_path):
    g = graph.rmat(12, 8, seed=21, weights=(1, 100))
    p = tmp_path / "g.spmcsr"
    g.save(p)
    h = graph.load(p)
    for a in ("indexofNodes", "edgeList", "edgeLen", "rev_indexofNodes", "srcList"):
        np.testing.assert_array_equal(getattr(g, a), getattr(h, a))
    data = bytearray(p.read_bytes())
    (tmp_path / "short").write_bytes(bytes(data[:-4]))
    data[100] ^= 0xFF
    (tmp_path / "flipped").write_bytes(bytes(data))
    for bad in ("short", "flipped", "missing"):
        with pytest.raises(OSError):
            graph.load(tmp_path / bad)


def test_edge_map_and_degrees():
    """edgeMap (graph.hpp:57, built at :835-846): for each reverse-CSR slot
    the CSR position of the same edge -- checked against the reference's
    rule restated by brute force (match the source in the row's edge list),
    and as a permutation that maps every slot to its own edge."""
    rng = np.random.default_rng(3)
    s = rng.integers(0, 40, 300).astype(np.int32)
    d = rng.integers(0, 40, 300).astype(np.int32)
    g = graph.from_edges(s, d)
    pos = {}
    for u in range(g.V):
        for j in range(g.indexofNodes[u], g.indexofNodes[u + 1]):
            pos[(u, int(g.edgeList[j]))] = j
    m = g.edgeMap
    for v in range(g.V):
        for p in range(g.rev_indexofNodes[v], g.rev_indexofNodes[v + 1]):
            assert m[p] == pos[(int(g.srcList[p]), v)]
    assert np.array_equal(np.sort(m), np.arange(g.E))
    for v in range(g.V):
        assert g.getOutDegree(v) == g.count_outNbrs(v) == len(g.neighbors(v))
        assert g.getInDegree(v) == len(g.nodes_to(v))


# ---- generators restated in numpy --------------------------------------------
# The generators draw from a counter-based stream: value(seed, stream, counter)
# = mix64(mix64(seed ^ mix64(stream * K)) + counter) with the splitmix64
# finaliser. Restated here with wrapping uint64 arithmetic, the samples are
# then reduced to a CSR by numpy's unique/lexsort, an algorithm independent
# of the library's bucketing.
M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _rng(seed, stream, counter):
    with np.errstate(over="ignore"):
        base = _mix64(np.uint64(seed) ^ _mix64(np.asarray(stream, np.uint64) * np.uint64(0xD1B54A32D192ED03)))
        return _mix64(base + np.asarray(counter, np.uint64))


def _csr_from_samples(s, d, seed, undirected, wts):
    keep = s != d
    s, d = s[keep].astype(np.int64), d[keep].astype(np.int64)
    if undirected:
        s, d = np.concatenate([s, d]), np.concatenate([d, s])
    V = int(max(s.max(initial=0), d.max(initial=0))) + 1
    key = np.unique(s * V + d)
    src, dst = key // V, key % V
    off = np.searchsorted(src, np.arange(V + 1)).astype(np.int32)
    w = np.zeros(len(key), np.int32)
    if wts:
        lo, hi = wts
        a, b = (np.minimum(src, dst), np.maximum(src, dst)) if undirected else (src, dst)
        h = _rng(seed, 0x57, (a.astype(np.uint64) << np.uint64(32)) | b.astype(np.uint64))
        w = (lo + (h % np.uint64(hi - lo + 1))).astype(np.int32)
    order = np.lexsort((src, dst))
    roff = np.searchsorted(dst[order], np.arange(V + 1)).astype(np.int32)
    return V, off, dst.astype(np.int32), w, roff, src[order].astype(np.int32)


def _check_against(g, ref):
    V, off, adj, w, roff, radj = ref
    assert g.V == V
    np.testing.assert_array_equal(g.indexofNodes, off)
    np.testing.assert_array_equal(g.edgeList, adj)
    np.testing.assert_array_equal(g.edgeLen, w)
    np.testing.assert_array_equal(g.rev_indexofNodes, roff)
    np.testing.assert_array_equal(g.srcList, radj)


@pytest.mark.parametrize("scale,ef,undirected,wts", [(9, 16, False, (1, 100)), (10, 8, True, None), (7, 30, True, (1, 9))])
def test_rmat_matches_numpy_restatement(scale, ef, undirected, wts):
    a, b, c, seed = 0.57, 0.19, 0.19, 11
    n = ef << scale
    ta, tb, tc = (np.uint64(int(x * 4294967296.0)) for x in (a, a + b, a + b + c))
    i = np.arange(n, dtype=np.uint64)
    s = np.zeros(n, np.uint64)
    d = np.zeros(n, np.uint64)
    for lev in range(0, scale, 2):
        h = _rng(seed, i, lev)
        for k in range(2):
            if lev + k >= scale:
                break
            r = (h >> np.uint64(32)) if k else (h & np.uint64(0xFFFFFFFF))
            bs = (r >= tb).astype(np.uint64)
            bd = (((r >= ta) & (r < tb)) | (r >= tc)).astype(np.uint64)
            s = (s << np.uint64(1)) | bs
            d = (d << np.uint64(1)) | bd
    g = graph.rmat(scale, ef, a, b, c, seed=seed, undirected=undirected, weights=wts)
    _check_against(g, _csr_from_samples(s.astype(np.int64), d.astype(np.int64), seed, undirected, wts))


@pytest.mark.parametrize("undirected", [False, True])
def test_er_matches_numpy_restatement(undirected):
    nv, m, seed = 3000, 12000, 6
    i = np.arange(m, dtype=np.uint64)
    mul = lambda h: ((h >> np.uint64(32)) * np.uint64(nv) + (((h & np.uint64(0xFFFFFFFF)) * np.uint64(nv)) >> np.uint64(32))) >> np.uint64(32)  # noqa: E731
    s, d = mul(_rng(seed, i, 1)), mul(_rng(seed, i, 2))
    g = graph.erdos_renyi(nv, m, seed=seed, undirected=undirected, weights=(1, 100))
    _check_against(g, _csr_from_samples(s.astype(np.int64), d.astype(np.int64), seed, undirected, (1, 100)))


def test_grid_matches_numpy_restatement():
    rows, cols = 13, 17
    v = np.arange(rows * cols).reshape(rows, cols)
    s = np.concatenate([v[:, :-1].ravel(), v[:-1, :].ravel()])
    d = np.concatenate([v[:, 1:].ravel(), v[1:, :].ravel()])
    g = graph.grid2d(rows, cols, seed=2, weights=(1, 50))
    _check_against(g, _csr_from_samples(s, d, 2, True, (1, 50)))


@pytest.mark.parametrize("undirected,nsrc", [(False, 5), (True, 64)])
def test_bc_gather_profile_matches_brute_force(undirected, nsrc):
    """spmh_bc_gather_profile against plain loops: BFS levels per source, then
    for every in-edge (v, w) and every level d at which w is listed, the set
    of sources with w at d and v at d - 1."""
    import ctypes as C
    from collections import deque
    from starplat_amd import _native as N
    g = graph.rmat(8, 6, seed=3, undirected=undirected)
    src = sources(g.V, nsrc, seed=2)
    off, adj = np.asarray(g.indexofNodes), np.asarray(g.edgeList)
    roff, radj = np.asarray(g.rev_indexofNodes), np.asarray(g.srcList)
    lev = np.full((nsrc, g.V), -1)
    for k, s in enumerate(src):
        lev[k, s] = 0
        q = deque([s])
        while q:
            v = q.popleft()
            for w in adj[off[v]:off[v + 1]]:
                if lev[k, w] < 0:
                    lev[k, w] = lev[k, v] + 1
                    q.append(w)
    dag = gathers = l64 = l32 = listed = rows = 0
    for w in range(g.V):
        levels = sorted({int(x) for x in lev[:, w] if x >= 1})
        listed += len(levels)
        rows += len(levels) * int(roff[w + 1] - roff[w])
        for v in radj[roff[w]:roff[w + 1]]:
            for d in levels:
                S = [k for k in range(nsrc) if lev[k, w] == d and lev[k, v] == d - 1]
                if S:
                    gathers += 1
                    dag += len(S)
                    l64 += len({k // 16 for k in S})
                    l32 += len({k // 32 for k in S})
    out = np.zeros(7)
    assert N.host_lib().spmh_bc_gather_profile(g._native(), N.ptr(src, C.c_int32), nsrc, N.ptr(out, C.c_double)) == 0
    assert list(out[:6]) == [dag, gathers, l64, l32, listed, rows]
    assert out[6] == lev.max() + 1
    # DAG edges summed over sources, the per-source count the byte model uses
    per_source = sum(int(((lev[k, adj] == np.repeat(lev[k], np.diff(off)) + 1) & (np.repeat(lev[k], np.diff(off)) >= 0)).sum())
                     for k in range(nsrc))
    assert dag == per_source
