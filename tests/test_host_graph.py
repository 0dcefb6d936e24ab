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
